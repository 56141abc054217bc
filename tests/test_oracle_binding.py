"""The oracle's obstacle skinning, nearest-face query and binding (oracle/binding_oracle.c,
oracle/bvh_oracle.c), pinned without the code under test:

  * LBS against rigid closed forms and two-joint blends written with numpy matrices, and
    against the capsule-chain body's own forward kinematics (scenes.ChainBody.at)
  * the triangle closest point against dense sampling of the triangle
  * the tree query against a brute-force nearest face over every face
  * Generate / Solve against closed forms
"""
import numpy as np
import pytest

from oracle import binding as ob
from oracle import oracle as orc
from xrtailor_amd import scenes


def _rot(axis, ang):
    c, s = np.cos(ang), np.sin(ang)
    if axis == "z":
        return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def _T(R, t):
    return np.concatenate([R, np.asarray(t, float)[:, None]], axis=1).reshape(-1)


def test_lbs_rigid_and_blend_closed_forms():
    rng = np.random.default_rng(0)
    p = rng.standard_normal((50, 3))
    one = (np.arange(51, dtype=np.int32), np.zeros(50, np.int32), np.ones(50))
    # identity
    np.testing.assert_allclose(ob.lbs(p, *one, _T(np.eye(3), [0, 0, 0])[None]), p, atol=1e-15)
    # translation
    np.testing.assert_allclose(ob.lbs(p, *one, _T(np.eye(3), [1, -2, 3])[None]), p + [1, -2, 3],
                               atol=1e-14)
    # rotation about z by 30 degrees
    R = _rot("z", np.pi / 6)
    np.testing.assert_allclose(ob.lbs(p, *one, _T(R, [0, 0, 0])[None]), p @ R.T, atol=1e-14)
    # two joints, complementary weights: w (R1 p + t1) + (1 - w)(R2 p + t2)
    w = rng.uniform(0, 1, 50)
    ptr = np.arange(0, 101, 2, dtype=np.int32)
    joint = np.tile([0, 1], 50).astype(np.int32)
    weight = np.stack([w, 1 - w], 1).reshape(-1)
    R1, t1, R2, t2 = _rot("x", 0.4), np.array([0.1, 0, 0]), _rot("z", -0.7), np.array([0, 0.2, 0.3])
    T = np.stack([_T(R1, t1), _T(R2, t2)])
    want = w[:, None] * (p @ R1.T + t1) + (1 - w)[:, None] * (p @ R2.T + t2)
    np.testing.assert_allclose(ob.lbs(p, ptr, joint, weight, T), want, atol=1e-14)


def test_lbs_matches_chain_body_kinematics():
    _, body = scenes.garment(seed=6, ring=24, rows=16, joints=4, gap=0.01, body_subdiv=2)
    ptr, joint, w = body.skinning()
    assert np.allclose(np.add.reduceat(w, ptr[:-1]), 1.0)  # partition of unity
    for t in (0.0, 0.25, 0.9):
        x = ob.lbs(body.pos, ptr, joint, w, body.transforms(t))
        np.testing.assert_allclose(x, body.at(t), atol=2e-7)


def test_triangle_closest_point_dense_sampling():
    rng = np.random.default_rng(1)
    for _ in range(200):
        a, b, c = rng.standard_normal((3, 3))
        p = rng.standard_normal(3) * 1.5
        d, uvw = ob.tri_nearest(p, a, b, c)
        assert uvw.min() >= -1e-6 and abs(uvw.sum() - 1) < 1e-5
        q = uvw[0] * a + uvw[1] * b + uvw[2] * c
        assert abs(np.linalg.norm(p - q) - d) < 1e-5
        s = rng.uniform(0, 1, (4000, 2))
        s = np.where(s.sum(1, keepdims=True) > 1, 1 - s, s)
        pts = a + s[:, :1] * (b - a) + s[:, 1:] * (c - a)
        assert d <= np.linalg.norm(pts - p, axis=1).min() + 1e-5


def _brute(points, v, tris):
    best = np.full(len(points), np.inf)
    arg = [set() for _ in points]
    for f, t in enumerate(tris):
        for k, p in enumerate(points):
            d, _ = ob.tri_nearest(p, v[t[0]], v[t[1]], v[t[2]])
            if d < best[k] - 1e-7:
                best[k], arg[k] = d, {f}
            elif abs(d - best[k]) <= 1e-7:
                arg[k].add(f)
    return best, arg


@pytest.mark.parametrize("tol", [0.0, 0.003])
def test_nearest_query_matches_brute_force(tol):
    pos, tris = scenes.icosphere(2, 0.3)
    n = orc.vertex_normals(pos, tris)
    bvh = orc.Bvh()
    bvh.Build(pos, tris, tol=tol, n=n)
    v = (np.asarray(pos, np.float32) + np.float32(tol) * n).astype(np.float32)  # corner points
    rng = np.random.default_rng(2)
    d = rng.standard_normal((150, 3))
    pts = (d / np.linalg.norm(d, axis=1, keepdims=True) * rng.uniform(0.05, 0.6, (150, 1)))
    pts = pts.astype(np.float32)
    faces, uvw, dist = ob.nearest(bvh, pts)
    best, arg = _brute(pts, v, tris)
    assert (faces >= 0).all()
    np.testing.assert_allclose(dist, best, atol=1e-6)
    assert all(int(f) in a for f, a in zip(faces, arg))
    for k in range(len(pts)):
        t = tris[faces[k]]
        q = uvw[k, 0] * v[t[0]] + uvw[k, 1] * v[t[1]] + uvw[k, 2] * v[t[2]]
        assert abs(np.linalg.norm(pts[k] - q) - dist[k]) < 1e-5
    ties = sum(len(a) > 1 for a in arg)
    print(f"  tol {tol}: {ties} of {len(pts)} points with tied nearest faces")


def test_binding_generate_and_solve_closed_forms():
    obs = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float)
    tri3 = np.array([[0, 1, 2]] * 4, np.int32)
    uvw = np.array([[1 / 3] * 3] * 4)
    c = obs.mean(0)
    cloth = np.array([c + [0, 0, 0.5], c + [0, 0, 0.2], c + [0, 0, 0.5], c + [0.3, 0, 0.4]])
    vert = np.arange(4, dtype=np.int32)
    # distance < EPSILON: taken from the rest state
    L = ob.generate(cloth, obs, vert, tri3, uvw, np.array([0.0, 0.0, 0.3, 1e-7]))
    np.testing.assert_allclose(L, [0.5, 0.2, 0.3, 0.5], atol=1e-12)
    x = cloth.copy()
    x[0, 2] += 0.1   # farther than L: back onto the sphere of radius L
    x[1, 2] -= 0.1   # inside: stays
    w = np.array([1.0, 1.0, 1.0, 0.0])  # the last vertex is pinned
    ob.solve(x, w, obs, vert, tri3, uvw, L, np.array([1.0, 1.0, 0.5, 1.0]))
    np.testing.assert_allclose(x[0], c + [0, 0, 0.5], atol=1e-12)
    np.testing.assert_allclose(x[1], cloth[1] - [0, 0, 0.1], atol=1e-12)
    # stiffness 0.5: half of the excess 0.5 - 0.3
    np.testing.assert_allclose(x[2], c + [0, 0, 0.4], atol=1e-12)
    np.testing.assert_allclose(x[3], cloth[3], atol=0)
