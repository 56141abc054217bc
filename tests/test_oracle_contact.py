"""Pins of the contact oracle (CPU only): narrow-phase routines against dense sampling and
closed forms, R-triangle ownership, single-contact response closed forms."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as orc
from oracle import scene as osc
from xrtailor_amd import scenes

_D = C.POINTER(C.c_double)


def _d(a):
    return np.ascontiguousarray(a, np.float64).ctypes.data_as(_D)


def _tri_samples(tri, n=300):
    u, v = np.meshgrid(np.linspace(0, 1, n), np.linspace(0, 1, n))
    m = (u + v) <= 1
    u, v = u[m], v[m]
    return tri[0] + u[:, None] * (tri[1] - tri[0]) + v[:, None] * (tri[2] - tri[0])


def test_nearest_point_against_dense_sampling():
    lib = osc._lib()
    rng = scenes.Rng(2)
    for _ in range(40):
        tri = rng.normal(9).reshape(3, 3)
        p = 1.5 * rng.normal(3)
        q = np.zeros(3)
        qq = np.ascontiguousarray(q)
        d = lib.orc_nearest_point(_d(p), _d(tri.reshape(-1)), qq.ctypes.data_as(_D))
        samp = _tri_samples(tri)
        brute = np.linalg.norm(samp - p, axis=1).min()
        assert d <= brute + 1e-12
        assert d >= brute - 5e-3 * np.linalg.norm(tri[1] - tri[0])
        assert abs(np.linalg.norm(p - qq) - d) < 1e-12


def test_ray_intersect_closed_form():
    lib = osc._lib()
    tri = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0]])  # z = 0 plane
    o = np.array([0.2, 0.3, 1.0])
    d = np.array([0.0, 0.0, -1.0])
    assert lib.orc_ray_intersect(_d(o), _d(d), 2.0, _d(tri.reshape(-1))) == pytest.approx(1.0)
    assert np.isinf(lib.orc_ray_intersect(_d(o), _d(d), 0.5, _d(tri.reshape(-1))))   # too short
    o2 = np.array([0.8, 0.8, 1.0])                                                    # outside
    assert np.isinf(lib.orc_ray_intersect(_d(o2), _d(d), 2.0, _d(tri.reshape(-1))))
    d2 = np.array([1.0, 0.0, 0.0])                                                    # parallel
    assert np.isinf(lib.orc_ray_intersect(_d(o), _d(d2), 2.0, _d(tri.reshape(-1))))


def test_closest_segment_segment_against_sampling():
    lib = osc._lib()
    rng = scenes.Rng(5)
    s = np.zeros(1)
    t = np.zeros(1)
    for _ in range(40):
        a, b, c, d = rng.normal(12).reshape(4, 3)
        dist = lib.orc_closest_seg_seg(_d(a), _d(b), _d(c), _d(d), s.ctypes.data_as(_D), t.ctypes.data_as(_D))
        g = np.linspace(0, 1, 401)
        P = a + g[:, None] * (b - a)
        Q = c + g[:, None] * (d - c)
        brute = np.linalg.norm(P[:, None] - Q[None], axis=2).min()
        assert dist <= brute + 1e-12
        assert dist >= brute - 1e-2
        assert 0 <= s[0] <= 1 and 0 <= t[0] <= 1
        assert abs(np.linalg.norm((a + s[0] * (b - a)) - (c + t[0] * (d - c))) - dist) < 1e-12


@pytest.mark.parametrize("mesh", ["grid", "sphere"])
def test_rtri_owns_every_vertex_and_edge_once(mesh):
    pos, tris = scenes.grid(9, 1.0) if mesh == "grid" else scenes.icosphere(2, 1.0)
    edges = orc.build_edges(tris)
    fe, ptr, ef, rt, fn = osc.contact_topology(pos, tris, edges)
    vert_owner = np.zeros(len(pos), int)
    edge_owner = np.zeros(len(edges), int)
    for f, bits in enumerate(rt):
        for i in range(3):
            if (bits >> i) & 1:
                vert_owner[tris[f][i]] += 1
            if (bits >> (3 + i)) & 1:
                edge_owner[fe[3 * f + i]] += 1
    assert np.all(vert_owner == 1)
    assert np.all(edge_owner == 1)
    # every face edge id names the edge between its i-th and (i+1)-th vertices
    for f in range(0, len(tris), 7):
        for i in range(3):
            a, b = tris[f][i], tris[f][(i + 1) % 3]
            assert set(edges[fe[3 * f + i]]) == {a, b}
    # faces per edge ascending
    for e in range(len(edges)):
        fs = ef[ptr[e]:ptr[e + 1]]
        assert list(fs) == sorted(fs)


def test_signed_distances_closed_form():
    lib = osc._lib()
    x = np.array([[0.2, 0.3, 0.5], [0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float64)
    n = np.zeros(3)
    w = np.zeros(4)
    h = lib.orc_signed_vf(_d(x.reshape(-1)), n.ctypes.data_as(_D), w.ctypes.data_as(_D))
    assert h == pytest.approx(0.5)
    np.testing.assert_allclose(n, [0, 0, 1])
    np.testing.assert_allclose(-w[1:], [0.5, 0.2, 0.3])  # barycentric of (0.2,0.3)
    # edge-edge: crossing edges 0.25 apart; n is the raw cross of unit directions
    e = np.array([[0, 0, 0.25], [1, 0, 0.25], [0.5, -0.5, 0], [0.5, 0.5, 0]], np.float64)
    h = lib.orc_signed_ee(_d(e.reshape(-1)), n.ctypes.data_as(_D), w.ctypes.data_as(_D))
    assert h == pytest.approx(0.25)
    np.testing.assert_allclose(n, [0, 0, 1])
    np.testing.assert_allclose(w, [0.5, 0.5, -0.5, -0.5])


def _plane_obstacle(tol=0.003):
    """Flat 2x2 m square at y = 0 (+y normal): built at tol 0, refit with tol, so its boxes
    span the shell y in [0, tol] and the contact face sits at y = tol (pred corners)."""
    opos = np.array([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], np.float32)
    otris = np.array([[0, 2, 1], [0, 3, 2]], np.int32)
    b = orc.Bvh()
    b.Build(opos, otris, tol=0.0)
    b.RefitMidstep(opos, orc.vertex_normals(opos, otris), tol)
    return b, otris


_IP = C.POINTER(C.c_int32)


def test_vertex_in_the_contact_shell_is_projected_above_it():
    """VF-DCD: a vertex inside the shell is pushed to 1e-3 above the pred face
    (lambda = n.(x - q) - 1e-3 < 0, x -= lambda n); one below the shell is not a DCD hit."""
    lib = osc._lib()
    b, otris = _plane_obstacle()
    x = np.array([[0.1, 0.001, 0.2], [0.3, -0.02, 0.1]], np.float64)
    face = np.zeros(2, np.int32)
    q = np.zeros(6)
    lib.orc_vf_dcd_generate(b.h, otris.ctypes.data_as(_IP), _d(x), 2, face.ctypes.data_as(_IP),
                            q.ctypes.data_as(_D))
    assert face[0] >= 0 and face[1] == -1
    xx = np.ascontiguousarray(x.reshape(-1))
    lib.orc_vf_solve(lib.orc_bvh_pred(b.h), otris.ctypes.data_as(_IP), face.ctypes.data_as(_IP),
                     q.ctypes.data_as(_D), xx.ctypes.data_as(_D), 2, 1e-3)
    assert xx[1] == pytest.approx(0.003 + 1e-3, abs=1e-9)
    assert xx[0] == pytest.approx(0.1) and xx[2] == pytest.approx(0.2)
    assert xx[4] == pytest.approx(-0.02)


def test_ccd_catches_a_vertex_that_tunnels_through():
    """VF-CCD: segment x0 -> x crossing the pred face; the solve (offset 0) puts the vertex
    back on the face plane."""
    lib = osc._lib()
    b, otris = _plane_obstacle()
    x0 = np.array([[0.1, 0.1, 0.2]], np.float64)
    x = np.array([[0.12, -0.05, 0.2]], np.float64)
    face = np.zeros(1, np.int32)
    q = np.zeros(3)
    lib.orc_vf_ccd_generate(b.h, otris.ctypes.data_as(_IP), _d(x0), _d(x), 1,
                            face.ctypes.data_as(_IP), q.ctypes.data_as(_D))
    assert face[0] >= 0
    assert q[1] == pytest.approx(0.003, abs=1e-7)
    xx = np.ascontiguousarray(x.reshape(-1))
    lib.orc_vf_solve(lib.orc_bvh_pred(b.h), otris.ctypes.data_as(_IP), face.ctypes.data_as(_IP),
                     q.ctypes.data_as(_D), xx.ctypes.data_as(_D), 1, 0.0)
    assert xx[1] == pytest.approx(0.003, abs=1e-7)
