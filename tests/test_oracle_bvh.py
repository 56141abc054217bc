"""Pins of the FP32 LBVH oracle (CPU only): Morton known answers, stable sort order, a
top-down brute-force radix tree, brute-force pre-order escape successors, leaf-box unions,
brute-force point / segment / pair sets, refit / update growth, normals."""
import numpy as np
import pytest

from oracle import oracle as orc
from xrtailor_amd import scenes


def test_morton_known_answers():
    assert orc.expand_bits(0x3FF) == 0x09249249
    assert orc.expand_bits(1) == 1 and orc.expand_bits(2) == 8 and orc.expand_bits(3) == 9
    assert orc.morton_code(0.0, 0.0, 0.0) == 0
    assert orc.morton_code(1.0, 1.0, 1.0) == 0x3FFFFFFF      # clamp to 1023 per axis
    assert orc.morton_code(-5.0, 2.0, 0.0) == 0x3FFFFFFF & 0x12492492  # y only, clamped
    # x is the most significant axis of each triple: code = 4*X + 2*Y + Z
    assert orc.morton_code(0.5, 0.0, 0.0) == 4 << 27
    assert orc.morton_code(0.0, 0.0, 0.5) == 1 << 27
    assert orc.morton_code(float("nan"), 0.0, 0.0) == 0       # NaN -> 0 (ternary max)


def _brute_tree(codes):
    """Top-down radix tree over sorted unique 64-bit keys with Karras numbering:
    range [i, j] splits after the last key whose highest differing bit is 0."""
    n = len(codes)
    ni = n - 1
    kids = {}

    def rec(i, j, node):
        if i == j:
            return
        hb = int(codes[i] ^ codes[j]).bit_length() - 1
        g = i
        while g + 1 <= j and ((int(codes[g + 1]) >> hb) & 1) == 0:
            g += 1
        left = g if g > i else ni + g
        right = g + 1 if g + 1 < j else ni + g + 1
        kids[node] = (left, right)
        if g > i:
            rec(i, g, g)
        if g + 1 < j:
            rec(g + 1, j, g + 1)

    rec(0, n - 1, 0)
    return kids


def _preorder_escape(nodes):
    nn = len(nodes)
    order, stack = [], [0]
    while stack:
        k = stack.pop()
        order.append(k)
        if nodes[k, 3] < 0:
            stack.append(nodes[k, 2])
            stack.append(nodes[k, 1])
    size = {}

    def sub(k):
        if nodes[k, 3] >= 0:
            size[k] = 1
        else:
            size[k] = 1 + sub(nodes[k, 1]) + sub(nodes[k, 2])
        return size[k]

    sub(0)
    pos = {k: i for i, k in enumerate(order)}
    esc = np.full(nn, -1)
    for k in range(nn):
        nxt = pos[k] + size[k]
        esc[k] = order[nxt] if nxt < nn else -1
    return esc


def _leaf_boxes(x, tris):
    p = np.asarray(x, np.float32)[np.asarray(tris)]
    return np.concatenate([p.min(1), p.max(1)], 1)


@pytest.mark.parametrize("mesh", ["grid", "sphere", "dup"])
def test_bvh_structure_against_brute_force(mesh):
    rng = scenes.Rng(3)
    if mesh == "grid":
        pos, tris = scenes.grid(17, 1.0)
        pos = pos + 0.01 * rng.normal(pos.size).reshape(-1, 3)
    elif mesh == "sphere":
        pos, tris = scenes.icosphere(2, 0.5)
    else:  # every face duplicated: equal Morton codes force the 64-bit keys
        pos, tris = scenes.grid(6, 1.0)
        tris = np.concatenate([tris, tris])
    b = orc.Bvh()
    b.Build(pos, tris)
    h = b.host()
    nf = len(tris)
    leaf = _leaf_boxes(pos, tris)
    # morton codes in sorted order from the leaf centers, stable order on ties
    root = h["root"]
    np.testing.assert_array_equal(root[:3], leaf[:, :3].min(0))
    np.testing.assert_array_equal(root[3:], leaf[:, 3:].max(0))
    m = np.array([orc.morton_code(*((np.float32(0.5) * (leaf[f, :3] + leaf[f, 3:]) - root[:3])
                                    / (root[3:] - root[:3]))) for f in range(nf)], np.uint32)
    order = np.argsort(m, kind="stable")
    np.testing.assert_array_equal(h["order"], order)
    np.testing.assert_array_equal(h["morton"], m[order])
    if mesh == "dup":
        assert len(np.unique(m)) < nf
    # topology equals the top-down radix tree of the 64-bit keys
    codes = (h["morton"].astype(np.uint64) << np.uint64(32)) | h["order"].astype(np.uint64)
    kids = _brute_tree(codes)
    nodes = h["nodes"]
    ni = nf - 1
    for k in range(ni):
        assert (nodes[k, 1], nodes[k, 2]) == kids[k], k
        assert nodes[k, 3] == -1
    np.testing.assert_array_equal(nodes[ni:, 3], order)
    assert nodes[0, 0] == -1
    for k in range(1, 2 * nf - 1):
        p = nodes[k, 0]
        assert k in (nodes[p, 1], nodes[p, 2])
    # boxes: leaves = face boxes in sorted order, internal = union of descendant leaves
    aabb = h["aabbs"]
    np.testing.assert_array_equal(aabb[ni:], leaf[order])

    def union(k):
        if nodes[k, 3] >= 0:
            return leaf[nodes[k, 3]]
        a, c = union(nodes[k, 1]), union(nodes[k, 2])
        return np.concatenate([np.minimum(a[:3], c[:3]), np.maximum(a[3:], c[3:])])

    for k in range(ni):
        np.testing.assert_array_equal(aabb[k], union(k))
    # escape tables equal the pre-order skip successor
    np.testing.assert_array_equal(h["escape"], _preorder_escape(nodes))


def test_single_face_tree():
    pos, tris = scenes.grid(2, 1.0)
    b = orc.Bvh()
    b.Build(pos, tris[:1])
    h = b.host()
    assert h["nodes"].shape == (1, 4)
    np.testing.assert_array_equal(h["nodes"][0], [-1, -1, -1, 0])
    assert h["escape"][0] == -1
    assert [list(q) for q in b.Query(np.array([[0.0, 5.0, 0.0]]))] == [[]]


def _brute_point(leaf, p):
    return set(np.nonzero(np.all((p > leaf[:, :3]) & (p < leaf[:, 3:]), 1))[0])


def test_point_and_segment_queries_against_brute_force():
    rng = scenes.Rng(12)
    pos, tris = scenes.icosphere(3, 0.5)
    b = orc.Bvh()
    b.Build(pos, tris, tol=0.01, n=orc.vertex_normals(pos, tris))
    leaf = b.host()["aabbs"][len(tris) - 1:]
    order = b.host()["order"]
    lf = np.empty_like(leaf)
    lf[order] = leaf
    pts = (0.52 * rng.normal(300).reshape(-1, 3) / 1.7).astype(np.float32)
    for p, hits in zip(pts, b.Query(pts)):
        assert set(hits) == _brute_point(lf, p)
        assert len(hits) == len(set(hits))
    p0 = (0.6 * rng.normal(150).reshape(-1, 3)).astype(np.float32)
    p1 = (p0 + 0.3 * rng.normal(150).reshape(-1, 3)).astype(np.float32)
    for a, c, hits in zip(p0, p1, b.Query(p0, p1)):
        brute = {f for f in range(len(tris)) if orc.overlap_segment(lf[f], a, c)}
        assert set(hits) == brute


def test_segment_box_predicate_against_slab_test():
    """Bounds::Overlap(p1, p2) vs an independent slab test on segments in general position
    (the reference's test reports an endpoint-inside or a face crossing)."""
    rng = scenes.Rng(21)
    box = np.array([-0.3, -0.2, -0.1, 0.4, 0.3, 0.2], np.float32)
    agree = 0
    for _ in range(400):
        a = rng.normal(3).astype(np.float32)
        c = rng.normal(3).astype(np.float32)
        t0, t1 = 0.0, 1.0
        ok = True
        for k in range(3):
            d = float(c[k]) - float(a[k])
            lo, hi = box[k] - a[k], box[k + 3] - a[k]
            if abs(d) < 1e-12:
                ok &= lo < 0 < hi
            else:
                s0, s1 = sorted((lo / d, hi / d))
                t0, t1 = max(t0, s0), min(t1, s1)
        slab = ok and t0 < t1
        agree += slab == orc.overlap_segment(box, a, c)
    assert agree == 400


def test_pair_traversal_equals_brute_force_overlap():
    rng = scenes.Rng(8)
    pos, tris = scenes.grid(14, 1.0)
    pos = pos + 0.03 * rng.normal(pos.size).reshape(-1, 3)
    spos, stris = scenes.icosphere(2, 0.3, (0.0, -0.1, 0.0))
    a, s = orc.Bvh(), orc.Bvh()
    a.Build(pos, tris)
    s.Build(spos, stris)
    for lhs, rhs, lp, lt, rp, rt in [(a, s, pos, tris, spos, stris), (a, a, pos, tris, pos, tris)]:
        th = 2e-3
        pairs = orc.traverse(lhs, rhs, th)
        L, R = _leaf_boxes(lp, lt), _leaf_boxes(rp, rt)
        ov = np.all((R[None, :, :3] <= (L[:, None, 3:] + np.float32(th))) &
                    (R[None, :, 3:] >= (L[:, None, :3] - np.float32(th))), 2)
        brute = {tuple(p) for p in np.argwhere(ov)}
        assert {tuple(p) for p in pairs} == brute
        assert len(pairs) == len(brute)
    # self traversal contains (i, i) and both (i, j), (j, i)
    selfp = {tuple(p) for p in orc.traverse(a, a, 0.0)}
    assert all((i, i) in selfp for i in range(len(tris)))
    assert all((j, i) in selfp for (i, j) in selfp)


def test_refit_midstep_and_update_grow_only():
    pos, tris = scenes.icosphere(2, 0.5)
    n = orc.vertex_normals(pos, tris)
    b = orc.Bvh()
    b.Build(pos, tris)
    before = b.host()["aabbs"].copy()
    moved = (pos * 1.1).astype(np.float32)
    b.RefitMidstep(moved, n, 0.003)
    after = b.host()["aabbs"]
    assert np.all(after[:, :3] <= before[:, :3]) and np.all(after[:, 3:] >= before[:, 3:])
    corners = (moved + np.float32(0.003) * n).astype(np.float32)
    leaf = b.host()["aabbs"][len(tris) - 1:]
    tri = b.host()["order"]
    for k in range(0, len(tris), 37):
        f = tri[k]
        c = corners[tris[f]]
        assert np.all(leaf[k, :3] <= c.min(0)) and np.all(leaf[k, 3:] >= c.max(0))
    # Update(ccd): leaves cover the faces at both x and x0
    b2 = orc.Bvh()
    b2.Build(pos, tris)
    x = (pos + np.array([0.0, 0.2, 0.0])).astype(np.float32)
    b2.Update(x, pos, ccd=True)
    h = b2.host()
    leaf = h["aabbs"][len(tris) - 1:]
    for k, f in enumerate(h["order"]):
        pts = np.concatenate([x[tris[f]], pos[tris[f]].astype(np.float32)])
        assert np.all(leaf[k, :3] <= pts.min(0)) and np.all(leaf[k, 3:] >= pts.max(0))
    assert np.all(h["aabbs"][0, :3] <= leaf[:, :3].min(0))


def test_vertex_normals_closed_forms():
    pos, tris = scenes.grid(9, 1.0)
    n = orc.vertex_normals(pos, tris)
    np.testing.assert_allclose(n, np.tile([0.0, 1.0, 0.0], (len(pos), 1)), atol=1e-6)
    sp, st = scenes.icosphere(3, 0.7)
    n = orc.vertex_normals(sp, st)
    radial = sp / np.linalg.norm(sp, axis=1, keepdims=True)
    assert np.min(np.sum(n * radial, 1)) > 0.995
    np.testing.assert_allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-6)
