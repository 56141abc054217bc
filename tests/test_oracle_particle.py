"""The oracle's swift-mode particle self-collision (oracle/particle_oracle.c), pinned
without the code under test:

  * HashCoords against hand-computed values (wrapping products, C remainder sign)
  * neighbor lists against a brute-force restatement that groups particles by hash
    directly (no sort, no cell table): per-cell truncation to the first max_nb particles
    by index, the full-list early return, the sentinel, hash collisions visited twice
  * the cell table against the sorted keys
  * CollideParticles + ApplyDeltas against two-particle closed forms
"""
import math

import numpy as np
import pytest

from oracle import particle as opart

SENT = 0xFFFFFFFF


def _hash_ref(x, y, z, table):
    h = ((x * 92837111) ^ (y * 689287499) ^ (z * 283923481)) & 0xFFFFFFFF
    if h >= 1 << 31:
        h -= 1 << 32
    r = int(math.fmod(h, table))  # C remainder: sign of the dividend
    return abs(r)


def test_hash_coords_known_answers():
    assert opart.hash_coords(0, 0, 0, 1000) == 0
    assert opart.hash_coords(1, 0, 0, 1000) == 92837111 % 1000 == 111
    # -92837111 % 100 in C is -11; abs -> 11
    assert opart.hash_coords(-1, 0, 0, 100) == 11
    # 100 * 92837111 = 9283711100 wraps to 9283711100 - 2 * 2**32 = 693776508
    assert opart.hash_coords(100, 0, 0, 1 << 30) == 693776508 % (1 << 30)
    # a product that wraps to a negative int32: 30 * 92837111 = 2785113330 -> -1509853966
    assert opart.hash_coords(30, 0, 0, 7) == abs(int(math.fmod(-1509853966, 7)))
    rng = np.random.default_rng(0)
    for x, y, z in rng.integers(-5000, 5000, size=(200, 3)):
        for table in (2, 17, 2 * 4096, 131072):
            assert opart.hash_coords(int(x), int(y), int(z), table) == _hash_ref(int(x), int(y), int(z), table)


def _f32(a):
    return np.asarray(a, np.float32)


def _brute_neighbors(x, orig, d, cell, max_nb):
    """Reference restatement by definition: the 27 cells around each particle (x-major),
    each cell's members (particles whose own cell hashes to the same slot) in index order,
    at most max_nb of them, filtered by the two FP32 distance tests."""
    x, orig = _f32(x), _f32(orig)
    n = len(x)
    table = 2 * n
    d = np.float32(d)
    sp = np.float32(d * np.float32(cell))
    sp2, d2 = np.float32(sp * sp), np.float32(d * d)
    ic = np.floor(x / sp).astype(np.int64)
    members = {}
    for i in range(n):
        members.setdefault(_hash_ref(*map(int, ic[i]), table), []).append(i)

    def len2(a, b):
        e = a - b
        return np.float32(np.float32(e[0] * e[0]) + np.float32(e[1] * e[1])) + np.float32(e[2] * e[2])

    out = np.zeros((max_nb, n), np.uint32)
    for i in range(n):
        lst = []
        full = False
        for cx in range(ic[i, 0] - 1, ic[i, 0] + 2):
            for cy in range(ic[i, 1] - 1, ic[i, 1] + 2):
                for cz in range(ic[i, 2] - 1, ic[i, 2] + 2):
                    if full:
                        break
                    for j in members.get(_hash_ref(cx, cy, cz, table), [])[:max_nb]:
                        if j == i or not len2(x[i], x[j]) < sp2 or not len2(orig[i], orig[j]) > d2:
                            continue
                        lst.append(j)
                        if len(lst) == max_nb:
                            full = True
                            break
        out[: len(lst), i] = lst
        if len(lst) < max_nb:
            out[len(lst), i] = SENT
    return out


def _cloud(n, box, seed):
    rng = np.random.default_rng(seed)
    return _f32(rng.uniform(-box, box, size=(n, 3)))


@pytest.mark.parametrize("case", ["cloud64", "truncate4", "crumpled", "orig_close", "small_table"])
def test_neighbors_match_brute_force(case):
    d, cell = 0.02, 1.5
    if case == "cloud64":
        x, max_nb = _cloud(400, 0.08, 1), 64
        orig = x + _f32(np.random.default_rng(2).uniform(-1, 1, x.shape))
    elif case == "truncate4":
        x, max_nb = _cloud(400, 0.06, 3), 4
        orig = x + _f32(np.random.default_rng(4).uniform(-1, 1, x.shape))
    elif case == "crumpled":  # every particle within a cell or two: long cells, full lists
        x, max_nb = _cloud(300, 0.015, 5), 16
        orig = _cloud(300, 10.0, 6)
    elif case == "orig_close":  # initial positions equal to the current ones
        x, max_nb = _cloud(300, 0.06, 7), 64
        orig = x.copy()
    else:  # 24 particles, a 48-slot table: neighborhood cells collide, members listed twice
        x, max_nb = _cloud(24, 0.03, 8), 64
        orig = _cloud(24, 10.0, 9)
    got = opart.neighbors(x, orig, d, cell, max_nb)
    want = _brute_neighbors(x, orig, d, cell, max_nb)
    np.testing.assert_array_equal(got["neighbors"], want)
    nb = got["neighbors"]
    listed = (nb != SENT) & (np.cumsum(nb == SENT, axis=0) == 0)
    print(f"  {case}: {listed.sum()} entries, {np.sum(listed.all(axis=0))} full lists")
    if case == "crumpled":
        assert listed.all(axis=0).any()  # some lists end full, without a sentinel
    if case == "small_table":
        assert any(len(set(nb[listed[:, i], i])) < listed[:, i].sum() for i in range(len(x)))
    if case == "orig_close":  # close pairs at the start are never listed
        for i in range(len(x)):
            for j in nb[listed[:, i], i]:
                e = orig[i] - orig[j]
                assert (e * e).sum() > np.float32(d * d) * 0.99


def test_cell_table_and_sort_order():
    x = _cloud(500, 0.1, 11)
    g = opart.neighbors(x, x + 5.0, 0.02, 1.5, 8)
    h, idx = g["particle_hash"], g["particle_index"]
    order = np.lexsort((idx, h))  # stable sort by hash: ascending index within a hash
    np.testing.assert_array_equal(order, np.arange(len(h)))
    counts = np.bincount(h, minlength=2 * len(x))
    cs, ce = g["cell_start"], g["cell_end"]
    for c in range(2 * len(x)):
        if counts[c] == 0:
            assert cs[c] == SENT
        else:
            assert ce[c] - cs[c] == counts[c] and (h[cs[c]:ce[c]] == c).all()


def _two(xa, xb, wa=1.0, wb=1.0, va=(0, 0, 0), vb=(0, 0, 0)):
    x = np.array([xa, xb], np.float64)
    x0 = x - np.array([va, vb], np.float64)
    nb = np.array([[1, 0], [SENT, SENT]], np.uint32)
    return x, x0, np.array([wa, wb]), nb


def test_collide_two_particles_closed_forms():
    d = 0.01
    # equal masses at 0.6 d: each moves 0.2 d * dist / (dist + eps) apart
    x, x0, w, nb = _two((0, 0, 0), (0.6 * d, 0, 0))
    delta, cnt = opart.collide(x, x0, w, nb, d, 0.0)
    s = 0.2 * d * (0.6 * d) / (0.6 * d + 1e-6)
    np.testing.assert_allclose(delta, [[-s, 0, 0], [s, 0, 0]], rtol=1e-12, atol=1e-18)
    np.testing.assert_array_equal(cnt, [1, 1])
    # a pinned partner: the free particle moves the whole overlap
    x, x0, w, nb = _two((0, 0, 0), (0, 0.5 * d, 0), wa=1.0, wb=0.0)
    delta, cnt = opart.collide(x, x0, w, nb, d, 0.0)
    s = 0.5 * d * (0.5 * d) / (0.5 * d + 1e-6)
    np.testing.assert_allclose(delta[0], [0, -s, 0], rtol=1e-12)
    np.testing.assert_allclose(delta[1], [0, 0, 0], atol=0)
    np.testing.assert_array_equal(cnt, [1, 1])  # the pinned one counts a zero delta
    # both pinned: skipped, no count
    x, x0, w, nb = _two((0, 0, 0), (0.5 * d, 0, 0), wa=0.0, wb=0.0)
    delta, cnt = opart.collide(x, x0, w, nb, d, 0.0)
    assert not delta.any() and not cnt.any()
    # apart: no contact
    x, x0, w, nb = _two((0, 0, 0), (1.01 * d, 0, 0))
    delta, cnt = opart.collide(x, x0, w, nb, d, 0.3)
    assert not delta.any() and not cnt.any()


def test_collide_friction_closed_form():
    d, mu = 0.01, 0.1
    v = 0.02  # tangential (z) motion of particle 0 this substep, partner at rest
    x, x0, w, nb = _two((0, 0, 0), (0.5 * d, 0, 0), va=(0, 0, v))
    delta, _ = opart.collide(x, x0, w, nb, d, mu)
    corr = 0.25 * d * (0.5 * d) / (0.5 * d + 1e-6)  # |lambda * gradient|
    # Coulomb cap: the tangential slip v is cut by min(mu |corr| / v, 1) of itself
    np.testing.assert_allclose(delta[0], [-corr, 0, -v * min(mu * corr / v, 1.0)], rtol=1e-12)
    # the partner sees the opposite relative velocity
    np.testing.assert_allclose(delta[1], [corr, 0, v * min(mu * corr / v, 1.0)], rtol=1e-12)
    # a large friction coefficient stops the slip entirely
    delta, _ = opart.collide(x, x0, w, nb, d, 1e3)
    np.testing.assert_allclose(delta[0], [-corr, 0, -v], rtol=1e-12)


def test_collide_stops_at_sentinel_and_counts():
    d = 0.01
    x = np.array([[0, 0, 0], [0.5 * d, 0, 0], [0, 0.5 * d, 0], [0, 0, 0.5 * d]], np.float64)
    nb = np.full((4, 4), SENT, np.uint32)
    nb[:, 0] = [1, 2, SENT, 3]  # the entry after the sentinel is never read
    delta, cnt = opart.collide(x, x.copy(), np.ones(4), nb, d, 0.0)
    assert cnt[0] == 2
    s = 0.25 * d * (0.5 * d) / (0.5 * d + 1e-6)
    np.testing.assert_allclose(delta[0], [-s, -s, 0], rtol=1e-12)


def test_particle_diameter_fp32():
    pos = np.array([[0, 0, 0], [0.1, 0.2, 0.3], [1, 0, 0]], np.float32)
    d = opart.particle_diameter(pos, np.array([[0, 1, 2]]), 1.3)
    assert d.dtype == np.float32
    assert d == np.float32(np.sqrt(np.float32(0.14), dtype=np.float32) * np.float32(1.3)) or \
        abs(float(d) - math.sqrt(0.14) * 1.3) < 1e-7
