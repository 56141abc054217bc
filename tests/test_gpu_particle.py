"""GPU swift-mode particle self-collision (spatial hash + CollideParticles) against the
oracle, through the C ABI: the hash state bit for bit from the same FP32 positions, one
collision pass and whole substeps against the FP64 twin, batched == plain bit for bit."""
import numpy as np
import pytest

from contact_helpers import oracle_scene
from gpu_helpers import build, rel_err
from oracle import particle as opart
from xrtailor_amd import XrtError, scenes
from xrtailor_amd.solver import make_params

pytestmark = pytest.mark.gpu

SENT = 0xFFFFFFFF
TOL = 1e-5


def particle_layers(n=16, side=0.3, rest_gap=0.1, gap=0.6, angle=0.3, v_down=-0.2, seed=7):
    """Two sheets in one mesh, rest_gap apart at rest (the initial positions), and a state
    with the upper sheet `gap` particle diameters above the lower one, turned by `angle`
    about y so that the two lattices do not line up, moving down at v_down."""
    pos, tris = scenes.grid(n, side, 0.0)
    rest = np.concatenate([pos, pos + np.array([0.0, rest_gap, 0.0])])
    T = np.concatenate([tris, tris + len(pos)]).astype(np.int32)
    spec = scenes.ClothSpec(rest, T, np.zeros(0, np.int32), 1e-7, 1e-3, rest.astype(np.float32))
    d = float(opart.particle_diameter(rest, T, 1.3))
    c, s = np.cos(angle), np.sin(angle)
    up = pos.copy()
    up[:, 0], up[:, 2] = c * pos[:, 0] - s * pos[:, 2], s * pos[:, 0] + c * pos[:, 2]
    up[:, 1] += gap * d
    state = np.concatenate([pos, up]) + 1e-4 * np.random.default_rng(seed).standard_normal(rest.shape)
    v = np.zeros_like(rest)
    v[len(pos):, 1] = v_down
    return spec, state.astype(np.float32), v.astype(np.float32)


def crumpled(n=24, side=0.5, radius=0.04, seed=9):
    """A flat sheet at rest whose state is scattered in a small ball: long cells, full
    neighbor lists, truncation at max_num_neighbors per cell."""
    pos, tris = scenes.grid(n, side, 0.0)
    spec = scenes.ClothSpec(pos, tris, np.zeros(0, np.int32), 1e-7, 1e-3, pos.astype(np.float32))
    rng = np.random.default_rng(seed)
    state = rng.uniform(-radius, radius, pos.shape)
    return spec, state.astype(np.float32), np.zeros_like(state, np.float32)


def _params(**kw):
    base = dict(enable_particle_collision=1, bend_model=1, enable_ground=0, ground_friction=0.1)
    base.update(kw)
    return make_params(**base)


def _hash_case(spec, state, v, xp):
    s, (i,) = build([spec], xp, seed=1)
    s.SetState(i, state, v)
    dt = xp.frame_time / xp.num_substeps
    s.PredictPositions(i, dt)
    sh = s.SpatialHash(i)
    sh.Hash()
    return s, i, sh.buffers(), s.GetPredicted(i)


def _listed(nb):
    return (nb != SENT) & (np.cumsum(nb == SENT, axis=0) == 0)


def _assert_lists_equal(got, want):
    """Neighbor lists equal up to and including each list's sentinel.  Entries after it are
    whatever an earlier hash of the same buffers left there (nothing reads them): the
    state reset of xrt_set_state hashes once before the test's own hash."""
    upto = (np.cumsum(want == SENT, axis=0) - (want == SENT)) == 0
    np.testing.assert_array_equal(got[upto], want[upto])


@pytest.mark.parametrize("scene,max_nb", [("layers", 64), ("layers", 4), ("crumpled", 16)])
def test_hash_matches_oracle(scene, max_nb):
    print(f"case particle hash {scene} max_nb={max_nb}", flush=True)
    spec, state, v = particle_layers() if scene == "layers" else crumpled()
    xp = _params(max_num_neighbors=max_nb)
    s, i, g, x = _hash_case(spec, state, v, xp)
    d = opart.particle_diameter(spec.pos, spec.tris, xp.particle_diameter)
    assert np.float32(g["diameter"]) == d
    want = opart.neighbors(x, spec.x0, d, xp.hash_cell_size, max_nb)
    np.testing.assert_array_equal(g["particle_hash"], want["particle_hash"])
    np.testing.assert_array_equal(g["particle_index"], want["particle_index"])
    np.testing.assert_array_equal(g["cell_start"], want["cell_start"])
    set_ = want["cell_start"] != SENT
    np.testing.assert_array_equal(g["cell_end"][set_], want["cell_end"][set_])
    _assert_lists_equal(g["neighbors"], want["neighbors"])
    lst = _listed(want["neighbors"])
    print(f"  {len(x)} particles, {lst.sum()} neighbor entries, {lst.all(axis=0).sum()} full "
          f"lists, {set_.sum()} cells", flush=True)
    assert lst.sum() > 0
    s.close()


def test_collide_matches_oracle():
    print("case particle collide: one pass vs FP64 oracle", flush=True)
    spec, state, v = particle_layers()
    xp = _params()
    s, i, g, x = _hash_case(spec, state, v, xp)
    x0, _ = s.GetState(i)
    s.CollideParticles(i)
    got = s.GetPredicted(i)
    w = s.inv_mass(i)
    delta, cnt = opart.collide(x, x0, w, g["neighbors"], g["diameter"], xp.ground_friction)
    ref = x.astype(np.float64)
    hit = cnt > 0
    ref[hit] += delta[hit] / cnt[hit, None] * xp.relaxation_factor
    err = rel_err(got, ref)
    print(f"  {hit.sum()} particles in contact, rel err {err:.3e}", flush=True)
    assert hit.sum() > 50
    assert err <= TOL
    # the untouched particles did not move at all
    np.testing.assert_array_equal(got[~hit], x[~hit])
    s.close()


def test_substeps_match_oracle_twin():
    """Plain substeps (hash every inter_leaved_hash = 3 substeps, collision every
    iteration, friction 0.1) against the oracle twin."""
    print("case particle substeps vs oracle twin", flush=True)
    spec, state, v = particle_layers()
    xp = _params()
    s, (i,) = build([spec], xp, seed=1)
    s.SetState(i, state, v)
    twin = oracle_scene(spec, xp, solver=s, inst=i)
    twin.p_orig = spec.x0.astype(np.float32)  # the state at xrt_start
    errs = []
    for _ in range(4):
        s.SubstepPlain(i)
        twin.substep()
        errs.append(rel_err(s.GetState(i)[0], twin.c.state()[0]))
    print(f"  rel err per substep {', '.join(f'{e:.2e}' for e in errs)}", flush=True)
    assert errs[0] <= TOL
    assert max(errs) <= 1e-4
    # the layers were pushed apart: no inter-layer pair closer than half a diameter
    x = s.GetState(i)[0]
    n = len(x) // 2
    d = float(opart.particle_diameter(spec.pos, spec.tris, xp.particle_diameter))
    dmin = np.min(np.linalg.norm(x[:n, None, :] - x[None, n:, :], axis=2))
    print(f"  closest inter-layer pair {dmin / d:.3f} diameters", flush=True)
    assert dmin > 0.5 * d
    s.close()


@pytest.mark.parametrize("self_contact", [0, 1])
def test_batched_equals_plain(self_contact):
    """Instances of three sizes: graph-captured batched frames equal plain frames bit for
    bit (with self contact the wide path also runs the PBD repulsion)."""
    print(f"case particle batched vs plain, self contact {self_contact}", flush=True)
    cases = [particle_layers(n=12, seed=1), particle_layers(n=16, seed=2),
             particle_layers(n=20, angle=0.5, seed=3)]
    xp = _params(enable_self_contact=self_contact, num_substeps=4, num_iterations=4,
                 inter_leaved_hash=3)
    runs = []
    for batched in (True, False):
        s, insts = build([c[0] for c in cases], xp, seed=1)
        for k, (_, st, v) in zip(insts, cases):
            s.SetState(k, st, v)
        if batched:
            s.FramesBatched(2)
            s.SubstepBatched()
        else:
            s.FramesPlain(2)
            for k in insts:
                s.SubstepPlain(k)
        runs.append([s.GetState(k)[0] for k in insts])
        s.close()
    for k, (a, b) in enumerate(zip(*runs)):
        print(f"  instance {k}: max |batched - plain| = {np.abs(a - b).max():.3e}", flush=True)
        np.testing.assert_array_equal(a, b)


def test_particle_params_and_errors():
    print("case particle parameters and error paths", flush=True)
    spec, state, v = particle_layers()
    xp = _params(enable_particle_collision=0)
    s, (i,) = build([spec], xp, seed=1)
    with pytest.raises(XrtError):
        s.SpatialHash(i).Hash()  # off
    with pytest.raises(XrtError):
        s.set_params(_params(max_num_neighbors=0))
    # switched on after xrt_start, with a different list length: buffers sized then
    s.set_params(_params(max_num_neighbors=8))
    s.SetState(i, state, v)
    s.PredictPositions(i, xp.frame_time / xp.num_substeps)
    sh = s.SpatialHash(i)
    sh.Hash()
    g = sh.buffers()
    assert g["neighbors"].shape == (8, len(state))
    want = opart.neighbors(s.GetPredicted(i), spec.x0, g["diameter"], 1.5, 8)
    _assert_lists_equal(g["neighbors"], want["neighbors"])
    with pytest.raises(XrtError):
        s.AdjointBegin(1, 1)  # the adjoint has no particle collision
    s.close()


def test_time_particles_counts():
    """The timing hook: device times of the collision kernel and the hash, the list-entry
    count equal to the lists' lengths read back."""
    print("case particle timing hook counts", flush=True)
    cases = [particle_layers(n=12, seed=1), particle_layers(n=16, seed=2)]
    xp = _params(num_substeps=3, num_iterations=2)
    s, insts = build([c[0] for c in cases], xp, seed=1)
    for k, (_, st, v) in zip(insts, cases):
        s.SetState(k, st, v)
    col_ms, hash_ms, c = s.time_particles(3)
    print(f"  collide {col_ms * 1e3:.1f} us, hash {hash_ms * 1e3:.1f} us, {c}", flush=True)
    assert col_ms > 0 and hash_ms > 0
    assert c["particles"] == sum(len(cc[1]) for cc in cases)
    total = sum(int(_listed(s.SpatialHash(k).buffers()["neighbors"]).sum()) for k in insts)
    assert c["entries"] == total > 0
    s.close()
