"""GPU parity of the XPBD substep (plain per-color launches and the batched persistent
launch) against the FP64 oracle, through the C ABI.

Bar: single-substep FP32 positions within 1e-5 relative of the oracle
(||x_gpu - x_orc||_inf / max(||x_orc||_inf, 1 m)); rollouts report max / RMS error.
"""
import numpy as np
import pytest

from gpu_helpers import build, make_params, oracle_of, rel_err
from xrtailor_amd import XrtError, scenes

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _print(msg):
    print(msg, flush=True)


def _bend_compliance(bend_model, c):
    """The reference bend formula is only stable for large compliance (the reference ships
    1e2): at 1e-4 it multiplies rounding differences by ~1e3 per substep, so its parity
    is checked at compliance >= 1 and the consistent model at the configured values."""
    return c if bend_model == 1 else max(1.0, c * 1e4)


@pytest.mark.parametrize("bend_model", [0, 1])
@pytest.mark.parametrize("path", ["plain", "batched"])
def test_hanging_cloth_single_substep_matches_oracle(bend_model, path):
    _print(f"case hanging 64x64 bend_model={bend_model} path={path}")
    spec = scenes.hanging_cloth(seed=1, n=64, bend=_bend_compliance(bend_model, 1e-4))
    params = make_params(bend_model=bend_model)
    s, (i,) = build([spec], params, seed=5)
    orc_cloth = oracle_of(s, i, spec.stretch_compliance, spec.bend_compliance)
    from oracle import oracle as orc
    op, dt = orc.params_from(params)
    if path == "plain":
        s.SubstepPlain(i)
    else:
        s.SubstepBatched()
    x0, v = s.GetState(i)
    orc_cloth.substep(op, dt)
    ox, ov = orc_cloth.state()
    err = rel_err(x0, ox)
    _print(f"  rel err {err:.3e}")
    assert err <= TOL
    # velocities follow from positions: (x - x0)/dt*(1 - damping*dt)
    assert np.abs(v - ov).max() <= TOL / dt * 2


@pytest.mark.parametrize("coloring", [0, 1])
def test_topology_and_colors_match_oracle(coloring):
    from oracle import oracle as orc
    spec = scenes.hanging_cloth(seed=1, n=32)
    s, (i,) = build([spec], make_params(coloring=coloring, bend_model=1), seed=11)
    edges, rest, scol = s.edges(i)
    np.testing.assert_array_equal(edges, orc.build_edges(spec.tris))
    col, n = orc.color(edges, len(spec.pos), 11, coloring)
    np.testing.assert_array_equal(scol, col)
    bends, theta, bcol = s.bends(i)
    col, n = orc.color(bends, len(spec.pos), 11, coloring)
    np.testing.assert_array_equal(bcol, col)
    # a substep under either coloring meets the parity bar
    oc = oracle_of(s, i, spec.stretch_compliance, spec.bend_compliance)
    op, dt = orc.params_from(s.params)
    s.SubstepBatched()
    oc.substep(op, dt)
    assert rel_err(s.GetState(i)[0], oc.state()[0]) <= TOL
    np.testing.assert_allclose(rest, orc.rest_lengths(spec.pos, edges), rtol=1e-6)
    np.testing.assert_allclose(theta, orc.rest_angles(spec.pos, bends), atol=1e-6)


def test_hanging_cloth_rollout_error_report():
    """10 frames (100 substeps): report max / RMS vertex error per frame vs the oracle."""
    from oracle import oracle as orc
    spec = scenes.hanging_cloth(seed=1, n=64)
    params = make_params(bend_model=1)
    s, (i,) = build([spec], params, seed=5)
    oc = oracle_of(s, i, spec.stretch_compliance, spec.bend_compliance)
    op, dt = orc.params_from(params)
    worst = 0.0
    for f in range(10):
        for _ in range(params.num_substeps):
            s.SubstepPlain(i)
            oc.substep(op, dt)
        x0, _ = s.GetState(i)
        ox, _ = oc.state()
        d = np.linalg.norm(x0 - ox, axis=1)
        _print(f"  frame {f}: max {d.max():.3e} m  rms {np.sqrt((d ** 2).mean()):.3e} m")
        worst = max(worst, d.max())
    assert worst < 1e-3


def _mixed_specs(bend_model=1):
    """Instances of uneven size; the largest exceeds the LDS-resident limit."""
    out = []
    for k, n in enumerate([12, 40, 100, 170]):
        spec = scenes.hanging_cloth(seed=20 + k, n=n, stretch=10 ** (-8 + k),
                                    bend=_bend_compliance(bend_model, 10 ** (-4 + k / 2)))
        out.append(spec)
    # one instance with LRA slots on its pinned row
    lra = scenes.hanging_cloth(seed=30, n=24, bend=_bend_compliance(bend_model, 1e-4))
    lra.lra_slots = np.array([0, 23], np.int32)
    out.append(lra)
    return out


@pytest.mark.parametrize("bend_model", [0, 1])
def test_batched_matches_plain_and_oracle_on_uneven_instances(bend_model):
    from oracle import oracle as orc
    specs = _mixed_specs(bend_model)
    params = make_params(bend_model=bend_model)
    sb, ib = build(specs, params, seed=3)
    sp, ip = build(specs, params, seed=3)
    info = sb.device_info()
    _print(f"device: {info}")
    oracles = [oracle_of(sb, i, sp_.stretch_compliance, sp_.bend_compliance)
               for i, sp_ in zip(ib, specs)]
    op, dt = orc.params_from(params)
    sb.SubstepBatched()
    for i in ip:
        sp.SubstepPlain(i)
    for k, (i, oc) in enumerate(zip(ib, oracles)):
        xb, _ = sb.GetState(i)
        xp, _ = sp.GetState(ip[k])
        oc.substep(op, dt)
        ox, _ = oc.state()
        eb, ep = rel_err(xb, ox), rel_err(xp, ox)
        _print(f"  inst {k} nv={len(specs[k].pos)} batched {eb:.3e} plain {ep:.3e} "
               f"|b-p| {np.abs(xb - xp).max():.3e}")
        assert eb <= TOL and ep <= TOL
        assert np.abs(xb - xp).max() <= 1e-6


def test_batched_graph_frames_match_substeps():
    specs = _mixed_specs()[:3]
    params = make_params(bend_model=1)
    sa, ia = build(specs, params, seed=1)
    sg, ig = build(specs, params, seed=1)
    for _ in range(2 * params.num_substeps):
        sa.SubstepBatched()
    sg.FramesBatched(1)
    sg.sync()
    sg.FramesBatched(1)
    for a, g in zip(ia, ig):
        np.testing.assert_array_equal(sa.GetState(a)[0], sg.GetState(g)[0])


def test_rollouts_are_bitwise_deterministic():
    specs = _mixed_specs()[:3]
    params = make_params(bend_model=1)
    runs = []
    for _ in range(2):
        s, ii = build(specs, params, seed=9)
        s.FramesBatched(2)
        s.FramesPlain(1)
        runs.append([s.GetState(i)[0].copy() for i in ii])
    for a, b in zip(*runs):
        np.testing.assert_array_equal(a, b)


def test_lra_and_per_edge_compliance_match_oracle():
    from oracle import oracle as orc
    spec = scenes.hanging_cloth(seed=41, n=40)
    spec.lra_slots = np.array([0, 39], np.int32)
    params = make_params(bend_model=1, long_range_stretchiness=1.02)
    s, (i,) = build([spec], params, seed=2)
    ne = s.counts(i).n_edges
    comp = (10.0 ** (-9 + 3 * scenes.Rng(8).uniform(ne))).astype(np.float32)
    s.SetStretchCompliance(i, comp)
    oc = oracle_of(s, i, comp.astype(np.float64), spec.bend_compliance)
    op, dt = orc.params_from(params)
    for _ in range(3):
        s.SubstepPlain(i)
        oc.substep(op, dt)
    err = rel_err(s.GetState(i)[0], oc.state()[0])
    _print(f"  LRA + per-edge compliance: rel err after 3 substeps {err:.3e}")
    assert err <= 3 * TOL


def test_ground_plane_contact_matches_oracle():
    from oracle import oracle as orc
    spec = scenes.hanging_cloth(seed=7, n=32)
    spec.x0[:, 1] -= 0.999  # just above the plane y = -1
    spec.pinned = np.zeros(0, np.int32)
    params = make_params(bend_model=1, ground_friction=0.3)
    s, (i,) = build([spec], params, seed=2)
    oc = oracle_of(s, i, spec.stretch_compliance, spec.bend_compliance)
    op, dt = orc.params_from(params)
    for _ in range(5):
        s.SubstepBatched()
        oc.substep(op, dt)
    x0, _ = s.GetState(i)
    assert x0[:, 1].min() >= -1.0 - 1e-6
    assert rel_err(x0, oc.state()[0]) <= TOL


def test_error_paths_on_device():
    spec = scenes.hanging_cloth(seed=1, n=8)
    s, (i,) = build([spec], make_params(), seed=0)
    with pytest.raises(XrtError):
        s.Pin(i, [0])                       # after start
    with pytest.raises(XrtError):
        s.SetState(i, np.full((64, 3), np.nan, np.float32))
    with pytest.raises(XrtError):
        s.StretchSolver(i).Solve(99, 1e-3, 0)
    with pytest.raises(XrtError):
        s.SubstepPlain(7)
    # reference-shaped calls in SimulateSwift order reproduce xrt_substep_plain
    s2, (j,) = build([spec], make_params(), seed=0)
    p = s.params
    dt = float(np.float32(p.frame_time) / np.float32(p.num_substeps))  # the library's FP32 dt
    s.PredictPositions(i, dt)
    for it in range(p.num_iterations):
        st = s.StretchSolver(i)
        for c in range(st.Coloring().n_colors):
            st.Solve(c, dt, it)
        bd = s.BendSolver(i)
        for c in range(bd.Coloring().n_colors):
            bd.Solve(c, dt, it)
        s.SolveLongRangeAttachments(i)
        s.SolveGround(i)
        s.ApplyDeltas(i)
    s.Finalize(i, dt)
    s2.SubstepPlain(j)
    np.testing.assert_array_equal(s.GetState(i)[0], s2.GetState(j)[0])


@pytest.mark.parametrize("path", ["plain", "batched"])
def test_reference_bend_model_at_configured_compliance(path):
    """The reference bend formula (bend_model 0) at the plain workload's configured bend
    compliance 1e-4: one substep of the 64 x 64 hanging cloth against the FP64 oracle.  The
    formula amplifies rounding there (a flat hinge measures pi/2 against a rest angle of 0,
    and its gradients come from the cross products of nearly opposite normals): the FP64
    oracle's own one-substep result moves by millimetres when its start state is perturbed at
    FP32 resolution (6e-8 relative).  The GPU is held to 1e-5 relative or to twice the
    largest such move over six perturbations."""
    from oracle import oracle as orc
    _print(f"case reference bend model at compliance 1e-4, {path}")
    spec = scenes.hanging_cloth(seed=1, n=64, bend=1e-4)
    params = make_params(bend_model=0)
    s, (i,) = build([spec], params, seed=5)
    op, dt = orc.params_from(params)
    oc = oracle_of(s, i, spec.stretch_compliance, spec.bend_compliance)
    x_start, v_start = s.GetState(i)
    s.SubstepPlain(i) if path == "plain" else s.SubstepBatched()
    x_gpu = s.GetState(i)[0]
    oc.substep(op, dt)
    x_ref = oc.state()[0].copy()
    err = rel_err(x_gpu, x_ref)
    dev = float(np.abs(x_gpu - x_ref).max())
    sens = 0.0
    for trial in range(1, 7):
        z = np.random.default_rng(trial).standard_normal(x_start.shape)
        xp = x_start.astype(np.float64) * (1.0 + 6e-8 * z)
        o2 = oracle_of(s, i, spec.stretch_compliance, spec.bend_compliance, x0=xp, v=v_start)
        o2.substep(op, dt)
        sens = max(sens, float(np.abs(o2.state()[0] - x_ref).max()))
    _print(f"  rel err {err:.3e}; deviation {dev:.3e} m, oracle self-sensitivity {sens:.3e} m")
    assert err <= TOL or dev <= 2.0 * sens
