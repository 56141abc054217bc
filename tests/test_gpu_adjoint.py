"""GPU adjoint (reverse-mode) kernels against the FP64 reference gradient and central
finite differences, through the C ABI.

Bar: gradients within 1e-3 relative (||g - g_ref||_inf / ||g_ref||_inf) of the FP64
oracle's reverse pass on the same FP32-rounded start, and within 1e-3 of central
finite differences (eps 1e-3 in log compliance) of the GPU's own forward."""
import numpy as np
import pytest

from gpu_helpers import oracle_of
from oracle import adjoint as oadj
from oracle import oracle as orc
from oracle import scene as osc
from xrtailor_amd import XrtError, scenes
from xrtailor_amd.solver import ClothSolver, make_params

pytestmark = pytest.mark.gpu
TOL = 1e-3


def _rel(g, ref):
    g, ref = np.asarray(g, np.float64), np.asarray(ref, np.float64)
    return float(np.abs(g - ref).max() / max(np.abs(ref).max(), 1e-30))


class Case:
    """One cloth instance on the GPU and its oracle twin (same topology, colors, state)."""

    def __init__(self, xp, n=12, side=0.4, height=0.0, vy=-3.0, pinned=True, obstacle=None,
                 bend=1e-3, seed=3, lra_slots=None, pins=None):
        self.xp = xp
        self.pos, self.tris = scenes.grid(n, side, height)
        rng = scenes.Rng(seed)
        self.x0 = (self.pos + 1e-3 * rng.normal(self.pos.size).reshape(-1, 3)).astype(np.float32)
        v0 = 0.2 * rng.normal(self.pos.size).reshape(-1, 3)
        v0[:, 1] += vy
        self.v0 = v0.astype(np.float32)
        self.obstacle = obstacle
        self.bend = bend
        s = ClothSolver(params=xp)
        self.i = s.AddCloth(self.pos, self.tris, 1e-6, bend)
        if pins is not None:
            s.Pin(self.i, np.asarray(pins, np.int32))
        elif pinned:
            s.Pin(self.i, np.array([0, n - 1], np.int32))
        if obstacle is not None:
            s.SetObstacle(self.i, *obstacle)
        if lra_slots is not None:
            s.AddLraSlots(self.i, np.asarray(lra_slots, np.int32))
        s.Start(seed)
        self.s = s
        self.ne = s.counts(self.i).n_edges
        self.sc = (10.0 ** (-7 + 2 * rng.uniform(self.ne))).astype(np.float32)
        s.SetStretchCompliance(self.i, self.sc)
        self.a = rng.normal(self.pos.size).reshape(-1, 3).astype(np.float32)
        self.b = (0.01 * rng.normal(self.pos.size)).reshape(-1, 3).astype(np.float32)

    def gpu_grad(self, n_frames, ckpt=1, sc=None):
        s, i = self.s, self.i
        s.SetStretchCompliance(i, self.sc if sc is None else sc)
        s.SetState(i, self.x0, self.v0)
        s.AdjointBegin(n_frames, ckpt)
        s.AdjointSeed(i, self.a, self.b)
        s.AdjointBackward()
        gs, gb = s.GradCompliance(i)
        gx0, gv0 = s.GradState(i)
        return dict(gs=gs, gb=gb, gx0=gx0, gv0=gv0)

    def oracle_grad(self, n_frames):
        cl = oracle_of(self.s, self.i, self.sc.astype(np.float64), self.bend, x0=self.x0, v=self.v0)
        edges = self.s.edges(self.i, colors=False)[0]
        op, _ = orc.params_from(self.xp)
        obs = None if self.obstacle is None else (np.asarray(self.obstacle[0], np.float32),
                                                  self.obstacle[1])
        sc = osc.ContactScene(cl, self.pos, self.tris, edges, op, self.xp, obstacle=obs)
        tape = oadj.record(sc, n_frames * self.xp.num_substeps)
        g = oadj.backward(sc, tape, self.a, self.b)
        return dict(gs=g["gs_edge"], gb=g["gb_elem"], gx0=g["gx0"], gv0=g["gv0"]), tape


def _compare(g, ref, tol=TOL):
    worst = 0.0
    for k in ("gs", "gb", "gx0", "gv0"):
        e = _rel(g[k], ref[k])
        print(f"  {k}: rel err {e:.3e} (|ref| max {np.abs(ref[k]).max():.3e})", flush=True)
        worst = max(worst, e)
    assert worst <= tol


@pytest.mark.parametrize("bend_model", [0, 1])
def test_gradient_matches_oracle(bend_model):
    print(f"case adjoint grid 12x12 bend_model={bend_model}, ground + friction, 2 frames", flush=True)
    xp = make_params(bend_model=bend_model, ground_height=-0.04, ground_friction=0.3,
                     num_substeps=5, num_iterations=5)
    c = Case(xp, bend=1e-3 if bend_model == 1 else 1.0)
    g = c.gpu_grad(2)
    ref, tape = c.oracle_grad(2)
    assert any(e[0] == "ground" for e in tape)
    _compare(g, ref)


def test_checkpoint_schedules_agree_bit_for_bit():
    """Recompute from checkpoints reproduces the recorded trajectory exactly, so every
    checkpoint interval gives the same gradient."""
    print("case adjoint checkpoint intervals 1 / 2 / 4 over 4 frames", flush=True)
    xp = make_params(bend_model=1, ground_height=-0.04, ground_friction=0.3,
                     num_substeps=4, num_iterations=4)
    c = Case(xp)
    g1 = c.gpu_grad(4, ckpt=1)
    for k in (2, 4):
        gk = c.gpu_grad(4, ckpt=k)
        for key in g1:
            np.testing.assert_array_equal(g1[key], gk[key])


def _oracle_final(c, sc, n_frames, v0=None):
    """FP64 oracle rollout of the case's instance with per-edge compliance sc."""
    cl = oracle_of(c.s, c.i, np.asarray(sc, np.float64), c.bend, x0=c.x0,
                   v=c.v0 if v0 is None else v0)
    edges = c.s.edges(c.i, colors=False)[0]
    op, _ = orc.params_from(c.xp)
    scene = osc.ContactScene(cl, c.pos, c.tris, edges, op, c.xp)
    for _ in range(n_frames * c.xp.num_substeps):
        scene.substep()
    return scene.c.state()


def _smooth_case():
    """Swinging cloth (pinned corners) that stays clear of the ground: no contact or
    friction switches, so central differences at eps 1e-3 are meaningful."""
    xp = make_params(bend_model=1, ground_height=-1.0, num_substeps=5, num_iterations=5)
    return Case(xp)


def test_gradient_matches_finite_differences():
    """Central differences (eps 1e-3 in log compliance and in v0) of the FP64 forward
    against the GPU adjoint, on random directions."""
    print("case adjoint vs FD (eps 1e-3), swinging 12x12, 2 frames", flush=True)
    c = _smooth_case()
    g = c.gpu_grad(2)
    rng = scenes.Rng(99)
    a, b = c.a.astype(np.float64), c.b.astype(np.float64)

    def loss(sc, v0=None):
        x, v = _oracle_final(c, sc, 2, v0)
        return float((a * x).sum() + (b * v).sum())
    eps = 1e-3
    sc64 = c.sc.astype(np.float64)
    for trial in range(2):
        d = rng.normal(c.ne)
        fd = (loss(sc64 * np.exp(eps * d)) - loss(sc64 * np.exp(-eps * d))) / (2 * eps)
        ad = float((g["gs"].astype(np.float64) * sc64 * d).sum())
        print(f"  log-compliance trial {trial}: fd {fd:.6e} adjoint {ad:.6e}", flush=True)
        assert abs(fd - ad) <= TOL * max(abs(fd), abs(ad))
    d = rng.normal(c.v0.size).reshape(-1, 3)
    v64 = c.v0.astype(np.float64)
    fd = (loss(sc64, v64 + eps * d) - loss(sc64, v64 - eps * d)) / (2 * eps)
    ad = float((g["gv0"].astype(np.float64) * d).sum())
    print(f"  v0: fd {fd:.6e} adjoint {ad:.6e}", flush=True)
    assert abs(fd - ad) <= TOL * max(abs(fd), abs(ad))
    # (differences of the GPU's own FP32 forward are not used: its rollout noise, ~1e-6 of
    # the loss, is of the order of the eps = 1e-3 difference itself)


def test_body_contact_gradient_matches_oracle():
    """Sphere drop with VF-DCD / VF-CCD / EE-DCD records replayed on both sides."""
    print("case adjoint sphere drop 9x9 over icosphere s3", flush=True)
    xp = make_params(bend_model=1, enable_ground=0, enable_obstacle_contact=1,
                     num_substeps=5, num_iterations=5, num_collision_passes=5)
    c = Case(xp, n=9, side=0.5, height=0.302, vy=-1.0, pinned=False,
             obstacle=scenes.icosphere(3, 0.3))
    g = c.gpu_grad(1)
    ref, tape = c.oracle_grad(1)
    n_vf = sum(int((e[2] >= 0).sum()) for e in tape if e[0] == "vf")
    print(f"  VF records over the frame: {n_vf}", flush=True)
    assert n_vf > 0
    _compare(g, ref)


def test_adjoint_substep_matches_oracle():
    print("case xrt_adjoint_substep vs one oracle substep", flush=True)
    xp = make_params(bend_model=1, ground_height=-0.004, ground_friction=0.3,
                     num_substeps=5, num_iterations=5)
    c = Case(xp)
    s, i = c.s, c.i
    s.SetStretchCompliance(i, c.sc)
    s.SetState(i, c.x0, c.v0)
    s.AdjointBegin(1, 1)           # allocates; the state is now one frame later
    s.SetState(i, c.x0, c.v0)
    s.AdjointSeed(i, c.a, c.b)
    s.AdjointSubstep()
    gs, gb = s.GradCompliance(i)
    gx0, gv0 = s.GradState(i)
    cl = oracle_of(s, i, c.sc.astype(np.float64), c.bend, x0=c.x0, v=c.v0)
    edges = s.edges(i, colors=False)[0]
    op, _ = orc.params_from(xp)
    sc = osc.ContactScene(cl, c.pos, c.tris, edges, op, xp)
    tape = oadj.record(sc, 1)
    ref = oadj.backward(sc, tape, c.a, c.b)
    _compare(dict(gs=gs, gb=gb, gx0=gx0, gv0=gv0),
             dict(gs=ref["gs_edge"], gb=ref["gb_elem"], gx0=ref["gx0"], gv0=ref["gv0"]))


def test_torch_function_rollout_loss():
    """Differentiable-rollout loss in torch (mean squared distance to a target made with
    log-normally perturbed compliances, sigma 0.5); the gradient that flows back through
    the autograd Function is checked by central differences (eps 1e-3) of the same loss
    on the FP64 forward."""
    import torch
    from torch_rollout import Rollout
    print("case torch.autograd.Function rollout 12x12, 2 frames", flush=True)
    c = _smooth_case()
    rng = scenes.Rng(7)
    log_c = np.log(c.sc.astype(np.float64))
    target_c = log_c + 0.5 * rng.normal(c.ne)
    x_t = Rollout.apply(torch.tensor(target_c), torch.tensor(c.v0.astype(np.float64)), c.s, c.i,
                        c.x0, 2, 1).detach()

    lc = torch.tensor(log_c, requires_grad=True)
    v0 = torch.tensor(c.v0.astype(np.float64), requires_grad=True)
    x = Rollout.apply(lc, v0, c.s, c.i, c.x0, 2, 1)
    L = ((x - x_t) ** 2).sum(dim=1).mean()
    L.backward()
    assert torch.isfinite(lc.grad).all() and torch.isfinite(v0.grad).all()
    xt = x_t.numpy()

    def loss64(lcn):
        xo, _ = _oracle_final(c, np.exp(lcn), 2)
        return float(((xo - xt) ** 2).sum(axis=1).mean())
    d = rng.normal(c.ne)
    eps = 1e-3
    fd = (loss64(log_c + eps * d) - loss64(log_c - eps * d)) / (2 * eps)
    ad = float((lc.grad.numpy() * d).sum())
    print(f"  loss {float(L):.6e} fd {fd:.6e} adjoint {ad:.6e}", flush=True)
    assert abs(fd - ad) <= TOL * max(abs(fd), abs(ad))


def test_unsupported_scenes_are_refused():
    xp = make_params(bend_model=1, enable_particle_collision=1)
    c = Case(xp, n=6)
    with pytest.raises(XrtError):
        c.s.AdjointBegin(1, 1)
    xp = make_params(bend_model=1, enable_obstacle_contact=1, contact_friction=0.3)
    c = Case(xp, n=6, obstacle=scenes.icosphere(2, 0.1, (0.0, -1.0, 0.0)))
    with pytest.raises(XrtError):
        c.s.AdjointBegin(1, 1)
    with pytest.raises(XrtError):
        c.s.AdjointBackward()


@pytest.mark.parametrize("frames,ckpt", [(1, 1), (2, 2)])
def test_lra_gradient_matches_oracle(frames, ckpt):
    """Long-range attachments from the top row of a sheet hanging from every third top
    vertex and pulled down (soft stretch, so the attachments bind; stretchiness 1.05): the
    reverse replays each target's slot sequence from the taped pre-LRA state and sums
    every slot's adjoint in constraint order."""
    print(f"case adjoint with geodesic LRA, 12x12, {frames} frame(s)", flush=True)
    xp = make_params(bend_model=1, ground_height=-1.0, num_substeps=4, num_iterations=4,
                     long_range_stretchiness=1.05)
    n = 12
    top = np.arange(n * (n - 1), n * n)
    c = Case(xp, n=n, lra_slots=top, pins=top[::3], vy=-1.0)
    c.sc = (c.sc * 1e2).astype(np.float32)
    g = c.gpu_grad(frames, ckpt=ckpt)
    ref, tape = c.oracle_grad(frames)
    assert any(e[0] == "lra" for e in tape)
    _compare(g, ref)


def test_batch_of_rollouts_matches_single_rollouts():
    """Several rollouts in one context (wide launches over every instance) give each
    instance the gradient it gets alone."""
    print("case adjoint batch of 3 rollouts (uneven sizes, sphere contact)", flush=True)
    xp = make_params(bend_model=1, enable_ground=0, enable_obstacle_contact=1,
                     num_substeps=4, num_iterations=5, num_collision_passes=4)
    sides = (7, 10, 8)
    specs, seeds = [], []
    for k, n in enumerate(sides):
        pos, tris = scenes.grid(n, 0.5, 0.302)
        rng = scenes.Rng(40 + k)
        x0 = (pos + 1e-4 * rng.normal(pos.size).reshape(-1, 3)).astype(np.float32)
        specs.append((pos, tris, x0, scenes.icosphere(3, 0.3)))

    def run(which):
        s = ClothSolver(params=xp)
        ids = [s.AddCloth(specs[k][0], specs[k][1], 1e-6, 1e-3) for k in which]
        for i, k in zip(ids, which):
            s.SetObstacle(i, *specs[k][3])
        s.Start(5)
        out = {}
        for i, k in zip(ids, which):
            pos, tris, x0, _ = specs[k]
            v = np.zeros_like(x0)
            v[:, 1] = -1.0
            s.SetState(i, x0, v)
            ne = s.counts(i).n_edges
            s.SetStretchCompliance(i, (10.0 ** (-7 + 2 * scenes.Rng(k).uniform(ne))).astype(np.float32))
        s.AdjointBegin(2, 1)
        for i, k in zip(ids, which):
            nv = len(specs[k][0])
            s.AdjointSeed(i, scenes.Rng(70 + k).normal(3 * nv).reshape(-1, 3).astype(np.float32), None)
        s.AdjointBackward()
        for i, k in zip(ids, which):
            out[k] = s.GradCompliance(i) + s.GradState(i)
        return out

    together = run([0, 1, 2])
    for k in range(3):
        alone = run([k])[k]
        for a, b in zip(together[k], alone):
            np.testing.assert_array_equal(a, b)


class TwoLayers:
    """Two 8 x 8 sheets 1.5 mm apart (inside the 2 mm repulsion thickness), the upper one
    turned and wrinkled and moving down, the lower one pinned at two corners: vertex-face
    and edge-edge repulsion responses (self contact) in every substep."""

    def __init__(self, xp, n=8, seed=5):
        from contact_helpers import two_layers
        self.xp = xp
        self.spec = two_layers(n=n, seed=seed)
        nv = len(self.spec.pos)
        v0 = np.zeros((nv, 3))
        v0[nv // 2:, 1] = -0.05
        v0 += 0.01 * scenes.Rng(seed + 1).normal(3 * nv).reshape(-1, 3)
        self.v0 = v0.astype(np.float32)
        s = ClothSolver(params=xp)
        self.i = s.AddCloth(self.spec.pos, self.spec.tris, 1e-7, 1e-3)
        s.Pin(self.i, np.array([0, n - 1], np.int32))
        s.Start(seed)
        self.s = s
        self.ne = s.counts(self.i).n_edges
        rng = scenes.Rng(seed + 2)
        self.sc = (10.0 ** (-7 + 2 * rng.uniform(self.ne))).astype(np.float32)
        self.a = rng.normal(3 * nv).reshape(-1, 3).astype(np.float32)
        self.b = (0.01 * rng.normal(3 * nv)).reshape(-1, 3).astype(np.float32)

    def gpu_grad(self, frames, ckpt, x0=None):
        s, i = self.s, self.i
        s.SetStretchCompliance(i, self.sc)
        s.SetState(i, self.spec.x0 if x0 is None else x0, self.v0)
        s.AdjointBegin(frames, ckpt)
        s.AdjointSeed(i, self.a, self.b)
        s.AdjointBackward()
        gx0, gv0 = s.GradState(i)
        return dict(gs=s.GradCompliance(i)[0], gx0=gx0, gv0=gv0)

    def oracle_grad(self, frames):
        cl = oracle_of(self.s, self.i, self.sc.astype(np.float64), 1e-3, x0=self.spec.x0,
                       v=self.v0)
        edges = self.s.edges(self.i, colors=False)[0]
        op, _ = orc.params_from(self.xp)
        sc = osc.ContactScene(cl, self.spec.pos, self.spec.tris, edges, op, self.xp)
        tape = oadj.record(sc, frames * self.xp.num_substeps)
        g = oadj.backward(sc, tape, self.a, self.b)
        n_resp = sum(len(e[1]) for e in tape if e[0] == "repulsion")
        kinds = {it.kind for e in tape if e[0] == "repulsion" for it in e[1]}
        return dict(gs=g["gs_edge"], gx0=g["gx0"], gv0=g["gv0"]), n_resp, kinds


def test_self_contact_reverse_matches_fp64_reverse():
    """PBD repulsion reversed on the GPU from the taped forward's recorded responses (their
    Jacobians by dual numbers, Jacobi counts replayed, sums in fixed point): a two-layer
    cloth's gradients against the FP64 reverse of the oracle twin, within max(1e-3, 4 x the
    GPU gradient's own move under a 1e-7 start perturbation)."""
    print("case adjoint with self contact: two layers 2 x 8 x 8, 3 x 4", flush=True)
    xp = make_params(bend_model=1, num_substeps=3, num_iterations=4, enable_ground=0,
                     enable_self_contact=1, pair_update_interval=2)
    c = TwoLayers(xp)
    g = c.gpu_grad(1, 1)
    ref, n_resp, kinds = c.oracle_grad(1)
    print(f"  {n_resp} repulsion responses in the oracle's frame, kinds {sorted(kinds)}",
          flush=True)
    assert n_resp > 0 and kinds == {0, 1}
    spread = {k: 0.0 for k in g}
    for t in (1, 2):
        rng = np.random.default_rng(t)
        x0 = (c.spec.x0 * (1.0 + 1e-7 * rng.standard_normal(c.spec.x0.shape))).astype(np.float32)
        gm = c.gpu_grad(1, 1, x0=x0)
        for k in spread:
            spread[k] = max(spread[k], _rel(gm[k], g[k]))
    for k in ("gs", "gx0", "gv0"):
        e = _rel(g[k], ref[k])
        bar = max(TOL, 4.0 * spread[k])
        print(f"  {k}: rel err {e:.3e}; GPU FP32 spread {spread[k]:.3e}; bar {bar:.3e}",
              flush=True)
        assert e <= bar


def test_self_contact_reverse_is_reproducible_and_schedule_independent():
    """The repulsion reverse sums in fixed point, so two backward passes and two checkpoint
    schedules (one 2-frame segment, two 1-frame segments) give the same bits."""
    print("case adjoint with self contact: bitwise reproducibility", flush=True)
    xp = make_params(bend_model=1, num_substeps=3, num_iterations=4, enable_ground=0,
                     enable_self_contact=1, pair_update_interval=2)
    c = TwoLayers(xp)
    g1 = c.gpu_grad(2, 2)
    g2 = c.gpu_grad(2, 2)
    g3 = c.gpu_grad(2, 1)
    for k in g1:
        assert np.isfinite(g1[k]).all() and np.abs(g1[k]).max() > 0
        np.testing.assert_array_equal(g1[k], g2[k])
        np.testing.assert_array_equal(g1[k], g3[k])
