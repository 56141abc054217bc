"""Pins of the FP64 reference gradient (oracle/adjoint.py, oracle/adjoint_oracle.c) by
central finite differences of whole oracle rollouts (CPU only).

Loss L = sum(a . x0_T) + sum(b . v_T) with fixed random a, b; the reverse pass gives
dL/d(stretch compliance), dL/d(bend compliance), dL/dv0, dL/dx0, and each is checked
along random directions (log-compliance directions as in the differentiable-rollout
workload) against (L(p + h d) - L(p - h d)) / 2h."""
import numpy as np
import pytest

from oracle import adjoint as oadj
from oracle import oracle as orc
from oracle import scene as osc
from xrtailor_amd import scenes
from xrtailor_amd.solver import make_params


def _grid_case(n=7, seed=3):
    pos, tris = scenes.grid(n, 0.3, 0.0)
    rng = scenes.Rng(seed)
    x0 = (pos + 1e-3 * rng.normal(pos.size).reshape(-1, 3)).astype(np.float32)
    v0 = 0.2 * rng.normal(pos.size).reshape(-1, 3)
    v0[:, 1] -= 6.0                      # reach the ground within the rollout
    pinned = np.array([0, n - 1], np.int32)
    return pos, tris, x0, v0, pinned, None


def _sphere_case(n=9, seed=5):
    pos, tris = scenes.grid(n, 0.5, 0.302)
    rng = scenes.Rng(seed)
    x0 = (pos + 1e-4 * rng.normal(pos.size).reshape(-1, 3)).astype(np.float32)
    v0 = np.zeros_like(x0, dtype=np.float64)
    v0[:, 1] = -1.0
    return pos, tris, x0, v0, np.zeros(0, np.int32), scenes.icosphere(3, 0.3)


def _lra_csr(pos, edges, slots):
    """Per-target LRA CSR of every slot (registration order) at its geodesic distance."""
    dist = [orc.geodesic(pos, edges, int(s)) for s in slots]
    ptr, slot, rest = [0], [], []
    for t in range(len(pos)):
        for k, d in enumerate(dist):
            if np.isfinite(d[t]):
                slot.append(k)
                rest.append(np.float32(d[t]))
        ptr.append(len(slot))
    return (np.asarray(slots, np.int32), np.asarray(ptr, np.int32), np.asarray(slot, np.int32),
            np.asarray(rest, np.float64))


class Rollout:
    def __init__(self, case, xp, seed=1, lra_slots=None):
        self.pos, self.tris, self.x0, self.v0, self.pinned, self.obs = case
        self.xp = xp
        self.op, _ = orc.params_from(xp)
        self.edges = orc.build_edges(self.tris)
        self.bends = orc.build_bends(self.tris, self.edges)
        self.rest = orc.rest_lengths(self.pos, self.edges).astype(np.float32).astype(np.float64)
        self.theta = orc.rest_angles(self.pos, self.bends).astype(np.float32).astype(np.float64)
        self.scol, self.sn = orc.color(self.edges, len(self.pos), seed)
        self.bcol, self.bn = orc.color(self.bends, len(self.pos), seed)
        rng = scenes.Rng(seed + 7)
        self.sc = 10.0 ** (-7 + 2 * rng.uniform(len(self.edges)))
        self.bc = 10.0 ** (-4 + 2 * rng.uniform(len(self.bends)))
        self.w = np.ones(len(self.pos))
        self.w[self.pinned] = 0.0
        self.lra = {}
        if lra_slots is not None:
            sv, ptr, sl, rs = _lra_csr(self.pos, self.edges, lra_slots)
            self.lra = dict(slot_vert=sv, lra_ptr=ptr, lra_slot=sl, lra_rest=rs)

    def scene(self, sc=None, bc=None, v0=None, x0=None):
        cl = orc.Cloth(self.x0 if x0 is None else x0, self.v0 if v0 is None else v0, self.w,
                       self.edges, self.rest, self.sc if sc is None else sc, self.scol, self.sn,
                       self.bends, self.theta, self.bc if bc is None else bc, self.bcol, self.bn,
                       **self.lra)
        obs = None if self.obs is None else (np.asarray(self.obs[0], np.float32), self.obs[1])
        return osc.ContactScene(cl, self.pos, self.tris, self.edges, self.op, self.xp, obstacle=obs)

    def loss(self, sc, n_sub, a, b, **kw):
        s = self.scene(sc=sc, **kw)
        for _ in range(n_sub):
            s.substep()
        x, v = s.c.state()
        return float((a * x).sum() + (b * v).sum())


def _check(r, n_sub, tol, h=1e-6, seed=11):
    """Step sizes: 1e-6 in log-compliance; 1e-7 in v0 and 1e-8 in x0, which reach the
    positions through 1/dt and near-rigid projections (large third derivatives: at 1e-6
    the truncation error of the central difference is ~10%)."""
    rng = scenes.Rng(seed)
    nv = len(r.pos)
    a = rng.normal(3 * nv).reshape(-1, 3)
    b = 0.01 * rng.normal(3 * nv).reshape(-1, 3)
    s = r.scene()
    tape = oadj.record(s, n_sub)
    g = oadj.backward(s, tape, a, b)
    out = {}
    L0 = abs(r.loss(r.sc, n_sub, a, b))
    # stretch compliance along a log direction
    d = rng.normal(len(r.sc))
    fd = (r.loss(r.sc * np.exp(h * d), n_sub, a, b) - r.loss(r.sc * np.exp(-h * d), n_sub, a, b)) / (2 * h)
    out["stretch"] = (fd, float((g["gs_edge"] * r.sc * d).sum()), h)
    d = rng.normal(len(r.bc))
    fd = (r.loss(r.sc, n_sub, a, b, bc=r.bc * np.exp(h * d)) -
          r.loss(r.sc, n_sub, a, b, bc=r.bc * np.exp(-h * d))) / (2 * h)
    out["bend"] = (fd, float((g["gb_elem"] * r.bc * d).sum()), h)
    d = rng.normal(3 * nv).reshape(-1, 3)
    hv = 1e-7
    fd = (r.loss(r.sc, n_sub, a, b, v0=r.v0 + hv * d) - r.loss(r.sc, n_sub, a, b, v0=r.v0 - hv * d)) / (2 * hv)
    out["v0"] = (fd, float((g["gv0"] * d).sum()), hv)
    d = rng.normal(3 * nv).reshape(-1, 3)
    x0 = r.x0.astype(np.float64)
    hx = 1e-8
    fd = (r.loss(r.sc, n_sub, a, b, x0=x0 + hx * d) - r.loss(r.sc, n_sub, a, b, x0=x0 - hx * d)) / (2 * hx)
    out["x0"] = (fd, float((g["gx0"] * d).sum()), hx)
    for k, (f, ad, hk) in out.items():
        roundoff = 50 * np.finfo(np.float64).eps * L0 / hk   # rounding of the difference
        print(f"  {k}: fd {f:.9e} adjoint {ad:.9e} (round-off {roundoff:.1e})", flush=True)
        assert abs(f - ad) <= tol * abs(f) + roundoff, k
    return g, tape


@pytest.mark.parametrize("bend_model", [0, 1])
def test_reverse_matches_fd_dynamics(bend_model):
    """Stretch (per-edge compliance, multiplier carry), bend, ground with friction."""
    xp = make_params(bend_model=bend_model, num_substeps=4, num_iterations=5, ground_height=-0.012,
                     ground_friction=0.3, enable_ground=1)
    r = Rollout(_grid_case(), xp)
    g, tape = _check(r, n_sub=4, tol=1e-6)
    assert any(e[0] == "ground" for e in tape)
    # the ground is reached: some vertex ends on the plane
    s = r.scene()
    for _ in range(4):
        s.substep()
    assert (s.c.state()[0][:, 1] <= -0.012 + 1e-9).any()


def test_reverse_matches_fd_with_long_range_attachments():
    """Geodesic LRA from the top row of a falling, stretching sheet (soft stretch, so the
    attachments bind): the reverse through the LRA passes against central differences.
    Stretchiness 1.05: at 1.0 an attachment to a mesh neighbour puts the vertex at exactly
    the edge's rest length, where the reference's stretch solve skips the edge (l == r)
    and drops its multiplier term -- a jump of the forward that no difference quotient
    resolves."""
    xp = make_params(bend_model=1, num_substeps=3, num_iterations=4, enable_ground=0,
                     long_range_stretchiness=1.05)
    n = 7
    r = Rollout(_grid_case(n), xp, lra_slots=np.arange(n * (n - 1), n * n))
    r.sc = r.sc * 1e3                      # soft stretch: the attachments take the load
    g, tape = _check(r, n_sub=3, tol=1e-6)
    assert any(e[0] == "lra" for e in tape)
    # the attachments act: the rollout differs from one without them
    s = r.scene()
    keep = r.lra
    r.lra = {}
    s2 = r.scene()
    r.lra = keep
    for _ in range(3):
        s.substep()
        s2.substep()
    assert np.abs(s.c.state()[0] - s2.c.state()[0]).max() > 1e-6


def _fd_jacobian_check(fwd, x, gbar_out, rev, h=1e-7, tol=1e-6):
    """Central FD of <gbar_out, fwd(x)> per coordinate against rev(x, gbar_out)."""
    ad = rev(x, gbar_out)
    fd = np.zeros_like(x)
    for k in range(x.size):
        e = np.zeros_like(x)
        e.flat[k] = h
        fd.flat[k] = ((gbar_out * fwd(x + e)).sum() - (gbar_out * fwd(x - e)).sum()) / (2 * h)
    np.testing.assert_allclose(ad, fd, rtol=tol, atol=tol * np.abs(fd).max())


def test_vf_reverse_matches_fd():
    """VF-DCD / VF-CCD solve with its record (face, q) fixed: x' = x - min(n.(x-q) - off, 0) n.
    Replaying the record is exact here: q lies in the face plane, so n.q does not move."""
    lib = oadj._lib()
    rng = scenes.Rng(21)
    ot = np.array([0, 1, 2, 0, 2, 3], np.int32)
    pred = np.array([[-1, 0.003, -1], [1, 0.003, -1], [1, 0.004, 1], [-1, 0.002, 1]], np.float32)
    nv = 12
    face = (np.arange(nv) % 3 - 1).astype(np.int32)          # -1 = no record
    x = rng.normal(3 * nv).reshape(-1, 3) * [0.3, 0.002, 0.3]
    q = x.copy()
    q[:, 1] = 0.003
    for off in (0.0, 1e-3):
        def fwd(xx):
            out = np.ascontiguousarray(xx.reshape(-1)).copy()
            lib.orc_vf_solve(pred.ctypes.data_as(oadj._F), ot.ctypes.data_as(oadj._I),
                             face.ctypes.data_as(oadj._I), oadj._d(np.ascontiguousarray(q.reshape(-1))),
                             oadj._d(out), nv, off, None, None, None, 0.1, 0.0)  # no friction
            return out.reshape(-1, 3)

        def rev(xx, gb):
            g = np.ascontiguousarray(gb.reshape(-1)).copy()
            lib.orc_adj_vf(pred.ctypes.data_as(oadj._F), ot.ctypes.data_as(oadj._I),
                           face.ctypes.data_as(oadj._I), oadj._d(np.ascontiguousarray(q.reshape(-1))),
                           oadj._d(np.ascontiguousarray(xx.reshape(-1))), nv, off, oadj._d(g))
            return g.reshape(-1, 3)
        gb = rng.normal(3 * nv).reshape(-1, 3)
        _fd_jacobian_check(fwd, x, gb, rev)


def test_ee_reverse_matches_fd():
    """EE-DCD solve + ApplyDeltas(relax) with the records (s, t, fn, obstacle edge) fixed,
    at states whose constraints are away from their switch (lambda = 0) so that the
    Jacobi counts are constant under the perturbation."""
    lib = oadj._lib()
    rng = scenes.Rng(33)
    # a strip of 6 cloth vertices / 5 edges just above the plane y = 0.003
    nv, ne = 6, 5
    edges = np.array([[k, k + 1] for k in range(ne)], np.int32)
    x = np.zeros((nv, 3))
    x[:, 0] = 0.1 * np.arange(nv)
    x[:, 1] = 0.003 + np.array([-2e-3, 5e-4, -5e-4, -1e-3, 2e-4, 4e-3])
    rec = np.zeros((ne, 12))
    rec[:, 0] = 0.5                                             # distance (a flag)
    rec[:, 1:4] = [[0.05 + 0.1 * k, 0.003, -0.2] for k in range(ne)]
    rec[:, 4:7] = [[0.05 + 0.1 * k, 0.003, 0.2] for k in range(ne)]
    rec[:, 7] = 0.25 + 0.1 * rng.uniform(ne)                    # s
    rec[:, 8] = rng.uniform(ne)                                 # t
    rec[:, 10] = 1.0                                            # fn = +y
    rec[3, 0] = 1e30                                            # one edge without a record
    e = np.ascontiguousarray(edges.reshape(-1))
    relax = 0.25

    def fwd(xx):
        out = np.ascontiguousarray(xx.reshape(-1)).copy()
        delta = np.zeros(3 * nv)
        cnt = np.zeros(nv, np.int32)
        lib.orc_ee_solve(oadj._d(rec.reshape(-1)), oadj._d(out), e.ctypes.data_as(oadj._I), ne,
                         oadj._d(delta), cnt.ctypes.data_as(oadj._I))
        lib.orc_apply_deltas(oadj._d(out), oadj._d(delta), cnt.ctypes.data_as(oadj._I), nv, relax)
        return out.reshape(-1, 3)

    def rev(xx, gb):
        g = np.ascontiguousarray(gb.reshape(-1)).copy()
        lib.orc_adj_ee(oadj._d(rec.reshape(-1)), oadj._d(np.ascontiguousarray(xx.reshape(-1))),
                       e.ctypes.data_as(oadj._I), ne, nv, relax, oadj._d(g))
        return g.reshape(-1, 3)
    moved = np.abs(fwd(x) - x).max(axis=1) > 0
    assert moved.sum() >= 3                                     # several active constraints
    _fd_jacobian_check(fwd, x, rng.normal(3 * nv).reshape(-1, 3), rev)


def test_body_contact_rollout_reverse_runs():
    """The tape of a sphere drop carries VF and EE records and reverses to finite
    gradients.  (Whole-rollout finite differences are not defined here: VF-DCD leaves
    vertices exactly at the 1e-3 offset that EE-DCD tests against, so the EE Jacobi count
    switches on rounding -- a discontinuity of the forward itself.)"""
    xp = make_params(bend_model=1, num_substeps=4, num_iterations=5, enable_ground=0,
                     enable_obstacle_contact=1, num_collision_passes=10)
    r = Rollout(_sphere_case(), xp)
    s = r.scene()
    tape = oadj.record(s, 4)
    nv = len(r.pos)
    g = oadj.backward(s, tape, np.ones((nv, 3)), np.zeros((nv, 3)))
    assert sum(int((e[2] >= 0).sum()) for e in tape if e[0] == "vf") > 0
    assert sum(int((e[1][0::12] < 1e6).sum()) for e in tape if e[0] == "ee") > 0
    for k in ("gx0", "gv0", "gs", "gb"):
        assert np.all(np.isfinite(g[k])), k


def _two_layers_case(n=8, seed=5):
    """Two sheets 1.5 mm apart (inside the 2 mm repulsion thickness), the upper one turned
    by 0.45 rad and wrinkled, the lower one pinned at two corners, the upper one moving
    down: vertex-face and edge-edge repulsion responses in every substep."""
    from contact_helpers import two_layers

    spec = two_layers(n=n, seed=seed)
    nv = len(spec.pos)
    v0 = np.zeros((nv, 3))
    v0[nv // 2:, 1] = -0.05
    v0 += 0.01 * scenes.Rng(seed + 1).normal(3 * nv).reshape(-1, 3)
    return spec.pos, spec.tris, spec.x0, v0, np.array([0, n - 1], np.int32), None


def test_repulsion_reverse_matches_fd():
    """PBD repulsion (vertex-face and edge-edge proximity responses, self contact, Jacobi
    apply): the reverse replays the pass's recorded responses and counts; gradients w.r.t.
    compliances, v0 and x0 against central differences of whole rollouts."""
    xp = make_params(bend_model=1, num_substeps=3, num_iterations=4, enable_ground=0,
                     enable_self_contact=1, pair_update_interval=2)
    r = Rollout(_two_layers_case(), xp)
    g, tape = _check(r, n_sub=3, tol=1e-5)
    reps = [e for e in tape if e[0] == "repulsion"]
    kinds = {it.kind for e in reps for it in e[1]}
    print(f"  {sum(len(e[1]) for e in reps)} recorded responses over {len(reps)} passes, "
          f"kinds {sorted(kinds)}", flush=True)
    assert kinds == {0, 1}   # both vertex-face and edge-edge responses were reversed


def test_repulsion_reverse_planted_mistake_fails():
    """The reverse pin is sensitive: with the responses' contributions scaled by 1.01 (a
    wrong relaxation or count), the x0 and v0 gradients leave the 1e-5 bar."""
    xp = make_params(bend_model=1, num_substeps=3, num_iterations=4, enable_ground=0,
                     enable_self_contact=1, pair_update_interval=2)
    r = Rollout(_two_layers_case(), xp)
    s = r.scene()
    tape = oadj.record(s, 3)
    bent = [(e[0], e[1], e[2], e[3], e[4], e[5], e[6] * 1.01) if e[0] == "repulsion" else e
            for e in tape]
    nv = len(r.pos)
    rng = scenes.Rng(11)
    a = rng.normal(3 * nv).reshape(-1, 3)
    b = 0.01 * rng.normal(3 * nv).reshape(-1, 3)
    g_ok = oadj.backward(s, tape, a, b)
    g_bad = oadj.backward(r.scene(), bent, a, b)
    rel = float(np.abs(g_bad["gv0"] - g_ok["gv0"]).max() / np.abs(g_ok["gv0"]).max())
    print(f"  relaxation x 1.01 in the reverse moves gv0 by {rel:.2e}", flush=True)
    assert rel > 1e-4


def test_fp32_forward_moves_the_gradient_beyond_fp64_spread():
    """Why the full-size GPU gradients are held to their own FP32 noise: the FP64 reverse
    along an FP32-arithmetic forward (predict, stiff stretch colors and finalize in float32,
    tests/fp32_forward.py) of a 48 x 48 cloth released above a sphere (compliance 1e-8
    log-normal, 10 x 10) departs from the FP64 reverse along the FP64 forward by more than
    1e-3, while the FP64 reverse itself moves by less than 5e-4 under a 1e-7 start
    perturbation; the departure stays within 4x the FP32 forward's own move under that
    perturbation (the bar of the GPU's full-size gradient tests)."""
    from fp32_forward import use_fp32_forward
    from oracle import scene as osc

    xp = make_params(bend_model=1, enable_obstacle_contact=1, ground_height=-1.0,
                     solve_bending=0, enable_ground=0)
    spec = scenes.differentiable_drop(seed=5, n=48)
    op, _ = orc.params_from(xp)
    edges = orc.build_edges(spec.tris)
    bends = orc.build_bends(spec.tris, edges)
    rest = orc.rest_lengths(spec.pos, edges).astype(np.float32).astype(np.float64)
    theta = orc.rest_angles(spec.pos, bends).astype(np.float32).astype(np.float64)
    scol, sn = orc.color(edges, len(spec.pos), 5)
    bcol, bn = orc.color(bends, len(spec.pos), 5)
    nv = len(spec.pos)
    sc = spec.stretch_compliance * np.exp(0.5 * scenes.Rng(100).normal(len(edges)))
    x0 = (spec.pos + 1e-3 * scenes.Rng(200).normal(3 * nv).reshape(-1, 3)).astype(np.float32)
    r = scenes.Rng(300)
    a = r.normal(3 * nv).reshape(-1, 3)
    b = (0.01 * r.normal(3 * nv)).reshape(-1, 3)
    obs = (np.asarray(spec.obstacle[0], np.float32), spec.obstacle[1])

    def grad(start, fp32):
        cl = orc.Cloth(start.astype(np.float64), np.zeros((nv, 3)), np.ones(nv), edges, rest, sc,
                       scol, sn, bends, theta, np.full(len(bends), 1e-4), bcol, bn)
        s = osc.ContactScene(cl, spec.pos, spec.tris, edges, op, xp, obstacle=obs)
        if fp32:
            use_fp32_forward(s)
        tape = oadj.record(s, xp.num_substeps)
        return oadj.backward(s, tape, a, b)

    moved = (x0 * (1.0 + 1e-7 * np.random.default_rng(1).standard_normal(x0.shape))).astype(np.float32)
    g64, g64m = grad(x0, False), grad(moved, False)
    g32, g32m = grad(x0, True), grad(moved, True)
    rel = lambda u, v: float(np.abs(u - v).max() / np.abs(v).max())
    for k in ("gs_edge", "gx0", "gv0"):
        dep, s64, s32 = rel(g32[k], g64[k]), rel(g64m[k], g64[k]), rel(g32m[k], g32[k])
        print(f"{k}: FP32 forward vs FP64 {dep:.2e}; FP64 spread {s64:.2e}; FP32 spread {s32:.2e}")
        assert s64 < 5e-4
        assert dep > 1e-3
        assert dep <= 4.0 * s32
