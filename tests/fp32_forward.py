"""An FP32-arithmetic forward for the oracle's contact scene: predict, the XPBD stretch
colors and finalize evaluated in float32 the way the GPU kernels evaluate them (same
operation order, no contraction), the state stored in float32, recording the same tape as
ContactScene.substep so that the FP64 reverse (oracle.adjoint.backward) runs along an FP32
trajectory.  Bending, LRA, ground and particle / self contact are not emulated (the
substep refuses scenes that use them); obstacle contact runs its record generation and
must find no contact (asserted), so the obstacle only has to be present."""
import ctypes as C
import types

import numpy as np

from oracle import oracle as orc
from oracle import scene as osc

F32 = np.float32
KEPS = F32(1e-6)


def _stretch_color(c, k, dt, it):
    lo, hi = int(c.s_off[k]), int(c.s_off[k + 1])
    idx = c.s_idx.reshape(-1, 2)[lo:hi]
    X = c.x.reshape(-1, 3)
    a, b = X[idx[:, 0]].astype(F32), X[idx[:, 1]].astype(F32)
    lam = np.zeros(hi - lo, F32) if it == 0 else c.s_lam[lo:hi].astype(F32)
    rest, comp = c.s_rest[lo:hi].astype(F32), c.s_comp[lo:hi].astype(F32)
    inv_dt2 = F32(F32(1.0) / (F32(dt) * F32(dt)))
    d = a - b
    dist = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
    w1, w2 = c.w[idx[:, 0]].astype(F32), c.w[idx[:, 1]].astype(F32)
    act = (dist != rest) & (w1 + w2 > 0)
    inv = F32(1.0) / (dist + KEPS)
    at = comp * inv_dt2
    dl = np.where(act, (-(dist - rest) - at * lam) / (w1 + w2 + at), F32(0.0))
    s = (dl * inv)[:, None]
    X[idx[:, 0]] = np.where(act[:, None], a + w1[:, None] * s * d, a)
    X[idx[:, 1]] = np.where(act[:, None], b - w2[:, None] * s * d, b)
    c.s_lam[lo:hi] = np.where(act, lam + dl, lam)


def _substep(self):
    lib, c, p, xp, dt = self.lib, self.c, self.p, self.xp, self.dt
    assert not p.solve_bending and not p.enable_ground and not xp.enable_self_contact
    assert len(c.slot_vert) == 0 and not self.particles and not self.fem and not self.bindings
    dt32 = F32(dt)
    if self.frame_substep == 0 and self.has_obs:
        self.obs_bvh.Build(self.o_x0, self.o_tris)
        self.pass_cnt = 0
    freq = max(1, xp.num_substeps * xp.num_iterations // xp.num_collision_passes)
    g = np.array(list(p.gravity), F32)
    V, X0 = c.v.reshape(-1, 3).astype(F32), c.x0.reshape(-1, 3).astype(F32)
    V = np.where((c.w > 0)[:, None], V + g * dt32, V)
    c.v[:] = V.reshape(-1)
    c.x[:] = (X0 + V * dt32).reshape(-1)
    self._rec("predict")
    d, i = osc._d, osc._i
    ot = np.ascontiguousarray(self.o_tris).reshape(-1)
    for it in range(p.num_iterations):
        if (self.frame_substep * p.num_iterations + it) % freq == 0:
            alpha = F32(F32(self.pass_cnt) + F32(1.0)) / F32(xp.num_collision_passes)
            x1 = ((F32(1.0) - alpha) * self.o_x0 + alpha * self.o_x).astype(F32)
            self.obs_bvh.RefitMidstep(x1, orc.vertex_normals(x1, self.o_tris), xp.bvh_tolerance)
            lib.orc_vf_dcd_generate(self.obs_bvh.h, i(ot), d(c.x), self.nv, i(self.vf_face),
                                    d(self.vf_q))
            lib.orc_vf_ccd_generate(self.obs_bvh.h, i(ot), d(c.x0), d(c.x), self.nv,
                                    i(self.rt_face), d(self.rt_q))
            lib.orc_ee_dcd_generate(self.obs_bvh.h, i(ot), self.o_rtri.ctypes.data_as(osc._U),
                                    d(c.x), i(self.edges), self.ne, d(self.ee))
            self.pass_cnt += 1
            assert (self.vf_face < 0).all() and (self.rt_face < 0).all()
        for k in range(c.s_ncol):
            self._rec("stretch", k, it, c.x.copy(), c.s_lam.copy())
            _stretch_color(c, k, dt, it)
        pred = np.ctypeslib.as_array(lib.orc_bvh_pred(self.obs_bvh.h), (3 * len(self.o_x0),)).copy()
        self._rec("vf", pred, self.rt_face.copy(), self.rt_q.copy(), c.x.copy(), 0.0)
        self._rec("vf", pred, self.vf_face.copy(), self.vf_q.copy(), c.x.copy(), 1e-3)
        self._rec("ee", self.ee.copy(), c.x.copy(), xp.relaxation_factor)
    self._rec("finalize")
    X = c.x.reshape(-1, 3).astype(F32)
    fac = F32(F32(1.0) - F32(p.damping) * dt32)
    c.v[:] = ((X - X0) * (F32(1.0) / dt32) * fac).reshape(-1)
    c.x0[:] = X.reshape(-1)
    self.frame_substep += 1
    if self.frame_substep >= xp.num_substeps:
        self.frame_substep = 0
        self.o_x0 = self.o_x.copy()


def use_fp32_forward(scene: osc.ContactScene) -> osc.ContactScene:
    """Switch `scene`'s substep to the FP32-arithmetic one; returns the scene."""
    scene.substep = types.MethodType(_substep, scene)
    return scene
