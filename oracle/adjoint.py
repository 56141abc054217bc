"""FP64 host reference gradient of an oracle rollout (test infrastructure only).

The forward is ContactScene.substep with its tape switched on (each operator's input
state is recorded); the reverse walks the tape backwards and applies each operator's
vector-Jacobian product from adjoint_oracle.c.  Predict and finalize are linear and are
reversed here:

  predict   x = x0 + (v + g dt [free]) dt      ->  gx0 += gx,  gv = gx dt
  finalize  v = (x - x0) / dt * (1 - damping dt), x0 = x
                                               ->  gx = gx0 + gv k / dt,  gx0 = -gv k / dt

Contact records (VF face and q, EE s, t, fn) are held fixed, as the library's adjoint
does.  Long-range attachments replay the slot snapshot of the pass (the pre-LRA state) and
each target's slot sequence.  PBD repulsion replays the pass's recorded responses (the
orientation of each VF response and each EE normal's source are held fixed; the normals of
free edge pairs follow the start-of-substep state) with the pass's Jacobi counts.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import oracle as orc
from . import scene as osc

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)
_F = C.POINTER(C.c_float)
_bound = False


def _lib():
    global _bound
    lib = osc._lib()
    if not _bound:
        P, Cl = C.POINTER(orc.OrcParams), C.POINTER(orc.OrcCloth)
        lib.orc_adj_stretch_color.argtypes = [Cl, C.c_int, C.c_double, C.c_int, _D, _D, _D, _D]
        lib.orc_adj_bend_color.argtypes = [Cl, P, C.c_int, C.c_double, _D, _D]
        lib.orc_adj_ground.argtypes = [Cl, P, _D, _D]
        lib.orc_adj_lra.argtypes = [Cl, P, _D]
        lib.orc_adj_vf.argtypes = [_F, _I, _I, _D, _D, C.c_int, C.c_double, _D]
        lib.orc_adj_ee.argtypes = [_D, _D, _I, C.c_int, C.c_int, C.c_double, _D]
        lib.orc_adj_repulsion.argtypes = [C.POINTER(osc.RepItem), C.c_long,
                                          C.POINTER(osc.RMeshIn), C.POINTER(osc.RMeshIn), _I,
                                          C.c_int, C.c_double, C.c_double, _D, _D]
        for f in ("orc_adj_stretch_color", "orc_adj_bend_color", "orc_adj_ground", "orc_adj_vf",
                  "orc_adj_ee", "orc_adj_lra", "orc_adj_repulsion"):
            getattr(lib, f).restype = None
        _bound = True
    return lib


def _d(a):
    return a.ctypes.data_as(_D)


def record(scene: osc.ContactScene, n_substeps: int) -> list:
    """Run n_substeps forward with the tape on; returns the tape."""
    scene.tape = []
    try:
        for _ in range(n_substeps):
            scene.substep()
        return scene.tape
    finally:
        tape, scene.tape = scene.tape, None
        del tape


def backward(scene: osc.ContactScene, tape: list, gx_end, gv_end) -> dict:
    """Adjoints of the rollout's initial state and compliances from the adjoints of its
    final state (gx_end = dL/dx0, gv_end = dL/dv, shape (nv, 3)).  Compliance gradients
    come back in constraint order (the cloth's color-sorted order) and in the mesh's
    original edge / bend order."""
    lib = _lib()
    c, p, dt = scene.c, scene.p, scene.dt
    keep_x, keep_x0 = c.x.copy(), c.x0.copy()
    gx0 = np.ascontiguousarray(gx_end, np.float64).reshape(-1).copy()
    gv = np.ascontiguousarray(gv_end, np.float64).reshape(-1).copy()
    gx = np.zeros_like(gx0)
    glam = np.zeros(len(c.s_rest))
    gs = np.zeros(len(c.s_rest))
    gb = np.zeros(len(c.b_theta))
    kd = 1.0 - p.damping * dt
    ot = np.ascontiguousarray(scene.o_tris).reshape(-1) if scene.has_obs else None
    try:
        for e in reversed(tape):
            op = e[0]
            if op == "finalize":
                gx = gx0 + gv * (kd / dt)
                gx0 = -gv * (kd / dt)
                glam[:] = 0.0
            elif op == "predict":
                gx0 = gx0 + gx
                gv = gx * dt
            elif op == "stretch":
                _, k, it, xpre, lam_pre = e
                c.x[:] = xpre
                lib.orc_adj_stretch_color(C.byref(c.s), k, dt, it, _d(lam_pre), _d(gx), _d(glam),
                                          _d(gs))
            elif op == "bend":
                _, k, xpre = e
                c.x[:] = xpre
                lib.orc_adj_bend_color(C.byref(c.s), C.byref(p), k, dt, _d(gx), _d(gb))
            elif op == "lra":
                _, xpre = e
                c.x[:] = xpre
                lib.orc_adj_lra(C.byref(c.s), C.byref(p), _d(gx))
            elif op == "ground":
                _, xpre, x0s = e
                c.x[:] = xpre
                c.x0[:] = x0s
                lib.orc_adj_ground(C.byref(c.s), C.byref(p), _d(gx), _d(gx0))
            elif op == "vf":
                _, pred, face, q, xpre, off = e
                pf = np.ascontiguousarray(pred, np.float32)
                lib.orc_adj_vf(pf.ctypes.data_as(_F), ot.ctypes.data_as(_I),
                               face.ctypes.data_as(_I), _d(q), _d(xpre), c.nv, off, _d(gx))
            elif op == "ee":
                _, rec, xpre, relax = e
                lib.orc_adj_ee(_d(rec), _d(xpre), scene.edges.ctypes.data_as(_I), scene.ne, c.nv,
                               relax, _d(gx))
            elif op == "repulsion":
                _, items, xpre, x0s, count, h, relax = e
                c.x[:] = xpre
                c.x0[:] = x0s
                arr = (osc.RepItem * max(len(items), 1))(*items)
                cm = scene._rmesh(False)
                om = scene._rmesh(True) if scene.has_obs else None
                lib.orc_adj_repulsion(arr, len(items), C.byref(cm),
                                      C.byref(om) if om is not None else None,
                                      count.ctypes.data_as(_I), c.nv, h, relax, _d(gx), _d(gx0))
            else:
                raise ValueError(f"no adjoint for operator {op!r}")
    finally:
        c.x[:] = keep_x
        c.x0[:] = keep_x0
    gs_edge = np.zeros_like(gs)
    gs_edge[c.s_order] = gs
    gb_elem = np.zeros_like(gb)
    gb_elem[c.b_order] = gb
    return dict(gx0=gx0.reshape(-1, 3), gv0=gv.reshape(-1, 3), gs=gs, gb=gb, gs_edge=gs_edge,
                gb_elem=gb_elem)
