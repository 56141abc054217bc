"""Oracle orchestration of a substep with contacts (test infrastructure only).

Restates the substep order over the oracle's C primitives: the union of
ClothSolverBase::SimulateSwift's body-cloth PBD collision
(xrtailor/src/pipeline/base/ClothSolverBase.cpp:190-240) and SimulateQuality's PBD repulsion
(:277-313):

  frame start: obstacle LBVH from its x0 (tol 0); cloth LBVH from x0
  predict; particle spatial hash when frame substep % inter_leaved_hash == 0
    (SimulateSwift :193-197)
  per iteration:
    particle collision: CollideParticles, ApplyDeltas x relaxation factor (:199-202)
    collision pass when (substep*iters + it) % (substeps*iters/passes) == 0:
      x1 = lerp(x0, x, (cnt+1)/passes) of the obstacle, midstep normals,
      RefitMidstep(bvh_tolerance), generate VF-DCD, VF-CCD, EE-DCD, cnt++
    pair refresh when it % interval == 0: cloth LBVH Update(x, x0, ccd) and
      Traverse(cloth, obstacle) ++ Traverse(cloth, cloth) at thickness 1e-6
    repulsion over the pairs, applied x relaxation rate
    stretch colors, bend colors (or, with setup_fem, the quality mode's FEM strain colors
      and isometric bending colors, :314-323), LRA, ground
    VF-CCD, VF-DCD solves; EE-DCD deltas applied x relaxation factor
    bindings: BindingConstraint::Solve against the obstacle's frame positions (:237)
  finalize; frame end: obstacle x0 = x
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import oracle as orc
from . import particle as opart
from . import binding as obind
from . import fem as ofem

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)
_U = C.POINTER(C.c_uint32)
_B = C.POINTER(C.c_uint8)


class RMeshIn(C.Structure):
    _fields_ = [("x", _D), ("x0", _D), ("w", _D), ("free_", _B), ("tris", _I), ("edges", _I),
                ("fe", _I), ("ef_ptr", _I), ("ef", _I), ("face_n", _D), ("rtri", _U)]


class RepItem(C.Structure):
    """One recorded repulsion response (orc_rep_item, xrt_oracle.h)."""
    _fields_ = [("kind", C.c_int), ("other_cloth", C.c_int), ("contrib", C.c_int),
                ("pn_x0", C.c_int), ("ids", C.c_int * 4), ("side", C.c_double),
                ("pn_sign", C.c_double), ("pn", C.c_double * 3)]


_bound = False


def _lib():
    global _bound
    lib = orc._bind_bvh()
    if not _bound:
        lib.orc_face_edges.argtypes = [_I, C.c_int, _I, C.c_int, _I]
        lib.orc_face_edges.restype = None
        lib.orc_edge_faces.argtypes = [_I, C.c_int, C.c_int, _I, _I]
        lib.orc_rtri.argtypes = [_I, C.c_int, C.c_int, _U]
        lib.orc_rtri.restype = None
        lib.orc_nearest_point.argtypes = [_D, _D, _D]
        lib.orc_nearest_point.restype = C.c_double
        lib.orc_ray_intersect.argtypes = [_D, _D, C.c_double, _D]
        lib.orc_ray_intersect.restype = C.c_double
        lib.orc_closest_seg_seg.argtypes = [_D, _D, _D, _D, _D, _D]
        lib.orc_closest_seg_seg.restype = C.c_double
        lib.orc_signed_vf.argtypes = [_D, _D, _D]
        lib.orc_signed_vf.restype = C.c_double
        lib.orc_signed_ee.argtypes = [_D, _D, _D]
        lib.orc_signed_ee.restype = C.c_double
        lib.orc_vf_dcd_generate.argtypes = [C.c_void_p, _I, _D, C.c_int, _I, _D]
        lib.orc_vf_dcd_generate.restype = None
        lib.orc_vf_ccd_generate.argtypes = [C.c_void_p, _I, _D, _D, C.c_int, _I, _D]
        lib.orc_vf_ccd_generate.restype = None
        lib.orc_ee_dcd_generate.argtypes = [C.c_void_p, _I, _U, _D, _I, C.c_int, _D]
        lib.orc_ee_dcd_generate.restype = None
        lib.orc_bvh_pred.argtypes = [C.c_void_p]
        lib.orc_bvh_pred.restype = C.POINTER(C.c_float)
        _Fp = C.POINTER(C.c_float)
        lib.orc_vf_solve.argtypes = [_Fp, _I, _I, _D, _D, C.c_int, C.c_double, _D, _Fp, _Fp,
                                     C.c_double, C.c_double]
        lib.orc_vf_solve.restype = None
        lib.orc_ee_solve.argtypes = [_D, _D, _I, C.c_int, _D, _I]
        lib.orc_ee_solve.restype = None
        lib.orc_apply_deltas.argtypes = [_D, _D, _I, C.c_int, C.c_double]
        lib.orc_apply_deltas.restype = None
        lib.orc_repulsion.argtypes = [C.POINTER(RMeshIn), C.POINTER(RMeshIn), _I, _I, C.c_long,
                                      C.c_double, _D, _I]
        lib.orc_repulsion.restype = None
        lib.orc_repulsion_record.argtypes = [C.POINTER(RMeshIn), C.POINTER(RMeshIn), _I, _I,
                                             C.c_long, C.c_double, _D, _I, C.POINTER(RepItem),
                                             C.c_long]
        lib.orc_repulsion_record.restype = C.c_long
        _bound = True
    return lib


def _i(a):
    return a.ctypes.data_as(_I)


def _d(a):
    return a.ctypes.data_as(_D)


def contact_topology(pos, tris, edges):
    """(face_edges, ef_ptr, ef, rtri, rest face normals) of a mesh."""
    lib = _lib()
    t = np.ascontiguousarray(tris, np.int32).reshape(-1)
    e = np.ascontiguousarray(edges, np.int32).reshape(-1)
    nt, ne, nv = t.size // 3, e.size // 2, len(pos)
    fe = np.zeros(3 * nt, np.int32)
    lib.orc_face_edges(_i(t), nt, _i(e), ne, _i(fe))
    ptr = np.zeros(ne + 1, np.int32)
    ef = np.zeros(3 * nt, np.int32)
    lib.orc_edge_faces(_i(fe), nt, ne, _i(ptr), _i(ef))
    rt = np.zeros(nt, np.uint32)
    lib.orc_rtri(_i(t), nt, nv, rt.ctypes.data_as(_U))
    p = np.asarray(pos, np.float64)[np.asarray(tris)]
    n = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    n = n / np.linalg.norm(n, axis=1, keepdims=True)
    return fe, ptr, ef, rt, np.ascontiguousarray(n.reshape(-1))


class ContactScene:
    """Oracle twin of one solver instance with its obstacle and self contact."""

    def __init__(self, cloth: orc.Cloth, rest_pos, tris, edges, params: orc.OrcParams, xp,
                 obstacle=None):
        self.lib = _lib()
        self.c = cloth
        self.p = params
        self.xp = xp
        self.tris = np.ascontiguousarray(tris, np.int32)
        self.edges = np.ascontiguousarray(edges, np.int32)
        self.nv, self.nf, self.ne = cloth.nv, len(self.tris), len(self.edges)
        self.fe, self.efp, self.ef, self.rtri, self.fn = contact_topology(rest_pos, self.tris, self.edges)
        self.free = np.ascontiguousarray((cloth.w > 0).astype(np.uint8))
        self.has_obs = obstacle is not None
        if self.has_obs:
            opos, otris = obstacle
            self.o_tris = np.ascontiguousarray(otris, np.int32)
            self.o_edges = orc.build_edges(self.o_tris)
            self.o_x0 = np.ascontiguousarray(opos, np.float32).copy()
            self.o_x = self.o_x0.copy()
            (self.o_fe, self.o_efp, self.o_ef, self.o_rtri, self.o_fn) = contact_topology(
                opos, self.o_tris, self.o_edges)
            self.o_w = np.full(len(opos), 1e-3)
            self.o_free = np.zeros(len(opos), np.uint8)
            self.obs_bvh = orc.Bvh()
        self.cloth_bvh = orc.Bvh()
        self.vf_face = np.full(self.nv, -1, np.int32)
        self.vf_q = np.zeros(3 * self.nv)
        self.rt_face = np.full(self.nv, -1, np.int32)
        self.rt_q = np.zeros(3 * self.nv)
        self.ee = np.zeros(12 * max(self.ne, 1))
        self.delta = np.zeros(3 * self.nv)
        self.count = np.zeros(self.nv, np.int32)
        self.pairs = np.zeros((0, 2), np.int32)
        self.kind = np.zeros(0, np.int32)
        self.n_obs_pairs = 0
        self.frame_substep = 0
        self.pass_cnt = 0
        self.dt = float(np.float32(xp.frame_time) / np.float32(xp.num_substeps))
        self.tape = None  # a list: substep() appends each operator's input state (adjoint.py)
        self.bindings = False
        self.fem = False
        # swift-mode particle self-collision: diameter from the rest mesh, initial positions
        # = the state now (FP32), neighbor lists zero until the first hash
        self.particles = bool(getattr(xp, "enable_particle_collision", 0))
        if self.particles:
            self.p_diam = opart.particle_diameter(rest_pos, tris, xp.particle_diameter)
            self.p_orig = cloth.x0.reshape(-1, 3).astype(np.float32)
            self.p_nb = np.zeros((xp.max_num_neighbors, self.nv), np.uint32)

    def setup_bindings(self, vert, stiffness, distance, tol):
        """QueryNearestBarycentricClothToObstacle over the cloth's current x0 (the obstacle
        tree from its x0 with tolerance `tol` along the vertex normals, FP32) and
        BindingConstraint::Generate (FP64) for the bound vertices."""
        ox0 = np.ascontiguousarray(self.o_x0, np.float32)
        tree = orc.Bvh()
        tree.Build(ox0, self.o_tris, tol=tol, n=orc.vertex_normals(ox0, self.o_tris))
        faces, uvw, dist = obind.nearest(tree, self.c.x0.reshape(-1, 3).astype(np.float32))
        self.near = (faces, uvw, dist)
        vert = np.ascontiguousarray(vert, np.int32)
        self.b_vert = vert
        self.b_tri3 = np.ascontiguousarray(self.o_tris[faces[vert]], np.int32)
        self.b_uvw = uvw[vert].astype(np.float64)
        self.b_stiff = np.broadcast_to(np.asarray(stiffness, np.float64), vert.shape).copy()
        d_in = np.broadcast_to(np.asarray(distance, np.float32), vert.shape)
        self.b_L = obind.generate(self.c.x0.reshape(-1, 3), ox0.astype(np.float64), vert,
                                  self.b_tri3, self.b_uvw, d_in.astype(np.float64))
        self.bindings = True

    def setup_fem(self, face_colors, n_colors, rest_pos, E, bend_stiffness):
        """The quality mode's FEM constraints in place of stretch and bend: faces sorted by
        (color, index), rest data in FP64 from the FP32 rest positions and rounded to FP32
        as the solver stores them, the isometric data over the bend elements in their color
        order.  E = (xx, yy, xy stiffness, xy, yx Poisson ratio) as FP32 values."""
        order, self.f_off = orc.color_sort(np.asarray(face_colors), n_colors)
        rp = np.asarray(rest_pos, np.float32).astype(np.float64)
        self.f_tris = np.ascontiguousarray(self.tris[order], np.int32)
        area, inv = ofem.strain_init(rp, self.f_tris)
        self.f_area = area.astype(np.float32).astype(np.float64)
        self.f_inv = np.ascontiguousarray(inv.astype(np.float32).astype(np.float64))
        self.f_E = np.asarray(E, np.float32).astype(np.float64)
        K, cf = ofem.iso_init(rp, self.c.b_idx.reshape(-1, 4))
        self.iso_K = np.ascontiguousarray(K.astype(np.float32).astype(np.float64))
        self.iso_c = cf.astype(np.float32).astype(np.float64)
        self.iso_lam = np.zeros(len(cf))
        self.f_bstiff = float(np.float32(bend_stiffness))
        self.f_ncol = n_colors
        self.fem = True

    def _vf_friction(self):
        """Friction arguments of orc_vf_solve: the cloth's start-of-substep positions, the
        obstacle's frame-start / frame-end positions, 1 / substeps, contact_friction."""
        F = C.POINTER(C.c_float)
        self._ox0 = np.ascontiguousarray(self.o_x0, np.float32)
        self._ox = np.ascontiguousarray(self.o_x, np.float32)
        return (_d(self.c.x0), self._ox0.ctypes.data_as(F), self._ox.ctypes.data_as(F),
                1.0 / self.xp.num_substeps, float(getattr(self.xp, "contact_friction", 0.0)))

    def _rec(self, *entry):
        if self.tape is not None:
            self.tape.append(entry)

    def _rmesh(self, obstacle: bool) -> RMeshIn:
        if not obstacle:
            return RMeshIn(_d(self.c.x), _d(self.c.x0), _d(self.c.w), self.free.ctypes.data_as(_B),
                           _i(self.tris), _i(self.edges), _i(self.fe), _i(self.efp), _i(self.ef),
                           _d(self.fn), self.rtri.ctypes.data_as(_U))
        self._ox = np.ascontiguousarray(self.o_x, np.float64).reshape(-1)
        self._ox0 = np.ascontiguousarray(self.o_x0, np.float64).reshape(-1)
        return RMeshIn(_d(self._ox), _d(self._ox0), _d(self.o_w), self.o_free.ctypes.data_as(_B),
                       _i(self.o_tris), _i(self.o_edges), _i(self.o_fe), _i(self.o_efp),
                       _i(self.o_ef), _d(self.o_fn), self.o_rtri.ctypes.data_as(_U))

    def substep(self):
        lib, c, p, xp, dt = self.lib, self.c, self.p, self.xp, self.dt
        has_obs = self.has_obs and xp.enable_obstacle_contact
        self_c = bool(xp.enable_self_contact)
        if self.frame_substep == 0:
            if self.has_obs:
                self.obs_bvh.Build(self.o_x0, self.o_tris)
            if self_c:
                self.cloth_bvh.Build(c.x0.reshape(-1, 3).astype(np.float32), self.tris)
            self.pass_cnt = 0
        freq = max(1, xp.num_substeps * xp.num_iterations // xp.num_collision_passes)
        interval = max(1, xp.pair_update_interval)
        lib.orc_predict(C.byref(c.s), C.byref(p), dt)
        self._rec("predict")
        if self.particles and self.frame_substep % xp.inter_leaved_hash == 0:
            self.p_nb = opart.neighbors(c.x.reshape(-1, 3).astype(np.float32), self.p_orig,
                                        self.p_diam, xp.hash_cell_size,
                                        xp.max_num_neighbors)["neighbors"]
        x0s = c.x0.copy() if self.tape is not None else None
        for it in range(p.num_iterations):
            step = self.frame_substep * p.num_iterations + it
            if self.particles:
                pd, pc = opart.collide(c.x, c.x0, c.w, self.p_nb, self.p_diam, xp.ground_friction)
                pd = np.ascontiguousarray(pd.reshape(-1))
                lib.orc_apply_deltas(_d(c.x), _d(pd), _i(pc), self.nv, xp.relaxation_factor)
            if has_obs and step % freq == 0:
                alpha = np.float32(np.float32(self.pass_cnt) + np.float32(1.0)) / np.float32(xp.num_collision_passes)
                x1 = (np.float32(1.0) - alpha) * self.o_x0 + alpha * self.o_x
                x1 = x1.astype(np.float32)
                n1 = orc.vertex_normals(x1, self.o_tris)
                self.obs_bvh.RefitMidstep(x1, n1, xp.bvh_tolerance)
                ot = np.ascontiguousarray(self.o_tris).reshape(-1)
                lib.orc_vf_dcd_generate(self.obs_bvh.h, _i(ot), _d(c.x), self.nv, _i(self.vf_face), _d(self.vf_q))
                lib.orc_vf_ccd_generate(self.obs_bvh.h, _i(ot), _d(c.x0), _d(c.x), self.nv,
                                        _i(self.rt_face), _d(self.rt_q))
                lib.orc_ee_dcd_generate(self.obs_bvh.h, _i(ot), self.o_rtri.ctypes.data_as(_U), _d(c.x),
                                        _i(self.edges), self.ne, _d(self.ee))
                self.pass_cnt += 1
            if self_c and it % interval == 0:
                self.cloth_bvh.Update(c.x.reshape(-1, 3).astype(np.float32),
                                      c.x0.reshape(-1, 3).astype(np.float32), ccd=True)
                po = orc.traverse(self.cloth_bvh, self.obs_bvh, 1e-6) if self.has_obs else np.zeros((0, 2), np.int32)
                ps = orc.traverse(self.cloth_bvh, self.cloth_bvh, 1e-6)
                self.n_obs_pairs = len(po)
                self.pairs = np.ascontiguousarray(np.concatenate([po, ps]), np.int32)
                self.kind = np.ascontiguousarray(np.r_[np.zeros(len(po)), np.ones(len(ps))], np.int32)
            if self_c:
                cm = self._rmesh(False)
                om = self._rmesh(True) if self.has_obs else None
                args = (C.byref(cm), C.byref(om) if om is not None else None,
                        _i(self.pairs.reshape(-1)), _i(self.kind), len(self.pairs),
                        xp.pbd_thickness, _d(self.delta), _i(self.count))
                if self.tape is not None:
                    # the responses and counts of this pass, for the reverse (the same
                    # evaluation as orc_repulsion)
                    xpre = c.x.copy()
                    cap = max(64, 12 * len(self.pairs))
                    items = (RepItem * cap)()
                    n = lib.orc_repulsion_record(*args, items, cap)
                    assert n <= cap
                    self._rec("repulsion", items[:n] if n else [], xpre, c.x0.copy(),
                              self.count.copy(), xp.pbd_thickness, xp.repulsion_relaxation)
                else:
                    lib.orc_repulsion(*args)
                lib.orc_apply_deltas(_d(c.x), _d(self.delta), _i(self.count), self.nv, xp.repulsion_relaxation)
            if self.fem:
                for k in range(self.f_ncol):
                    ofem.strain_solve(c.x, c.w, self.f_tris, self.f_area, self.f_inv,
                                      int(self.f_off[k]), int(self.f_off[k + 1]), self.f_E)
                if p.solve_bending:
                    for k in range(c.b_ncol):
                        ofem.iso_solve(c.x, c.w, c.b_idx, self.iso_K, self.iso_c, self.iso_lam,
                                       int(c.b_off[k]), int(c.b_off[k + 1]), it, dt, self.f_bstiff)
            for k in range(0 if self.fem else c.s_ncol):
                self._rec("stretch", k, it, c.x.copy(), c.s_lam.copy())
                lib.orc_stretch_color(C.byref(c.s), k, dt, it)
            if p.solve_bending and not self.fem:
                for k in range(c.b_ncol):
                    self._rec("bend", k, c.x.copy())
                    lib.orc_bend_color(C.byref(c.s), C.byref(p), k, dt)
            if len(c.slot_vert) > 0:
                self._rec("lra", c.x.copy())
            lib.orc_lra(C.byref(c.s), C.byref(p))
            if p.enable_ground:
                self._rec("ground", c.x.copy(), x0s)
            lib.orc_ground(C.byref(c.s), C.byref(p))
            if has_obs:
                pred = lib.orc_bvh_pred(self.obs_bvh.h)
                ot = np.ascontiguousarray(self.o_tris).reshape(-1)
                if self.tape is not None:
                    predc = np.ctypeslib.as_array(pred, (3 * len(self.o_x0),)).copy()
                    self._rec("vf", predc, self.rt_face.copy(), self.rt_q.copy(), c.x.copy(), 0.0)
                fr = self._vf_friction()
                lib.orc_vf_solve(pred, _i(ot), _i(self.rt_face), _d(self.rt_q), _d(c.x), self.nv, 0.0, *fr)
                if self.tape is not None:
                    self._rec("vf", predc, self.vf_face.copy(), self.vf_q.copy(), c.x.copy(), 1e-3)
                lib.orc_vf_solve(pred, _i(ot), _i(self.vf_face), _d(self.vf_q), _d(c.x), self.nv, 1e-3, *fr)
                self._rec("ee", self.ee.copy(), c.x.copy(), xp.relaxation_factor)
                lib.orc_ee_solve(_d(self.ee), _d(c.x), _i(self.edges), self.ne, _d(self.delta), _i(self.count))
                lib.orc_apply_deltas(_d(c.x), _d(self.delta), _i(self.count), self.nv, xp.relaxation_factor)
            if self.bindings:
                obind.solve(c.x.reshape(-1, 3), c.w, np.asarray(self.o_x, np.float64), self.b_vert,
                            self.b_tri3, self.b_uvw, self.b_L, self.b_stiff)
        self._rec("finalize")
        lib.orc_finalize(C.byref(c.s), C.byref(p), dt)
        self.frame_substep += 1
        if self.frame_substep >= xp.num_substeps:
            self.frame_substep = 0
            if self.has_obs:
                self.o_x0 = self.o_x.copy()
