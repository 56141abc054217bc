"""Python mirror of the solver's C ABI (used by the tests and the benchmark).

Class and method names follow the reference's C++ interfaces for this path:
``ClothSolver`` ~ ``ClothSolverBase`` (xrtailor/src/pipeline/base/ClothSolverBase.cpp),
``StretchSolver`` / ``BendSolver`` ~ ``BasicConstraint::{Stretch,Bend}Solver``
(xrtailor/include/xrtailor/physics/dynamics/BasicSolver.cuh:15-54).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _ffi
from ._ffi import XrtParams, check, fptr, iptr


def default_params() -> XrtParams:
    p = XrtParams()
    _ffi.load().xrt_default_params(C.byref(p))
    return p


def make_params(**kw) -> XrtParams:
    """default_params() with fields overridden by keyword (gravity as a 3-sequence)."""
    p = default_params()
    for k, v in kw.items():
        if k == "gravity":
            for i in range(3):
                p.gravity[i] = v[i]
        else:
            setattr(p, k, v)
    return p


def build_solver(specs, params: XrtParams, seed: int = 0, device: int = 0,
                 obstacles: bool = True, before_start=None):
    """A started ClothSolver with one instance per scenes.ClothSpec (pins, LRA slots, the
    obstacle when `obstacles` and the jittered initial state applied; `before_start(solver,
    instance ids)` runs last before Start, e.g. to add bindings).
    Returns (solver, instance ids)."""
    s = ClothSolver(device=device, params=params)
    insts = []
    for spec in specs:
        i = s.AddCloth(spec.pos, spec.tris, spec.stretch_compliance, spec.bend_compliance)
        if len(spec.pinned):
            s.Pin(i, spec.pinned)
        if len(spec.lra_slots):
            s.AddLraSlots(i, spec.lra_slots)
        if obstacles and spec.obstacle is not None:
            s.SetObstacle(i, *spec.obstacle)
        s.SetState(i, spec.x0)
        insts.append(i)
    if before_start is not None:
        before_start(s, insts)
    s.Start(seed)
    return s, insts


@dataclass
class Counts:
    nv: int
    nt: int
    n_edges: int
    n_bends: int
    stretch_colors: int
    bend_colors: int
    lra_slots: int
    lra_constraints: int


class _Coloring:
    def __init__(self, n_colors: int):
        self.n_colors = n_colors


class StretchSolver:
    """BasicConstraint::StretchSolver over one instance of a started ClothSolver."""

    def __init__(self, solver: "ClothSolver", inst: int):
        self._s, self._i = solver, inst

    def Coloring(self) -> _Coloring:
        return _Coloring(self._s.counts(self._i).stretch_colors)

    def Solve(self, color: int, dt: float, iteration: int) -> None:
        check(self._s.lib.xrt_stretch_solve(self._s.ctx, self._i, color, dt, iteration))


class BendSolver:
    """BasicConstraint::BendSolver over one instance of a started ClothSolver."""

    def __init__(self, solver: "ClothSolver", inst: int):
        self._s, self._i = solver, inst

    def Coloring(self) -> _Coloring:
        return _Coloring(self._s.counts(self._i).bend_colors)

    def Solve(self, color: int, dt: float, iteration: int = 0) -> None:
        check(self._s.lib.xrt_bend_solve(self._s.ctx, self._i, color, dt, iteration))


class StrainSolver:
    """FEM::StrainSolver (FEMSolver.cu:9-42) over one instance of a started ClothSolver
    whose constraint_model is FEM."""

    def __init__(self, solver: "ClothSolver", inst: int):
        self._s, self._i = solver, inst

    def Coloring(self) -> _Coloring:
        return _Coloring(self._s.fem(self._i)["n_colors"])

    def Solve(self, color: int) -> None:
        check(self._s.lib.xrt_strain_solve(self._s.ctx, self._i, color))


class IsometricBendingSolver:
    """FEM::IsometricBendingSolver (FEMSolver.cu:44-115) over one instance of a started
    ClothSolver whose constraint_model is FEM; it shares the bend elements and colors."""

    def __init__(self, solver: "ClothSolver", inst: int):
        self._s, self._i = solver, inst

    def Coloring(self) -> _Coloring:
        return _Coloring(self._s.counts(self._i).bend_colors)

    def Solve(self, color: int, iteration: int, dt: float) -> None:
        check(self._s.lib.xrt_isometric_bending_solve(self._s.ctx, self._i, color, iteration, dt))


def write_obj(path: str, pos, tris, normals=None) -> None:
    """SaveObj: one mesh as OBJ text (the C ABI's writer)."""
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1)
    t = np.ascontiguousarray(tris, np.int32).reshape(-1)
    nrm = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1)
    check(_ffi.load().xrt_write_obj(str(path).encode(), fptr(pos), fptr(nrm), pos.size // 3,
                                    iptr(t), t.size // 3))


def write_obj_sequence(directory: str, frames, tris, normals=None, first_frame: int = 0) -> None:
    """ExportObjSequence: frames (n, nv, 3) of one mesh as directory/%06d.obj."""
    fr = np.ascontiguousarray(frames, np.float32)
    n, nv = fr.shape[0], fr.shape[1]
    t = np.ascontiguousarray(tris, np.int32).reshape(-1)
    nrm = None if normals is None else np.ascontiguousarray(normals, np.float32)
    check(_ffi.load().xrt_write_obj_sequence(str(directory).encode(), first_frame, n,
                                             fptr(fr.reshape(-1)),
                                             None if nrm is None else fptr(nrm.reshape(-1)),
                                             nv, iptr(t), t.size // 3))


class SpatialHash:
    """SpatialHashGPU of one instance (xrtailor/include/xrtailor/physics/broad_phase/
    spatial_hash/SpatialHashGPU.cuh:19-35): Hash() over the predicted positions and its
    public buffers."""

    def __init__(self, solver: "ClothSolver", inst: int):
        self._s, self._i = solver, inst

    def Hash(self) -> None:
        check(self._s.lib.xrt_particle_hash(self._s.ctx, self._i))

    def buffers(self) -> dict:
        """neighbors (max_num_neighbors, nv): row k holds every particle's k-th neighbor;
        particle_hash / particle_index: the sorted pairs; cell_start / cell_end (2 nv)."""
        s, i = self._s, self._i
        nv, m = s.counts(i).nv, s.params.max_num_neighbors
        u32 = C.POINTER(C.c_uint32)
        nb = np.zeros(nv * m, np.uint32)
        ph = np.zeros(nv, np.uint32)
        pi = np.zeros(nv, np.int32)
        cs = np.zeros(2 * nv, np.uint32)
        ce = np.zeros(2 * nv, np.uint32)
        info = np.zeros(2, np.float32)
        check(s.lib.xrt_get_particle_hash(s.ctx, i, nb.ctypes.data_as(u32), ph.ctypes.data_as(u32),
                                          iptr(pi), cs.ctypes.data_as(u32), ce.ctypes.data_as(u32),
                                          fptr(info)))
        return {"neighbors": nb.reshape(m, nv), "particle_hash": ph, "particle_index": pi,
                "cell_start": cs, "cell_end": ce, "diameter": float(info[0]),
                "cell_spacing": float(info[1])}


class _InstanceCall:
    def __init__(self, solver: "ClothSolver", inst: int):
        self._s, self._i = solver, inst

    def _call(self, name: str, *args) -> None:
        check(getattr(self._s.lib, name)(self._s.ctx, self._i, *args))


class VFConstraint(_InstanceCall):
    """PBDCollision::DCD::VFConstraint of one instance against its obstacle tree
    (xrtailor/include/xrtailor/physics/pbd_collision/Constraint.cuh:15-30)."""

    def Generate(self) -> None:
        self._call("xrt_vf_dcd_generate")

    def Solve(self) -> None:
        self._call("xrt_vf_dcd_solve")


class RTConstraint(_InstanceCall):
    """PBDCollision::CCD::RTConstraint (VF continuous collision), Constraint.cuh:54-69."""

    def Generate(self) -> None:
        self._call("xrt_vf_ccd_generate")

    def Solve(self) -> None:
        self._call("xrt_vf_ccd_solve")


class EEConstraint(_InstanceCall):
    """PBDCollision::DCD::EEConstraint, Constraint.cuh:32-48; Solve accumulates the deltas
    that ClothSolver.ApplyDeltas applies."""

    def Generate(self) -> None:
        self._call("xrt_ee_dcd_generate")

    def Solve(self) -> None:
        self._call("xrt_ee_dcd_solve")


class PBDRepulsionSolver(_InstanceCall):
    """PBDRepulsionSolver of one instance (xrtailor/include/xrtailor/physics/repulsion/
    PBDRepulsion.cuh:11-40)."""

    def UpdatePairs(self, thickness: float) -> None:
        self._call("xrt_repulsion_update_pairs", thickness)

    def SolveNoProximity(self, thickness: float, relaxation_rate: float) -> None:
        self._call("xrt_repulsion_solve_no_proximity", thickness, relaxation_rate)


class ObstacleMesh(_InstanceCall):
    """The obstacle PhysicsMesh of one instance and its LBVH (PhysicsMesh.cuh:84-128,
    BVH.cuh:90-96)."""

    def Interpolate(self, step: int, num_steps: int) -> None:
        self._call("xrt_obstacle_interpolate", step, num_steps)

    def UpdateMidstepNormals(self) -> None:
        self._call("xrt_obstacle_update_midstep_normals")

    def PredictedToPositions(self) -> None:
        self._call("xrt_obstacle_predicted_to_positions")

    def ConstructBVH(self) -> None:
        self._call("xrt_obstacle_bvh_construct")

    def RefitMidstep(self, tolerance: float) -> None:
        self._call("xrt_obstacle_bvh_refit_midstep", tolerance)


class ClothSolver:
    """Owns one xrt_ctx (device state of one or more cloth instances)."""

    def __init__(self, device: int = 0, params: XrtParams | None = None):
        self.lib = _ffi.load()
        ctx = C.c_void_p()
        check(self.lib.xrt_create(C.byref(ctx), device))
        self.ctx = ctx
        self.params = params if params is not None else default_params()
        self._nbind = {}
        check(self.lib.xrt_set_params(self.ctx, C.byref(self.params)))

    # -------------------------------------------------------------- lifetime
    def close(self) -> None:
        if self.ctx:
            self.lib.xrt_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ----------------------------------------------------------------- setup
    def set_params(self, params: XrtParams) -> None:
        check(self.lib.xrt_set_params(self.ctx, C.byref(params)))
        self.params = params

    def AddCloth(self, pos: np.ndarray, tris: np.ndarray, stretch_compliance: float,
                 bend_compliance: float) -> int:
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1)
        tris = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1)
        out = C.c_int32(-1)
        check(self.lib.xrt_add_cloth(self.ctx, fptr(pos), pos.size // 3, iptr(tris), tris.size // 3,
                                     stretch_compliance, bend_compliance, C.byref(out)))
        return out.value

    def Pin(self, inst: int, idx) -> None:
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        check(self.lib.xrt_pin_vertices(self.ctx, inst, iptr(idx), idx.size))

    def SetStretchCompliance(self, inst: int, comp: np.ndarray) -> None:
        comp = np.ascontiguousarray(comp, dtype=np.float32)
        check(self.lib.xrt_set_stretch_compliance(self.ctx, inst, fptr(comp)))

    def AddLraSlots(self, inst: int, slots) -> None:
        slots = np.ascontiguousarray(slots, dtype=np.int32)
        check(self.lib.xrt_add_lra_slots(self.ctx, inst, iptr(slots), slots.size))

    def SetObstacle(self, inst: int, pos: np.ndarray, tris: np.ndarray) -> None:
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1)
        tris = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1)
        check(self.lib.xrt_set_obstacle(self.ctx, inst, fptr(pos), pos.size // 3, iptr(tris),
                                        tris.size // 3))

    def AddBindings(self, inst: int, idx, stiffness, distance) -> None:
        """BindingConstraint::Add (before Start): cloth vertices bound to the obstacle."""
        idx = np.ascontiguousarray(idx, np.int32)
        st = np.ascontiguousarray(np.broadcast_to(stiffness, idx.shape), np.float32)
        di = np.ascontiguousarray(np.broadcast_to(distance, idx.shape), np.float32)
        check(self.lib.xrt_add_bindings(self.ctx, inst, iptr(idx), fptr(st), fptr(di), len(idx)))
        self._nbind[inst] = self._nbind.get(inst, 0) + len(idx)

    def SetObstacleSkinning(self, inst: int, ptr, joint, weight, n_joints: int) -> None:
        """Sparse LBS weights of the obstacle's rest vertices (before Start)."""
        ptr = np.ascontiguousarray(ptr, np.int32)
        joint = np.ascontiguousarray(joint, np.int32)
        weight = np.ascontiguousarray(weight, np.float32)
        check(self.lib.xrt_set_obstacle_skinning(self.ctx, inst, iptr(ptr), iptr(joint),
                                                 fptr(weight), n_joints))

    def SkinObstacles(self, transforms) -> None:
        """LBS of every skinned obstacle: transforms of all of them, instance order."""
        t = np.ascontiguousarray(np.concatenate([np.asarray(a, np.float32).reshape(-1)
                                                 for a in transforms]), np.float32)
        check(self.lib.xrt_skin_obstacles(self.ctx, fptr(t)))

    def SkinObstacle(self, inst: int, transforms) -> None:
        t = np.ascontiguousarray(transforms, np.float32).reshape(-1)
        check(self.lib.xrt_skin_obstacle(self.ctx, inst, fptr(t)))

    def QueryNearestBarycentric(self, inst: int):
        """(idx (nv, 4): face vertices u, v, w and face; uvw (nv, 4): weights, distance)."""
        nv = self.counts(inst).nv
        idx = np.zeros(4 * nv, np.int32)
        uvw = np.zeros(4 * nv, np.float32)
        check(self.lib.xrt_query_nearest_barycentric(self.ctx, inst, iptr(idx), fptr(uvw)))
        return idx.reshape(-1, 4), uvw.reshape(-1, 4)

    def Bindings(self, inst: int):
        """(idx (nb, 4): cloth vertex, body u, v, w; par (nb, 5): u, v, w, L, stiffness)."""
        nb = self._nbind.get(inst, 0)
        idx = np.zeros(4 * max(nb, 1), np.int32)
        par = np.zeros(5 * max(nb, 1), np.float32)
        check(self.lib.xrt_get_bindings(self.ctx, inst, iptr(idx), fptr(par)))
        return idx[: 4 * nb].reshape(-1, 4), par[: 5 * nb].reshape(-1, 5)

    def SolveBindings(self, inst: int) -> None:
        check(self.lib.xrt_solve_bindings(self.ctx, inst))

    def GetObstacle(self, inst: int, nv: int):
        x0 = np.zeros(3 * nv, np.float32)
        x = np.zeros(3 * nv, np.float32)
        check(self.lib.xrt_get_obstacle(self.ctx, inst, fptr(x0), fptr(x)))
        return x0.reshape(-1, 3), x.reshape(-1, 3)

    def UpdateObstacle(self, inst: int, pos: np.ndarray) -> None:
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1)
        check(self.lib.xrt_update_obstacle(self.ctx, inst, fptr(pos)))

    def Contacts(self, inst: int) -> dict:
        """Contact records of the last collision pass and the pair statistics."""
        n = self.counts(inst)
        vf = np.zeros(n.nv, np.int32)
        vq = np.zeros(3 * n.nv, np.float32)
        rt = np.zeros(n.nv, np.int32)
        rq = np.zeros(3 * n.nv, np.float32)
        ee = np.zeros(16 * max(n.n_edges, 1), np.float32)
        st = np.zeros(4, np.int32)
        check(self.lib.xrt_get_contacts(self.ctx, inst, iptr(vf), fptr(vq), iptr(rt), fptr(rq),
                                        fptr(ee), iptr(st)))
        return dict(vf_face=vf, vf_q=vq.reshape(-1, 3), rt_face=rt, rt_q=rq.reshape(-1, 3),
                    ee=ee[:16 * n.n_edges].reshape(-1, 16), obs_pairs=int(st[0]),
                    pairs=int(st[1]), pair_cap=int(st[2]), overflow=bool(st[3]))

    def Start(self, seed: int = 0) -> None:
        check(self.lib.xrt_start(self.ctx, C.c_uint64(seed)))

    # ------------------------------------------ reference-shaped collider objects
    def VFConstraint(self, inst: int) -> VFConstraint:
        return VFConstraint(self, inst)

    def RTConstraint(self, inst: int) -> RTConstraint:
        return RTConstraint(self, inst)

    def EEConstraint(self, inst: int) -> EEConstraint:
        return EEConstraint(self, inst)

    def PBDRepulsionSolver(self, inst: int) -> PBDRepulsionSolver:
        return PBDRepulsionSolver(self, inst)

    def Obstacle(self, inst: int) -> ObstacleMesh:
        return ObstacleMesh(self, inst)

    def ConstructClothBVH(self, inst: int) -> None:
        check(self.lib.xrt_cloth_bvh_construct(self.ctx, inst))

    # ----------------------------------------------------------------- query
    def num_instances(self) -> int:
        n = C.c_int32()
        check(self.lib.xrt_num_instances(self.ctx, C.byref(n)))
        return n.value

    def device_info(self) -> dict:
        info = (C.c_int64 * 4)()
        check(self.lib.xrt_device_info(self.ctx, info))
        return {"cu_count": info[0], "max_lds_bytes": info[1], "wave_size": info[2],
                "l2_bytes": info[3]}

    def counts(self, inst: int) -> Counts:
        c = np.zeros(8, np.int32)
        check(self.lib.xrt_get_counts(self.ctx, inst, iptr(c)))
        return Counts(*[int(v) for v in c])

    def edges(self, inst: int, colors: bool = True):
        n = self.counts(inst)
        e = np.zeros(2 * n.n_edges, np.int32)
        r = np.zeros(n.n_edges, np.float32)
        col = np.zeros(n.n_edges, np.int32) if colors else None
        check(self.lib.xrt_get_edges(self.ctx, inst, iptr(e), fptr(r), iptr(col)))
        return e.reshape(-1, 2), r, col

    def bends(self, inst: int, colors: bool = True):
        n = self.counts(inst)
        b = np.zeros(4 * n.n_bends, np.int32)
        t = np.zeros(n.n_bends, np.float32)
        col = np.zeros(n.n_bends, np.int32) if colors else None
        check(self.lib.xrt_get_bends(self.ctx, inst, iptr(b), fptr(t), iptr(col)))
        return b.reshape(-1, 4), t, col

    def lra(self, inst: int):
        n = self.counts(inst)
        sv = np.zeros(max(n.lra_slots, 1), np.int32)
        ptr = np.zeros(n.nv + 1, np.int32)
        sl = np.zeros(max(n.lra_constraints, 1), np.int32)
        rest = np.zeros(max(n.lra_constraints, 1), np.float32)
        check(self.lib.xrt_get_lra(self.ctx, inst, iptr(sv), iptr(ptr), iptr(sl), fptr(rest)))
        return sv[:n.lra_slots], ptr, sl[:n.lra_constraints], rest[:n.lra_constraints]

    def inv_mass(self, inst: int) -> np.ndarray:
        w = np.zeros(self.counts(inst).nv, np.float32)
        check(self.lib.xrt_get_inv_mass(self.ctx, inst, fptr(w)))
        return w

    # ----------------------------------------------------------------- state
    def SetState(self, inst: int, x0=None, v=None) -> None:
        x0 = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float32).reshape(-1)
        v = None if v is None else np.ascontiguousarray(v, dtype=np.float32).reshape(-1)
        check(self.lib.xrt_set_state(self.ctx, inst, fptr(x0), fptr(v)))

    def GetState(self, inst: int):
        nv = self.counts(inst).nv
        x0 = np.zeros(3 * nv, np.float32)
        v = np.zeros(3 * nv, np.float32)
        check(self.lib.xrt_get_state(self.ctx, inst, fptr(x0), fptr(v)))
        return x0.reshape(-1, 3), v.reshape(-1, 3)

    def GetPredicted(self, inst: int) -> np.ndarray:
        x = np.zeros(3 * self.counts(inst).nv, np.float32)
        check(self.lib.xrt_get_predicted(self.ctx, inst, fptr(x)))
        return x.reshape(-1, 3)

    # ------------------------------------------------------ reference-shaped
    def StretchSolver(self, inst: int = 0) -> StretchSolver:
        return StretchSolver(self, inst)

    def BendSolver(self, inst: int = 0) -> BendSolver:
        return BendSolver(self, inst)

    def StrainSolver(self, inst: int = 0) -> StrainSolver:
        return StrainSolver(self, inst)

    def IsometricBendingSolver(self, inst: int = 0) -> IsometricBendingSolver:
        return IsometricBendingSolver(self, inst)

    def fem(self, inst: int) -> dict:
        """The FEM rest data of one instance in face / bend order: face colors and their
        count, areas, inverse rest matrices (nt, 4), isometric weights K (nb, 4), coeff."""
        n = self.counts(inst)
        col = np.zeros(n.nt, np.int32)
        ncol = np.zeros(1, np.int32)
        area = np.zeros(n.nt, np.float32)
        inv = np.zeros(4 * n.nt, np.float32)
        k = np.zeros(4 * n.n_bends, np.float32)
        cf = np.zeros(n.n_bends, np.float32)
        check(self.lib.xrt_get_fem(self.ctx, inst, iptr(col), iptr(ncol), fptr(area), fptr(inv),
                                   fptr(k), fptr(cf)))
        return {"colors": col, "n_colors": int(ncol[0]), "area": area, "inv": inv.reshape(-1, 4),
                "K": k.reshape(-1, 4), "coeff": cf}

    def PredictPositions(self, inst: int, dt: float) -> None:
        check(self.lib.xrt_predict_positions(self.ctx, inst, dt))

    def SolveLongRangeAttachments(self, inst: int) -> None:
        check(self.lib.xrt_solve_lra(self.ctx, inst))

    def SolveGround(self, inst: int) -> None:
        check(self.lib.xrt_solve_ground(self.ctx, inst))

    def ApplyDeltas(self, inst: int) -> None:
        check(self.lib.xrt_apply_deltas(self.ctx, inst))

    def Finalize(self, inst: int, dt: float) -> None:
        check(self.lib.xrt_finalize(self.ctx, inst, dt))

    def SpatialHash(self, inst: int = 0) -> SpatialHash:
        return SpatialHash(self, inst)

    def CollideParticles(self, inst: int) -> None:
        """CollideParticles + ApplyDeltas x relaxation_factor (ClothSolverHelper.cu:171-219)."""
        check(self.lib.xrt_collide_particles(self.ctx, inst))

    # --------------------------------------------------------------- drivers
    def SubstepPlain(self, inst: int) -> None:
        check(self.lib.xrt_substep_plain(self.ctx, inst))

    def FramesPlain(self, n: int) -> None:
        check(self.lib.xrt_frames_plain(self.ctx, n))

    def SubstepBatched(self) -> None:
        check(self.lib.xrt_substep_batched(self.ctx))

    def FramesBatched(self, n: int) -> None:
        check(self.lib.xrt_frames_batched(self.ctx, n))

    def sync(self) -> None:
        check(self.lib.xrt_sync(self.ctx))

    def time_frames(self, batched: bool, n: int) -> float:
        ms = C.c_float()
        check(self.lib.xrt_time_frames(self.ctx, int(batched), n, C.byref(ms)))
        return ms.value

    # ---------------------------------------------------------------- tuning
    def autotune(self, substeps: int = 3) -> dict:
        """Times the candidate launch shapes of the batched substep on this context's
        instances and keeps the fastest (state restored); returns the choice."""
        t = _ffi.XrtTuning()
        check(self.lib.xrt_autotune(self.ctx, substeps, C.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def tuning(self) -> dict:
        t = _ffi.XrtTuning()
        check(self.lib.xrt_get_tuning(self.ctx, C.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def set_tuning(self, persistent_block: int = 1024, persistent_unroll: int = 0,
                   wide_per_thread: int = 1, path: int = 0) -> None:
        t = _ffi.XrtTuning(persistent_block, persistent_unroll, wide_per_thread, path, 0.0, 0.0,
                           0)
        check(self.lib.xrt_set_tuning(self.ctx, C.byref(t)))

    def time_batched_substep(self, n: int) -> float:
        ms = C.c_float()
        check(self.lib.xrt_time_batched_substep(self.ctx, n, C.byref(ms)))
        return ms.value

    # --------------------------------------------------------------- adjoint
    def AdjointBegin(self, n_frames: int, checkpoint_every: int = 10) -> None:
        """Run n_frames frames of every instance, checkpointing for the reverse pass."""
        check(self.lib.xrt_adjoint_begin(self.ctx, n_frames, checkpoint_every))

    def AdjointSeed(self, inst: int, grad_x=None, grad_v=None) -> None:
        gx = None if grad_x is None else np.ascontiguousarray(grad_x, np.float32).reshape(-1)
        gv = None if grad_v is None else np.ascontiguousarray(grad_v, np.float32).reshape(-1)
        check(self.lib.xrt_adjoint_seed(self.ctx, inst, fptr(gx), fptr(gv)))

    def AdjointBackward(self) -> None:
        check(self.lib.xrt_adjoint_backward(self.ctx))

    def AdjointSubstep(self) -> None:
        check(self.lib.xrt_adjoint_substep(self.ctx))

    def GradCompliance(self, inst: int):
        """(d/d stretch compliance per edge, d/d bend compliance per element)."""
        n = self.counts(inst)
        gs = np.zeros(max(n.n_edges, 1), np.float32)
        gb = np.zeros(max(n.n_bends, 1), np.float32)
        check(self.lib.xrt_adjoint_get_grad_compliance(self.ctx, inst, fptr(gs), fptr(gb)))
        return gs[:n.n_edges], gb[:n.n_bends]

    def GradState(self, inst: int):
        """(d/d initial positions, d/d initial velocities), each (nv, 3)."""
        nv = self.counts(inst).nv
        gx = np.zeros(3 * nv, np.float32)
        gv = np.zeros(3 * nv, np.float32)
        check(self.lib.xrt_adjoint_get_grad_state(self.ctx, inst, fptr(gx), fptr(gv)))
        return gx.reshape(-1, 3), gv.reshape(-1, 3)

    def time_repulsion(self, n: int) -> tuple[float, dict]:
        """ms per launch of the batched EE repulsion kernel over n un-captured batched
        substeps (the state advances), and the item / pair counts averaged over them."""
        ms = C.c_float()
        cnt = (C.c_int64 * 4)()
        check(self.lib.xrt_time_repulsion(self.ctx, n, C.byref(ms), cnt))
        return float(ms.value), dict(ee_items=int(cnt[0]), vf_items=int(cnt[1]),
                                     pairs=int(cnt[2]), item_path_instances=int(cnt[3]))

    def time_adjoint(self, n: int) -> tuple[float, float, float]:
        """ms per substep: (forward, taped forward, reverse) over every instance."""
        ms = (C.c_float * 3)()
        check(self.lib.xrt_time_adjoint(self.ctx, n, ms))
        return float(ms[0]), float(ms[1]), float(ms[2])

    def time_particles(self, n: int) -> tuple[float, float, dict]:
        """(collision kernel ms, hash ms, counts) averaged over n batched substeps (the
        state advances); counts: neighbor-list entries after the last hash, particles."""
        ms = np.zeros(2, np.float32)
        cnt = np.zeros(2, np.int64)
        check(self.lib.xrt_time_particles(self.ctx, n, fptr(ms),
                                          cnt.ctypes.data_as(C.POINTER(C.c_int64))))
        return float(ms[0]), float(ms[1]), {"entries": int(cnt[0]), "particles": int(cnt[1])}

    def time_binding(self, n: int) -> tuple[float, float]:
        """ms per launch of the obstacle skinning (every skinned obstacle) and of the
        binding solve (every instance), over n launches each (the state moves)."""
        ms = np.zeros(2, np.float32)
        check(self.lib.xrt_time_binding(self.ctx, n, fptr(ms)))
        return float(ms[0]), float(ms[1])

    def time_fem(self, n: int) -> tuple[float, float]:
        """ms per launch of the wide FEM strain launches (one face color of every instance)
        and of the wide isometric bending launches, over n passes of every color each (the
        state moves)."""
        ms = np.zeros(2, np.float32)
        check(self.lib.xrt_time_fem(self.ctx, n, fptr(ms)))
        return float(ms[0]), float(ms[1])

    def RecordFrames(self, n: int, normals: bool = False):
        """n batched frames with the per-frame cache: positions (n, N, 3) of every instance
        in instance order (and their vertex normals when asked)."""
        N = sum(self.counts(i).nv for i in range(self.num_instances()))
        pos = np.zeros((n, N, 3), np.float32)
        nrm = np.zeros((n, N, 3), np.float32) if normals else None
        check(self.lib.xrt_record_frames(self.ctx, n, fptr(pos.reshape(-1)),
                                         None if nrm is None else fptr(nrm.reshape(-1))))
        return pos, nrm
