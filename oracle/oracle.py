"""ctypes binding of the serial CPU oracle (oracle/_build/libxrt_oracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker / timed CPU baseline, never by the product path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_build" / "libxrt_oracle.so"

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)


class OrcParams(C.Structure):
    _fields_ = [
        ("gravity", C.c_double * 3),
        ("damping", C.c_double),
        ("relaxation_factor", C.c_double),
        ("long_range_stretchiness", C.c_double),
        ("ground_height", C.c_double),
        ("ground_friction", C.c_double),
        ("num_iterations", C.c_int32),
        ("solve_bending", C.c_int32),
        ("enable_ground", C.c_int32),
        ("bend_model", C.c_int32),
    ]


class OrcCloth(C.Structure):
    _fields_ = [
        ("nv", C.c_int32),
        ("x0", _D), ("v", _D), ("x", _D), ("inv_mass", _D),
        ("ns", C.c_int32), ("s_idx", _I), ("s_rest", _D), ("s_compliance", _D), ("s_lambda", _D),
        ("s_ncolors", C.c_int32), ("s_color_off", _I),
        ("nb", C.c_int32), ("b_idx", _I), ("b_theta", _D), ("b_compliance", _D),
        ("b_ncolors", C.c_int32), ("b_color_off", _I),
        ("nslots", C.c_int32), ("lra_slot_vert", _I), ("lra_ptr", _I), ("lra_slot", _I),
        ("lra_rest", _D),
    ]


_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} missing: run `python -m xrtailor_amd.build`")
        lib = C.CDLL(str(LIB_PATH))
        lib.orc_build_edges.argtypes = [_I, C.c_int, _I, C.c_int]
        lib.orc_build_bends.argtypes = [_I, C.c_int, _I, C.c_int, _I, C.c_int]
        lib.orc_rest_lengths.argtypes = [_D, _I, C.c_int, _D]
        lib.orc_rest_lengths.restype = None
        lib.orc_rest_angles.argtypes = [_D, _I, C.c_int, _D]
        lib.orc_rest_angles.restype = None
        lib.orc_color_constraints.argtypes = [_I, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_int,
                                              _I]
        for fn in ("orc_predict", "orc_lra", "orc_ground", "orc_finalize", "orc_substep"):
            getattr(lib, fn).restype = None
        lib.orc_predict.argtypes = [C.POINTER(OrcCloth), C.POINTER(OrcParams), C.c_double]
        lib.orc_finalize.argtypes = [C.POINTER(OrcCloth), C.POINTER(OrcParams), C.c_double]
        lib.orc_substep.argtypes = [C.POINTER(OrcCloth), C.POINTER(OrcParams), C.c_double]
        lib.orc_lra.argtypes = [C.POINTER(OrcCloth), C.POINTER(OrcParams)]
        lib.orc_ground.argtypes = [C.POINTER(OrcCloth), C.POINTER(OrcParams)]
        lib.orc_stretch_color.argtypes = [C.POINTER(OrcCloth), C.c_int, C.c_double, C.c_int]
        lib.orc_stretch_color.restype = None
        lib.orc_bend_color.argtypes = [C.POINTER(OrcCloth), C.POINTER(OrcParams), C.c_int,
                                       C.c_double]
        lib.orc_bend_color.restype = None
        lib.orc_geodesic.argtypes = [_D, C.c_int, _I, C.c_int, C.c_int, _D]
        lib.orc_geodesic.restype = None
        lib.orc_bend_probe.argtypes = [_D, _D, _D]
        lib.orc_splitmix64.argtypes = [C.POINTER(C.c_uint64)]
        lib.orc_splitmix64.restype = C.c_uint64
        _lib = lib
    return _lib


def _d(a):
    return a.ctypes.data_as(_D)


def _i(a):
    return a.ctypes.data_as(_I)


# ------------------------------------------------------------------- topology
def build_edges(tris: np.ndarray) -> np.ndarray:
    t = np.ascontiguousarray(tris, np.int32).reshape(-1)
    out = np.zeros(2 * t.size, np.int32)
    ne = load().orc_build_edges(_i(t), t.size // 3, _i(out), t.size)
    assert ne >= 0
    return out[: 2 * ne].reshape(-1, 2)


def build_bends(tris: np.ndarray, edges: np.ndarray) -> np.ndarray:
    t = np.ascontiguousarray(tris, np.int32).reshape(-1)
    e = np.ascontiguousarray(edges, np.int32).reshape(-1)
    cap = 3 * t.size + 16
    out = np.zeros(4 * cap, np.int32)
    nb = load().orc_build_bends(_i(t), t.size // 3, _i(e), e.size // 2, _i(out), cap)
    assert nb >= 0
    return out[: 4 * nb].reshape(-1, 4)


def rest_lengths(x: np.ndarray, edges: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64).reshape(-1)
    e = np.ascontiguousarray(edges, np.int32).reshape(-1)
    out = np.zeros(e.size // 2)
    load().orc_rest_lengths(_d(x), _i(e), e.size // 2, _d(out))
    return out


def rest_angles(x: np.ndarray, bends: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64).reshape(-1)
    b = np.ascontiguousarray(bends, np.int32).reshape(-1)
    out = np.zeros(b.size // 4)
    load().orc_rest_angles(_d(x), _i(b), b.size // 4, _d(out))
    return out


def color(cons: np.ndarray, nv: int, seed: int, mode: int = 0):
    """mode 0: first fit in index order; 1: seeded palette coloring."""
    c = np.ascontiguousarray(cons, np.int32)
    arity = c.shape[1]
    out = np.zeros(c.shape[0], np.int32)
    n = load().orc_color_constraints(_i(c.reshape(-1)), arity, c.shape[0], nv, C.c_uint64(seed),
                                     mode, _i(out))
    return out, n


def geodesic(x: np.ndarray, edges: np.ndarray, src: int) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64).reshape(-1)
    e = np.ascontiguousarray(edges, np.int32).reshape(-1)
    out = np.zeros(x.size // 3)
    load().orc_geodesic(_d(x), x.size // 3, _i(e), e.size // 2, src, _d(out))
    return out


def bend_probe(x4: np.ndarray):
    """(d_raw, q[4,3], active) of one hinge (v0, v1, v2, v3)."""
    x = np.ascontiguousarray(x4, np.float64).reshape(-1)
    d = np.zeros(1)
    q = np.zeros(12)
    ok = load().orc_bend_probe(_d(x), _d(d), _d(q))
    return float(d[0]), q.reshape(4, 3), bool(ok)


def color_sort(colors: np.ndarray, ncol: int):
    order = np.argsort(colors, kind="stable").astype(np.int32)
    off = np.zeros(ncol + 1, np.int32)
    off[1:] = np.cumsum(np.bincount(colors, minlength=ncol))
    return order, off


# ------------------------------------------------------------------ dynamics
class Cloth:
    """Oracle state of one cloth instance in the solver's shared layout (constraints
    in color order, per-target LRA CSR)."""

    def __init__(self, x0, v, inv_mass, edges, rest, s_comp, s_colors, s_ncol,
                 bends, theta, b_comp, b_colors, b_ncol,
                 slot_vert=None, lra_ptr=None, lra_slot=None, lra_rest=None):
        self.nv = len(inv_mass)
        self.x0 = np.ascontiguousarray(x0, np.float64).reshape(-1).copy()
        self.v = np.ascontiguousarray(v, np.float64).reshape(-1).copy()
        self.x = self.x0.copy()
        self.w = np.ascontiguousarray(inv_mass, np.float64)
        so, self.s_off = color_sort(np.asarray(s_colors), s_ncol)
        bo, self.b_off = color_sort(np.asarray(b_colors), b_ncol)
        self.s_order, self.b_order = so, bo
        self.s_idx = np.ascontiguousarray(np.asarray(edges, np.int32)[so].reshape(-1))
        self.s_rest = np.ascontiguousarray(np.asarray(rest, np.float64)[so])
        self.s_comp = np.ascontiguousarray(np.asarray(s_comp, np.float64)[so])
        self.s_lam = np.zeros(len(so))
        self.b_idx = np.ascontiguousarray(np.asarray(bends, np.int32)[bo].reshape(-1))
        self.b_theta = np.ascontiguousarray(np.asarray(theta, np.float64)[bo])
        self.b_comp = np.ascontiguousarray(np.asarray(b_comp, np.float64)[bo])
        self.s_ncol, self.b_ncol = s_ncol, b_ncol
        self.slot_vert = np.ascontiguousarray(slot_vert if slot_vert is not None else np.zeros(0), np.int32)
        self.lra_ptr = np.ascontiguousarray(lra_ptr if lra_ptr is not None else np.zeros(self.nv + 1), np.int32)
        self.lra_slot = np.ascontiguousarray(lra_slot if lra_slot is not None else np.zeros(0), np.int32)
        self.lra_rest = np.ascontiguousarray(lra_rest if lra_rest is not None else np.zeros(0), np.float64)
        self._keep = []
        self.s = OrcCloth(
            self.nv, _d(self.x0), _d(self.v), _d(self.x), _d(self.w),
            len(so), _i(self.s_idx), _d(self.s_rest), _d(self.s_comp), _d(self.s_lam),
            s_ncol, _i(self.s_off),
            len(bo), _i(self.b_idx), _d(self.b_theta), _d(self.b_comp), b_ncol, _i(self.b_off),
            len(self.slot_vert), _i(self.slot_vert), _i(self.lra_ptr), _i(self.lra_slot),
            _d(self.lra_rest))

    def substep(self, params: OrcParams, dt: float) -> None:
        load().orc_substep(C.byref(self.s), C.byref(params), dt)

    def state(self):
        return self.x0.reshape(-1, 3), self.v.reshape(-1, 3)


def params_from(xp) -> tuple[OrcParams, float]:
    """Oracle parameters (and substep dt) from a solver XrtParams."""
    p = OrcParams()
    for k in range(3):
        p.gravity[k] = xp.gravity[k]
    p.damping = xp.damping
    p.relaxation_factor = xp.relaxation_factor
    p.long_range_stretchiness = xp.long_range_stretchiness
    p.ground_height = xp.ground_height
    p.ground_friction = xp.ground_friction
    p.num_iterations = xp.num_iterations
    p.solve_bending = xp.solve_bending
    p.enable_ground = xp.enable_ground
    p.bend_model = xp.bend_model
    dt = float(np.float32(xp.frame_time) / np.float32(xp.num_substeps))
    return p, dt


# ---------------------------------------------------------------------- LBVH
_F = C.POINTER(C.c_float)
_U = C.POINTER(C.c_uint32)
_bvh_bound = False


def _bind_bvh():
    global _bvh_bound
    lib = load()
    if _bvh_bound:
        return lib
    lib.orc_vertex_normals.argtypes = [_F, C.c_int, _I, C.c_int, _F]
    lib.orc_vertex_normals.restype = None
    lib.orc_expand_bits.argtypes = [C.c_uint32]
    lib.orc_expand_bits.restype = C.c_uint32
    lib.orc_morton_code.argtypes = [C.c_float, C.c_float, C.c_float]
    lib.orc_morton_code.restype = C.c_uint32
    lib.orc_bvh_create.restype = C.c_void_p
    lib.orc_bvh_destroy.argtypes = [C.c_void_p]
    lib.orc_bvh_destroy.restype = None
    lib.orc_bvh_build.argtypes = [C.c_void_p, _F, _F, _F, C.c_int, _I, C.c_int, C.c_float, C.c_int,
                                  C.c_float]
    lib.orc_bvh_num_objects.argtypes = [C.c_void_p]
    lib.orc_bvh_get.argtypes = [C.c_void_p, _U, _I, _I, _F, _I, _F]
    lib.orc_bvh_get.restype = None
    lib.orc_bvh_refit_midstep.argtypes = [C.c_void_p, _F, _F, C.c_float]
    lib.orc_bvh_refit_midstep.restype = None
    lib.orc_bvh_update.argtypes = [C.c_void_p, _F, _F, C.c_int]
    lib.orc_bvh_update.restype = None
    lib.orc_bvh_query.argtypes = [C.c_void_p, _F, _F, _I, C.c_int]
    lib.orc_overlap_segment.argtypes = [_F, _F, _F]
    lib.orc_bvh_traverse.argtypes = [C.c_void_p, C.c_void_p, C.c_float, _I, C.c_long]
    lib.orc_bvh_traverse.restype = C.c_long
    _bvh_bound = True
    return lib


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, np.float32).reshape(-1)


def _fp(a):
    return None if a is None else a.ctypes.data_as(_F)


def vertex_normals(x, tris) -> np.ndarray:
    lib = _bind_bvh()
    x = _f32(x)
    t = np.ascontiguousarray(tris, np.int32).reshape(-1)
    out = np.zeros_like(x)
    lib.orc_vertex_normals(_fp(x), x.size // 3, _i(t), t.size // 3, _fp(out))
    return out.reshape(-1, 3)


def expand_bits(v: int) -> int:
    return int(_bind_bvh().orc_expand_bits(v))


def morton_code(x: float, y: float, z: float) -> int:
    return int(_bind_bvh().orc_morton_code(x, y, z))


def overlap_segment(box6, p0, p1) -> bool:
    b, a, c = _f32(box6), _f32(p0), _f32(p1)
    return bool(_bind_bvh().orc_overlap_segment(_fp(b), _fp(a), _fp(c)))


class Bvh:
    """FP32-strict oracle LBVH with the same surface as xrtailor_amd.bvh.BVH."""

    def __init__(self):
        self.lib = _bind_bvh()
        self.h = C.c_void_p(self.lib.orc_bvh_create())

    def __del__(self):
        try:
            self.lib.orc_bvh_destroy(self.h)
        except Exception:
            pass

    def Build(self, x0, tris, tol=0.0, n=None, x=None, ccd=False, scr=0.0):
        x0, x, n = _f32(x0), _f32(x), _f32(n)
        t = np.ascontiguousarray(tris, np.int32).reshape(-1)
        self._keep = (x0, x, n, t)
        self.lib.orc_bvh_build(self.h, _fp(x0), _fp(x), _fp(n), x0.size // 3, _i(t), t.size // 3,
                               tol, int(ccd), scr)

    def RefitMidstep(self, x1, n1, tol):
        x1, n1 = _f32(x1), _f32(n1)
        self.lib.orc_bvh_refit_midstep(self.h, _fp(x1), _fp(n1), tol)

    def Update(self, x, x0=None, ccd=False):
        x, x0 = _f32(x), _f32(x0)
        self.lib.orc_bvh_update(self.h, _fp(x), _fp(x0), int(ccd))

    def NumObjects(self) -> int:
        return int(self.lib.orc_bvh_num_objects(self.h))

    def host(self) -> dict:
        f = self.NumObjects()
        nn = max(2 * f - 1, 0)
        m = np.zeros(f, np.uint32)
        o = np.zeros(f, np.int32)
        nodes = np.zeros(4 * nn, np.int32)
        box = np.zeros(6 * nn, np.float32)
        esc = np.zeros(max(nn, 1), np.int32)
        root = np.zeros(6, np.float32)
        self.lib.orc_bvh_get(self.h, m.ctypes.data_as(_U), _i(o), _i(nodes), _fp(box), _i(esc),
                             _fp(root))
        ni = f - 1
        ie = np.full(f, -1, np.int32)
        ie[:ni] = esc[:ni]
        return dict(morton=m, order=o, nodes=nodes.reshape(-1, 4), aabbs=box.reshape(-1, 6),
                    i_escape=ie, e_escape=esc[ni:nn].copy(), root=root, escape=esc[:nn])

    def Query(self, p0, p1=None, cap=256):
        p0, p1 = _f32(p0), _f32(p1)
        n = p0.size // 3
        out = []
        buf = np.zeros(cap, np.int32)
        for q in range(n):
            a = p0[3 * q:3 * q + 3].copy()
            b = None if p1 is None else p1[3 * q:3 * q + 3].copy()
            k = self.lib.orc_bvh_query(self.h, _fp(a), _fp(b), _i(buf), cap)
            out.append(buf[:k].copy())
        return out


def traverse(lhs: Bvh, rhs: Bvh, thickness: float) -> np.ndarray:
    lib = _bind_bvh()
    n = lib.orc_bvh_traverse(lhs.h, rhs.h, thickness, None, 0)
    pairs = np.zeros(2 * max(n, 1), np.int32)
    lib.orc_bvh_traverse(lhs.h, rhs.h, thickness, _i(pairs), n)
    return pairs[: 2 * n].reshape(-1, 2)
