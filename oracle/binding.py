"""ctypes binding of the oracle's obstacle skinning, nearest-face query and binding
(binding_oracle.c, bvh_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker of the steps either side of the substep in the
garment pipeline (LBS, QueryNearestBarycentricClothToObstacle, BindingConstraint).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import oracle as orc

_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int32)

_bound = False


def _lib():
    global _bound
    lib = orc._bind_bvh()
    if not _bound:
        lib.orc_tri_nearest.argtypes = [_F, _F, _F, _F, _F]
        lib.orc_tri_nearest.restype = C.c_float
        lib.orc_bvh_nearest.argtypes = [C.c_void_p, _F, _F, _F]
        lib.orc_bvh_nearest.restype = C.c_int
        lib.orc_lbs.argtypes = [_D, C.c_int, _I, _I, _D, _D, _D]
        lib.orc_lbs.restype = None
        lib.orc_binding_generate.argtypes = [_D, _D, _I, _I, _D, _D, C.c_int, _D]
        lib.orc_binding_generate.restype = None
        lib.orc_binding_solve.argtypes = [_D, _D, _D, _I, _I, _D, _D, _D, C.c_int]
        lib.orc_binding_solve.restype = None
        _bound = True
    return lib


def _d(a):
    return a.ctypes.data_as(_D)


def _f(a):
    return a.ctypes.data_as(_F)


def _i(a):
    return a.ctypes.data_as(_I)


def lbs(rest, ptr, joint, weight, T):
    """Linear blend skinning in FP64: rest (nv, 3), weights CSR, T (n_joints, 12)."""
    rest = np.ascontiguousarray(rest, np.float64)
    ptr = np.ascontiguousarray(ptr, np.int32)
    joint = np.ascontiguousarray(joint, np.int32)
    weight = np.ascontiguousarray(weight, np.float64)
    T = np.ascontiguousarray(T, np.float64).reshape(-1)
    out = np.zeros_like(rest)
    _lib().orc_lbs(_d(rest), len(rest), _i(ptr), _i(joint), _d(weight), _d(T), _d(out))
    return out


def tri_nearest(p, a, b, c):
    """(distance, uvw) of the closest point of triangle abc to p, FP32."""
    vals = [np.ascontiguousarray(v, np.float32) for v in (p, a, b, c)]
    uvw = np.zeros(3, np.float32)
    d = _lib().orc_tri_nearest(*[_f(v) for v in vals], _f(uvw))
    return float(d), uvw


def nearest(bvh: orc.Bvh, points):
    """QueryDevice over a built oracle tree: faces (-1 none), uvw (n, 3), distances."""
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    lib = _lib()
    faces = np.zeros(len(pts), np.int32)
    uvw = np.zeros((len(pts), 3), np.float32)
    dist = np.zeros(len(pts), np.float32)
    for k in range(len(pts)):
        w = np.zeros(3, np.float32)
        dk = np.zeros(1, np.float32)
        faces[k] = lib.orc_bvh_nearest(bvh.h, _f(pts[k].copy()), _f(w), _f(dk))
        uvw[k], dist[k] = w, dk[0]
    return faces, uvw, dist


def generate(cloth_x0, obs_x0, vert, tri3, uvw, dist_in):
    """BindingConstraint::Generate: rest distance L per binding (FP64)."""
    cx = np.ascontiguousarray(cloth_x0, np.float64)
    ox = np.ascontiguousarray(obs_x0, np.float64)
    vert = np.ascontiguousarray(vert, np.int32)
    tri3 = np.ascontiguousarray(tri3, np.int32)
    uvw = np.ascontiguousarray(uvw, np.float64)
    dist_in = np.ascontiguousarray(dist_in, np.float64)
    L = np.zeros(len(vert))
    _lib().orc_binding_generate(_d(cx), _d(ox), _i(vert), _i(tri3), _d(uvw), _d(dist_in),
                                len(vert), _d(L))
    return L


def solve(x, w, obs_x, vert, tri3, uvw, L, stiffness):
    """BindingConstraint::Solve in place on x (nv, 3) float64 (C-contiguous)."""
    assert x.dtype == np.float64 and x.flags.c_contiguous
    w = np.ascontiguousarray(w, np.float64)
    ox = np.ascontiguousarray(obs_x, np.float64)
    vert = np.ascontiguousarray(vert, np.int32)
    tri3 = np.ascontiguousarray(tri3, np.int32)
    uvw = np.ascontiguousarray(uvw, np.float64)
    L = np.ascontiguousarray(L, np.float64)
    st = np.ascontiguousarray(stiffness, np.float64)
    _lib().orc_binding_solve(_d(x), _d(w), _d(ox), _i(vert), _i(tri3), _d(uvw), _d(L), _d(st),
                             len(vert))
