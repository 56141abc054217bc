"""ctypes binding of the oracle's quality-mode FEM constraints (fem_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker of the strain and isometric bending kernels.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import oracle as orc

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)
_bound = False


def _lib():
    global _bound
    lib = orc.load()
    if not _bound:
        lib.orc_strain_init.argtypes = [_D, _I, _D, _D]
        lib.orc_strain_init.restype = C.c_int
        lib.orc_strain_energy.argtypes = [_D, _I, C.c_double, _D, _D, _D]
        lib.orc_strain_energy.restype = C.c_double
        lib.orc_strain_solve.argtypes = [_D, _D, _I, _D, _D, C.c_int, C.c_int, _D]
        lib.orc_strain_solve.restype = None
        lib.orc_iso_init.argtypes = [_D, _I, _D, _D]
        lib.orc_iso_init.restype = None
        lib.orc_iso_energy.argtypes = [_D, _I, _D, C.c_double, _D]
        lib.orc_iso_energy.restype = C.c_double
        lib.orc_iso_solve.argtypes = [_D, _D, _I, _D, _D, _D, C.c_int, C.c_int, C.c_int,
                                      C.c_double, C.c_double]
        lib.orc_iso_solve.restype = None
        _bound = True
    return lib


def _d(a):
    return a.ctypes.data_as(_D)


def _i(a):
    return a.ctypes.data_as(_I)


def strain_init(x, tris):
    """(area (nt,), inv (nt, 4) glm order) of every face from rest positions x."""
    x = np.ascontiguousarray(x, np.float64)
    t = np.ascontiguousarray(tris, np.int32)
    area = np.zeros(len(t))
    inv = np.zeros((len(t), 4))
    lib = _lib()
    for f in range(len(t)):
        a = np.zeros(1)
        lib.orc_strain_init(_d(x), _i(t[f].copy()), _d(a), _d(inv[f]))
        area[f] = a[0]
    return area, inv


def strain_energy(x, tri, area, inv, E):
    x = np.ascontiguousarray(x, np.float64)
    g = np.zeros(9)
    e = _lib().orc_strain_energy(_d(x), _i(np.ascontiguousarray(tri, np.int32)), float(area),
                                 _d(np.ascontiguousarray(inv, np.float64)),
                                 _d(np.ascontiguousarray(E, np.float64)), _d(g))
    return e, g.reshape(3, 3)


def strain_solve(x, w, tris, area, inv, beg, end, E):
    assert x.dtype == np.float64 and x.flags.c_contiguous
    _lib().orc_strain_solve(_d(x), _d(np.ascontiguousarray(w, np.float64)),
                            _i(np.ascontiguousarray(tris, np.int32)),
                            _d(np.ascontiguousarray(area, np.float64)),
                            _d(np.ascontiguousarray(inv, np.float64)), beg, end,
                            _d(np.ascontiguousarray(E, np.float64)))


def iso_init(x, bends):
    x = np.ascontiguousarray(x, np.float64)
    b = np.ascontiguousarray(bends, np.int32)
    K = np.zeros((len(b), 4))
    coeff = np.zeros(len(b))
    lib = _lib()
    for e in range(len(b)):
        c = np.zeros(1)
        lib.orc_iso_init(_d(x), _i(b[e].copy()), _d(K[e]), _d(c))
        coeff[e] = c[0]
    return K, coeff


def iso_energy(x, b4, K, coeff):
    x = np.ascontiguousarray(x, np.float64)
    g = np.zeros(12)
    e = _lib().orc_iso_energy(_d(x), _i(np.ascontiguousarray(b4, np.int32)),
                              _d(np.ascontiguousarray(K, np.float64)), float(coeff), _d(g))
    return e, g.reshape(4, 3)


def iso_solve(x, w, bidx, K, coeff, lam, beg, end, it, dt, stiffness):
    assert x.dtype == np.float64 and x.flags.c_contiguous and lam.dtype == np.float64
    _lib().orc_iso_solve(_d(x), _d(np.ascontiguousarray(w, np.float64)),
                         _i(np.ascontiguousarray(bidx, np.int32)),
                         _d(np.ascontiguousarray(K, np.float64)),
                         _d(np.ascontiguousarray(coeff, np.float64)), _d(lam), beg, end, it,
                         float(dt), float(stiffness))
