"""ctypes binding of the oracle's particle self-collision (particle_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker of the swift-mode spatial hash and CollideParticles.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import oracle as orc

_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int32)
_U = C.POINTER(C.c_uint32)

_bound = False


def _lib():
    global _bound
    lib = orc.load()
    if not _bound:
        lib.orc_hash_coords.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32]
        lib.orc_hash_coords.restype = C.c_int32
        lib.orc_particle_neighbors.argtypes = [_F, _F, C.c_int, C.c_float, C.c_float, C.c_int,
                                               _U, _U, _U, _U, _U]
        lib.orc_particle_neighbors.restype = None
        lib.orc_collide_particles.argtypes = [_D, _D, _D, C.c_int, _U, C.c_int, C.c_double,
                                              C.c_double, _D, _I]
        lib.orc_collide_particles.restype = None
        _bound = True
    return lib


def hash_coords(x: int, y: int, z: int, table: int) -> int:
    return int(_lib().orc_hash_coords(x, y, z, table))


def particle_diameter(rest_pos, tris, factor: float) -> np.float32:
    """First triangle's first edge length x factor, in FP32 (ClothObject.hpp:36-37)."""
    p = np.asarray(rest_pos, np.float32)
    t = np.asarray(tris).reshape(-1)
    d = p[t[0]] - p[t[1]]
    s = np.float32(d[0] * d[0]) + np.float32(d[1] * d[1])
    s = np.float32(s + np.float32(d[2] * d[2]))
    return np.float32(np.sqrt(s, dtype=np.float32) * np.float32(factor))


def neighbors(x, orig, diameter, cell_size, max_nb):
    """SpatialHashGPU::Hash from FP32 positions: dict of neighbors (max_nb, n), the sorted
    particle_hash / particle_index, cell_start / cell_end (2n)."""
    x = np.ascontiguousarray(x, np.float32).reshape(-1)
    o = np.ascontiguousarray(orig, np.float32).reshape(-1)
    n = x.size // 3
    nb = np.zeros(n * max_nb, np.uint32)
    sh = np.zeros(n, np.uint32)
    si = np.zeros(n, np.uint32)
    cs = np.zeros(2 * n, np.uint32)
    ce = np.zeros(2 * n, np.uint32)
    _lib().orc_particle_neighbors(x.ctypes.data_as(_F), o.ctypes.data_as(_F), n,
                                  float(diameter), float(cell_size), max_nb,
                                  nb.ctypes.data_as(_U), sh.ctypes.data_as(_U),
                                  si.ctypes.data_as(_U), cs.ctypes.data_as(_U),
                                  ce.ctypes.data_as(_U))
    return {"neighbors": nb.reshape(max_nb, n), "particle_hash": sh,
            "particle_index": si.astype(np.int32), "cell_start": cs, "cell_end": ce}


def collide(x, x0, w, nb, diameter, mu):
    """CollideParticles: (delta (n, 3), count (n)) in FP64; nb as returned by neighbors()."""
    x = np.ascontiguousarray(x, np.float64).reshape(-1)
    x0 = np.ascontiguousarray(x0, np.float64).reshape(-1)
    w = np.ascontiguousarray(w, np.float64)
    nbf = np.ascontiguousarray(nb, np.uint32)
    max_nb, n = nbf.shape
    delta = np.zeros(3 * n)
    count = np.zeros(n, np.int32)
    _lib().orc_collide_particles(x.ctypes.data_as(_D), x0.ctypes.data_as(_D),
                                 w.ctypes.data_as(_D), n, nbf.reshape(-1).ctypes.data_as(_U),
                                 max_nb, float(diameter), float(mu), delta.ctypes.data_as(_D),
                                 count.ctypes.data_as(_I))
    return delta.reshape(n, 3), count
