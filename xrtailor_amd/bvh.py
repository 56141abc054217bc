"""Python mirror of the LBVH C ABI (the reference's BVH class and Traverse,
xrtailor/include/xrtailor/physics/broad_phase/lbvh/BVH.cuh:77-152)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi
from ._ffi import XRT_OK, XrtError, check, fptr, iptr

MAX_HITS = 256  # LBVH_MAX_BUFFER_SIZE


def _f3(a):
    return None if a is None else np.ascontiguousarray(a, np.float32).reshape(-1)


class BVH:
    def __init__(self, device: int = 0):
        self.lib = _ffi.load()
        h = C.c_void_p()
        check(self.lib.xrt_bvh_create(C.byref(h), device))
        self.h = h
        self.nv = 0

    def close(self):
        if self.h:
            self.lib.xrt_bvh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # BVH::UpdateData + BVH::Construct
    def Build(self, x0, tris, tol=0.0, n=None, x=None, ccd=False, scr=0.0):
        x0, x, n = _f3(x0), _f3(x), _f3(n)
        t = np.ascontiguousarray(tris, np.int32).reshape(-1)
        self.nv = x0.size // 3
        check(self.lib.xrt_bvh_build(self.h, fptr(x0), fptr(x), fptr(n), self.nv, iptr(t),
                                     t.size // 3, tol, int(ccd), scr))

    def RefitMidstep(self, x1, n1, tol):
        x1, n1 = _f3(x1), _f3(n1)
        check(self.lib.xrt_bvh_refit_midstep(self.h, fptr(x1), fptr(n1), tol))

    def Update(self, x, x0=None, ccd=False):
        x, x0 = _f3(x), _f3(x0)
        check(self.lib.xrt_bvh_update(self.h, fptr(x), fptr(x0), int(ccd)))

    def NumObjects(self) -> int:
        n = C.c_int32()
        check(self.lib.xrt_bvh_num_objects(self.h, C.byref(n)))
        return n.value

    def host(self) -> dict:
        """mortons, order, nodes (2F-1,4), aabbs (2F-1,6), i_escape (F), e_escape (F), root (6)"""
        f = self.NumObjects()
        nn = max(2 * f - 1, 0)
        m = np.zeros(f, np.uint32)
        o = np.zeros(f, np.int32)
        nodes = np.zeros(4 * nn, np.int32)
        aabbs = np.zeros(6 * nn, np.float32)
        ie = np.zeros(f, np.int32)
        ee = np.zeros(f, np.int32)
        root = np.zeros(6, np.float32)
        check(self.lib.xrt_bvh_get(self.h, m.ctypes.data_as(C.POINTER(C.c_uint32)), iptr(o),
                                   iptr(nodes), fptr(aabbs), iptr(ie), iptr(ee), fptr(root)))
        return dict(morton=m, order=o, nodes=nodes.reshape(-1, 4), aabbs=aabbs.reshape(-1, 6),
                    i_escape=ie, e_escape=ee, root=root)

    def Query(self, p0, p1=None):
        """QueryDeviceStackless: list of hit arrays per query (strict point / segment)."""
        p0, p1 = _f3(p0), _f3(p1)
        n = p0.size // 3
        hits = np.zeros(n * MAX_HITS, np.int32)
        cnt = np.zeros(n, np.int32)
        check(self.lib.xrt_bvh_query(self.h, fptr(p0), fptr(p1), n, iptr(hits), iptr(cnt)))
        hits = hits.reshape(n, MAX_HITS)
        return [hits[i, :cnt[i]].copy() for i in range(n)]


def Traverse(lhs: BVH, rhs: BVH, thickness: float) -> np.ndarray:
    """Pairs (lhs face, rhs face) in the reference's layout (lhs leaves in Morton order,
    each leaf's pairs in reverse traversal order)."""
    lib = _ffi.load()
    n = C.c_int64()
    rc = lib.xrt_bvh_traverse(lhs.h, rhs.h, thickness, None, 0, C.byref(n))
    if rc not in (XRT_OK, -4):
        check(rc)
    pairs = np.zeros(2 * max(n.value, 1), np.int32)
    check(lib.xrt_bvh_traverse(lhs.h, rhs.h, thickness, iptr(pairs), n.value, C.byref(n)))
    return pairs[: 2 * n.value].reshape(-1, 2)


def vertex_normals(bvh: BVH, x, tris) -> np.ndarray:
    x = _f3(x)
    t = np.ascontiguousarray(tris, np.int32).reshape(-1)
    out = np.zeros_like(x)
    check(_ffi.load().xrt_vertex_normals(bvh.h, fptr(x), x.size // 3, iptr(t), t.size // 3,
                                         fptr(out)))
    return out.reshape(-1, 3)


__all__ = ["BVH", "Traverse", "vertex_normals", "XrtError", "MAX_HITS"]
