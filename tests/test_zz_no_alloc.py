"""xrt_bvh_traverse_noalloc (device pair staging in a fixed array, no allocation) against
xrt_bvh_traverse through the C ABI: identical pairs (same layout) and counts on two calls in
a row, the second on a smaller mesh; a pair count beyond the fixed staging maximum is
refused with XRT_ERR_CAPACITY and the count."""
import ctypes as C

import numpy as np
import pytest

from xrtailor_amd import _ffi, scenes
from xrtailor_amd._ffi import XRT_ERR_CAPACITY, XrtError, check, iptr
from xrtailor_amd.bvh import BVH, Traverse

pytestmark = pytest.mark.gpu

MAX_PAIRS = 1 << 21


def _traverse_noalloc(lhs, rhs, thickness, cap=None):
    lib = _ffi.load()
    n = C.c_int64()
    if cap is None:
        rc = lib.xrt_bvh_traverse_noalloc(lhs.h, rhs.h, thickness, None, 0, C.byref(n))
        if rc not in (_ffi.XRT_OK, XRT_ERR_CAPACITY):
            check(rc)
        cap = n.value
    pairs = np.zeros(2 * max(cap, 1), np.int32)
    check(lib.xrt_bvh_traverse_noalloc(lhs.h, rhs.h, thickness, iptr(pairs), cap, C.byref(n)))
    return pairs[: 2 * n.value].reshape(-1, 2)


def _sheet(n, side, seed):
    rng = scenes.Rng(seed)
    pos, tris = scenes.grid(n, side)
    pos = pos + 1e-3 * rng.normal(pos.size).reshape(-1, 3)
    return pos.astype(np.float32), np.asarray(tris, np.int32)


def test_noalloc_traverse_matches_traverse_twice_second_smaller():
    sphere = BVH()
    spos, stris = scenes.icosphere(4, 0.5, (0.0, -0.4, 0.0))
    sphere.Build(spos.astype(np.float32), stris)
    for n, th in ((128, 0.004), (40, 0.002)):       # second call: a smaller mesh
        print(f"case noalloc traverse sheet {n}x{n} thickness {th}", flush=True)
        cloth = BVH()
        cloth.Build(*_sheet(n, 1.0, seed=n))
        for lhs, rhs in ((cloth, cloth), (cloth, sphere)):
            ref = Traverse(lhs, rhs, th)
            got = _traverse_noalloc(lhs, rhs, th)
            print(f"  {len(ref)} pairs", flush=True)
            assert len(ref) > 0
            np.testing.assert_array_equal(got, ref)
        cloth.close()


def test_noalloc_traverse_refuses_beyond_staging_maximum():
    print("case noalloc traverse beyond the staging maximum", flush=True)
    cloth = BVH()
    cloth.Build(*_sheet(256, 2.0, seed=7))
    lib = _ffi.load()
    n = C.c_int64()
    rc = lib.xrt_bvh_traverse(cloth.h, cloth.h, 0.012, None, 0, C.byref(n))  # the count
    assert rc == XRT_ERR_CAPACITY and n.value > MAX_PAIRS, (rc, n.value)
    pairs = np.zeros(2 * n.value, np.int32)         # room for all: a refusal writes nothing
    pairs[:] = -7
    rc = lib.xrt_bvh_traverse_noalloc(cloth.h, cloth.h, 0.012, iptr(pairs), n.value, C.byref(n))
    print(f"  rc {rc}, count {n.value}", flush=True)
    assert rc == XRT_ERR_CAPACITY and n.value > MAX_PAIRS
    assert (pairs == -7).all()
    assert b"2,097,152" in lib.xrt_last_error()
    small = BVH()                                  # still usable afterwards
    small.Build(*_sheet(64, 1.0, seed=8))
    got = _traverse_noalloc(small, small, 0.002)
    np.testing.assert_array_equal(got, Traverse(small, small, 0.002))
    assert len(got) > 0
