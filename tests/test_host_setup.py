"""CPU tests of the product's host logic and C ABI surface (no device needed):
topology and coloring equal the oracle's, the library exports every declared symbol,
and device-free error paths return status codes."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as orc
from xrtailor_amd import _ffi, scenes

ROOT = Path(__file__).resolve().parent.parent


def _lib_edges(tris):
    lib = _ffi.load()
    t = np.ascontiguousarray(tris, np.int32).reshape(-1)
    out = np.zeros(2 * t.size, np.int32)
    ne = C.c_int32()
    _ffi.check(lib.xrt_build_edges(_ffi.iptr(t), t.size // 3, _ffi.iptr(out), t.size, C.byref(ne)))
    return out[: 2 * ne.value].reshape(-1, 2)


def _lib_bends(tris, edges):
    lib = _ffi.load()
    t = np.ascontiguousarray(tris, np.int32).reshape(-1)
    e = np.ascontiguousarray(edges, np.int32).reshape(-1)
    cap = t.size * 3
    out = np.zeros(4 * cap, np.int32)
    nb = C.c_int32()
    _ffi.check(lib.xrt_build_bends(_ffi.iptr(t), t.size // 3, _ffi.iptr(e), e.size // 2,
                                   _ffi.iptr(out), cap, C.byref(nb)))
    return out[: 4 * nb.value].reshape(-1, 4)


def _lib_color(cons, nv, seed, mode=0):
    lib = _ffi.load()
    c = np.ascontiguousarray(cons, np.int32)
    out = np.zeros(c.shape[0], np.int32)
    n = C.c_int32()
    _ffi.check(lib.xrt_color_constraints(_ffi.iptr(c.reshape(-1)), c.shape[1], c.shape[0], nv,
                                         C.c_uint64(seed), mode, _ffi.iptr(out), C.byref(n)))
    return out, n.value


def test_abi_exports_every_declared_symbol():
    header = (ROOT / "include" / "xrtailor_amd.h").read_text()
    declared = set(re.findall(r"^(?:int|void|const char\*)\s+(xrt_\w+)\(", header, re.M))
    assert len(declared) > 30
    lib = C.CDLL(str(_ffi.LIB_PATH))
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    # the ctypes mirror binds exactly the declared set
    assert set(_ffi.SIGNATURES) == declared


def test_abi_version_and_defaults():
    lib = _ffi.load()
    assert lib.xrt_abi_version() == 5
    p = _ffi.XrtParams()
    lib.xrt_default_params(C.byref(p))
    assert p.num_substeps == 10 and p.num_iterations == 10
    assert p.damping == pytest.approx(0.98) and p.relaxation_factor == pytest.approx(0.25)
    assert p.bend_model == 0
    assert p.pair_capacity_per_face == 64
    assert p.repulsion_vf_items_per_face == 64 and p.repulsion_ee_items_per_face == 128
    # ParticleCollisionSettings defaults (xrtailor/include/xrtailor/config/Base.hpp:59-66)
    assert p.enable_particle_collision == 0 and p.max_num_neighbors == 64
    assert p.inter_leaved_hash == 3
    assert p.particle_diameter == pytest.approx(1.3) and p.hash_cell_size == pytest.approx(1.5)
    # FEMFabricSettings defaults (xrtailor/include/xrtailor/config/Base.hpp:47-57)
    assert p.constraint_model == 0
    assert (p.fem_xx_stiffness, p.fem_yy_stiffness, p.fem_xy_stiffness) == (1.0, 1.0, 1.0)
    assert p.fem_xy_poisson_ratio == pytest.approx(0.3) and p.fem_yx_poisson_ratio == pytest.approx(0.3)
    assert p.fem_bending_stiffness == 1.0
    assert p.contact_friction == 0.0             # the reference applies none on mesh contacts
    assert C.sizeof(p) == 4 * 38  # 38 four-byte fields
    assert C.sizeof(_ffi.XrtTuning) == 4 * 7


@pytest.mark.parametrize("n", [2, 5, 64])
def test_grid_topology_counts_and_match_oracle(n):
    pos, tris = scenes.grid(n, 1.0)
    e_lib = _lib_edges(tris)
    e_orc = orc.build_edges(tris)
    np.testing.assert_array_equal(e_lib, e_orc)
    assert len(e_lib) == 3 * (n - 1) ** 2 + 2 * (n - 1)
    b_lib = _lib_bends(tris, e_lib)
    np.testing.assert_array_equal(b_lib, orc.build_bends(tris, e_orc))
    assert len(b_lib) == len(e_lib) - 4 * (n - 1)  # one per interior edge
    # every element's (v0, v1) runs in the first face's winding and v2 / v3 sit opposite
    faces = {tuple(sorted(t)): t for t in tris.tolist()}
    for v0, v1, v2, v3 in b_lib[:200]:
        f1 = faces[tuple(sorted((v0, v1, v2)))]
        k = f1.index(v0)
        assert f1[(k + 1) % 3] == v1
        assert tuple(sorted((v0, v1, v3))) in faces


def test_icosphere_topology_closed_manifold():
    pos, tris = scenes.icosphere(3, 0.5)
    e = _lib_edges(tris)
    b = _lib_bends(tris, e)
    assert len(pos) - len(e) + len(tris) == 2  # Euler characteristic
    assert len(b) == len(e)  # every edge interior
    np.testing.assert_array_equal(e, orc.build_edges(tris))


def test_nonmanifold_edge_yields_face_pairs():
    # three triangles on one edge (0,1): bends for each pair of faces (ClothSolverHelper.cpp:22-24)
    tris = np.array([[0, 1, 2], [1, 0, 3], [0, 1, 4]], np.int32)
    e = _lib_edges(tris)
    b = _lib_bends(tris, e)
    assert len(b) == 3
    np.testing.assert_array_equal(b, orc.build_bends(tris, orc.build_edges(tris)))


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("arity", [2, 4])
@pytest.mark.parametrize("seed", [0, 7])
def test_coloring_valid_deterministic_and_equal_to_oracle(mode, arity, seed):
    pos, tris = scenes.grid(24, 1.0)
    e = orc.build_edges(tris)
    cons = e if arity == 2 else orc.build_bends(tris, e)
    col, n = _lib_color(cons, len(pos), seed, mode)
    col_o, n_o = orc.color(cons, len(pos), seed, mode)
    np.testing.assert_array_equal(col, col_o)
    assert n == n_o
    assert set(np.unique(col)) == set(range(n))  # compact
    # validity: no vertex appears twice within one color
    for c in range(n):
        verts = cons[col == c].reshape(-1)
        assert len(np.unique(verts)) == len(verts)
    col2, _ = _lib_color(cons, len(pos), seed, mode)
    np.testing.assert_array_equal(col, col2)
    assert n <= (12 if arity == 2 else 24)


def test_greedy_coloring_near_lower_bound_on_grids():
    """Edges at a grid vertex (degree 6) need 6 colors; the greedy pass reaches that, and
    the bend set (>= 12 elements per interior vertex) needs at most 14."""
    pos, tris = scenes.grid(48, 1.0)
    e = orc.build_edges(tris)
    _, n = _lib_color(e, len(pos), 0, 0)
    assert n == 6
    _, nb = _lib_color(orc.build_bends(tris, e), len(pos), 0, 0)
    assert 12 <= nb <= 14


def test_coloring_is_greedy_first_fit_on_a_path():
    # chain of edges 0-1-2-3-4: first fit alternates 0,1,0,1
    cons = np.array([[0, 1], [1, 2], [2, 3], [3, 4]], np.int32)
    col, n = _lib_color(cons, 5, 0, 0)
    np.testing.assert_array_equal(col, [0, 1, 0, 1])
    # a star (all edges share vertex 0) needs one color per edge
    star = np.array([[0, k] for k in range(1, 6)], np.int32)
    col, n = _lib_color(star, 6, 0, 0)
    assert n == 5 and len(set(col)) == 5
    col, n = _lib_color(star, 6, 3, 1)
    assert n == 5 and len(set(col)) == 5


def test_palette_coloring_depends_on_seed():
    pos, tris = scenes.grid(16, 1.0)
    e = orc.build_edges(tris)
    a, _ = _lib_color(e, len(pos), 1, 1)
    b, _ = _lib_color(e, len(pos), 2, 1)
    assert (a != b).any()


def test_host_helpers_reject_bad_input():
    lib = _ffi.load()
    cons = np.array([[0, 5]], np.int32)
    out = np.zeros(1, np.int32)
    n = C.c_int32()
    rc = lib.xrt_color_constraints(_ffi.iptr(cons.reshape(-1)), 2, 1, 3, C.c_uint64(0), 0,
                                   _ffi.iptr(out), C.byref(n))
    assert rc == -1
    ok = np.array([[0, 1]], np.int32)
    rc = lib.xrt_color_constraints(_ffi.iptr(ok.reshape(-1)), 2, 1, 3, C.c_uint64(0), 9,
                                   _ffi.iptr(out), C.byref(n))
    assert rc == -1
    pos, tris = scenes.grid(4, 1.0)
    t = np.ascontiguousarray(tris, np.int32).reshape(-1)
    e = np.zeros(4, np.int32)
    ne = C.c_int32()
    assert lib.xrt_build_edges(_ffi.iptr(t), t.size // 3, _ffi.iptr(e), 2, C.byref(ne)) == -4


def test_context_calls_without_context_fail_cleanly():
    lib = _ffi.load()
    assert lib.xrt_sync(None) == -1
    assert lib.xrt_start(None, C.c_uint64(0)) == -1
    assert lib.xrt_frames_batched(None, 1) == -3
    assert b"xrt_start" in lib.xrt_last_error() or lib.xrt_last_error()


def test_scenes_rng_is_reproducible():
    a = scenes.Rng(3).normal(1001)
    b = scenes.Rng(3).normal(1001)
    np.testing.assert_array_equal(a, b)
    assert abs(a.mean()) < 0.1 and abs(a.std() - 1) < 0.1
    # SplitMix64 stream matches the oracle's mixer
    st = C.c_uint64(3)
    v = [orc.load().orc_splitmix64(C.byref(st)) for _ in range(3)]
    np.testing.assert_array_equal(scenes.Rng(3).uint64(3), np.array(v, np.uint64))


def test_batched_workload_shape():
    specs = scenes.batched(seed=4, n_instances=256)
    sides = np.array([int(round(np.sqrt(len(s.pos)))) for s in specs])
    assert sides.min() >= 32 and sides.max() <= 192
    total = sum(len(s.pos) for s in specs)
    assert 1.5e6 < total < 4.0e6
