"""The C++ binding shown in INTEGRATION.md compiles against include/xrtailor_amd.h and
links against the built library (CPU only: nothing is called).  The engine-side types the
snippets use (glm::vec3, SimParams, FEMFabricSettings, Mesh, SkinWeights) are reduced to
the fields the snippets read; every xrt_* call in the snippets must exist with the
signature shown."""
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent

ENGINE_TYPES = r"""
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>
#include "xrtailor_amd.h"
namespace glm { struct vec3 { float x, y, z; }; }
struct SimParams {
  glm::vec3 gravity; float damping; int num_substeps, num_iterations;
  float long_range_stretchiness, bvh_tolerance;
};
struct FEMFabricSettings {
  float xx_stiffness, yy_stiffness, xy_stiffness, xy_poisson_ratio, yx_poisson_ratio;
  float bending_stiffness, long_range_stretchiness; bool solve_bending;
};
struct Mesh { int nv, nt; std::vector<int32_t> tris; };
struct SkinWeights { std::vector<int32_t> ptr, joint; std::vector<float> weight; int n_joints; };
"""


def _blocks():
    text = (ROOT / "INTEGRATION.md").read_text()
    return re.findall(r"```cpp\n(.*?)```", text, flags=re.S)


def _assemble():
    blocks = _blocks()
    head = blocks[0]
    cut = head.index(" private:")
    body = head[:cut] + "".join(blocks[1:]) + head[cut:]
    body = body.replace('#include "xrtailor_amd.h"', "")
    names = re.findall(r"^  (?:static )?(?:void|int|bool|float)\s+(\w+)\(", body, flags=re.M)
    names = [n for n in dict.fromkeys(names) if n != "Check"]
    uses = "\n".join(f"  volatile auto p{k} = &AmdClothSolver::{n}; (void)p{k};"
                     for k, n in enumerate(names))
    return ENGINE_TYPES + body + "\nint main() {\n" + uses + "\n  return 0;\n}\n", names


def test_integration_snippets_are_one_class():
    src, names = _assemble()
    assert len(_blocks()) >= 6
    for n in ("AddCloth", "Start", "FixedUpdate", "SimulateFrame", "UseFem", "Gradients"):
        assert n in names, n


def test_integration_binding_compiles_and_links(tmp_path):
    lib = ROOT / "xrtailor_amd" / "libxrtailor_amd.so"
    if not lib.exists():
        pytest.skip("library not built")
    src, names = _assemble()
    cpp = tmp_path / "amd_solver.cpp"
    cpp.write_text(src)
    exe = tmp_path / "amd_solver"
    cmd = ["g++", "-std=c++17", "-Wall", "-Werror", "-Wno-unused-variable",
           f"-I{ROOT / 'include'}", str(cpp), "-o", str(exe),
           f"-L{lib.parent}", "-lxrtailor_amd", "-Wl,-rpath-link,/opt/rocm/lib",
           "-Wl,--no-undefined"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    nm = subprocess.run(["nm", "-u", str(exe)], capture_output=True, text=True).stdout
    used = set(re.findall(r"\b(xrt_\w+)", nm))
    exported = set(re.findall(r"\b(xrt_\w+)\b", subprocess.run(
        ["nm", "-D", "--defined-only", str(lib)], capture_output=True, text=True).stdout))
    print(f"{len(names)} member functions, {len(used)} xrt_ entry points referenced")
    assert used and used <= exported
    assert {"xrt_vf_dcd_generate", "xrt_repulsion_update_pairs", "xrt_frames_batched"} <= used
