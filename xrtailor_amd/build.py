"""Build the HIP library (gfx950) in-tree and the CPU oracle used by the tests.

    python -m xrtailor_amd.build            # library + oracle
    python -m xrtailor_amd.build --lib      # library only

Outputs:
    xrtailor_amd/libxrtailor_amd.so   the product (C ABI declared in include/xrtailor_amd.h)
    oracle/_build/libxrt_oracle.so    the FP64 serial checker (test infrastructure)
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "xrtailor_amd" / "csrc"
OBJ = ROOT / "xrtailor_amd" / "_build"
LIB = ROOT / "xrtailor_amd" / "libxrtailor_amd.so"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "_build" / "libxrt_oracle.so"

ARCH = os.environ.get("XRT_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 VALU forms are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]
# kernels whose float results feed integer outputs (Morton codes, AABB tests) are built
# without FMA contraction so that they agree bit for bit with the FP32-strict oracle
STRICT_FP = {"bvh.hip", "contact.hip", "adjoint_contact.hip", "particle.hip", "particle_api.hip",
             "binding.hip"}
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")


def build_lib(jobs: int = 8) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    steps, objs = [], []
    for src in sorted(CSRC.glob("*.hip")):
        strict = ["-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt"]
        flags = HIP_FLAGS + (strict if src.name in STRICT_FP else [])
        objs.append(str(OBJ / (src.stem + ".o")))
        steps.append([HIPCC, *flags, "-c", str(src), "-o", objs[-1]])
    for src in sorted(CSRC.glob("*.cpp")):
        objs.append(str(OBJ / (src.stem + ".cpp.o")))
        steps.append(["g++", *CXX_FLAGS, "-c", str(src), "-o", objs[-1]])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, steps))
    # link exactly the objects of this build: an object left behind by a renamed or
    # deleted source must not end up in the library
    tmp = LIB.with_suffix(f".so.tmp{os.getpid():x}")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    out = ORACLE_LIB
    out.parent.mkdir(parents=True, exist_ok=True)
    srcs = sorted(str(p) for p in ORACLE_DIR.glob("*.c"))
    tmp = out.with_suffix(f".so.tmp{os.getpid():x}")
    _run(["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-ffp-contract=off", "-fno-fast-math",
          "-Wall", "-o", str(tmp), *srcs, "-lm"])
    os.replace(tmp, out)
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="library only")
    ap.add_argument("--oracle", action="store_true", help="oracle only")
    ap.add_argument("--clean", action="store_true")
    a = ap.parse_args(argv)
    if a.clean and OBJ.exists():
        shutil.rmtree(OBJ)
    if not a.oracle:
        print("built", build_lib())
    if not a.lib:
        print("built", build_oracle())
    return 0


if __name__ == "__main__":
    sys.exit(main())
