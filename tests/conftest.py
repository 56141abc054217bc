import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")


def _newest(paths):
    return max((p.stat().st_mtime for p in paths), default=0.0)


def _ensure_built():
    """Build what is missing; rebuild the CPU oracle (the checker) when a source of it is
    newer. The HIP library is never rebuilt here once it exists (tests load what is built):
    a library older than its sources is reported so that a stale build is not tested
    silently."""
    from xrtailor_amd import build
    lib = ROOT / "xrtailor_amd" / "libxrtailor_amd.so"
    orc = ROOT / "oracle" / "_build" / "libxrt_oracle.so"
    if not lib.exists():
        build.build_lib()
    elif _newest(list(build.CSRC.glob("*")) + list((ROOT / "include").glob("*.h"))) > lib.stat().st_mtime:
        sys.stderr.write("warning: xrtailor_amd/libxrtailor_amd.so is older than its sources; "
                         "run python -m xrtailor_amd.build\n")
    orc_src = list(build.ORACLE_DIR.glob("*.c")) + list(build.ORACLE_DIR.glob("*.h"))
    if not orc.exists() or _newest(orc_src) > orc.stat().st_mtime:
        build.build_oracle()


_ensure_built()
