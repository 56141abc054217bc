import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")


def _ensure_built():
    lib = ROOT / "xrtailor_amd" / "libxrtailor_amd.so"
    orc = ROOT / "oracle" / "_build" / "libxrt_oracle.so"
    if not lib.exists() or not orc.exists():
        from xrtailor_amd import build
        build.main([])


_ensure_built()
