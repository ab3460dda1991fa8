import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _ensure_built():
    from oracle import build as ob
    if not ob.LIB.exists():
        ob.build()
    lib = ROOT / "fast_ldpc_decoder_for_gpu_fixed_amd" / "libldpc_mi355x.so"
    if not lib.exists():
        from fast_ldpc_decoder_for_gpu_fixed_amd import build
        build.build_lib()


_ensure_built()


@pytest.fixture(scope="session")
def gpu_available():
    import fast_ldpc_decoder_for_gpu_fixed_amd as L
    if L.device_count() < 1:
        pytest.fail("no gfx950 device visible for a test marked gpu")
    return True


REFERENCE = Path(os.environ.get("LDPC_REFERENCE_ROOT", "/root/reference"))
