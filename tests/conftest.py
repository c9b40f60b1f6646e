import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


def _has_gpu() -> bool:
    try:
        import lumen_amd
        import ctypes
        h = ctypes.c_void_p()
        st = lumen_amd.lib().lt_context_create(0, ctypes.byref(h))
        if st == 0:
            lumen_amd.lib().lt_context_destroy(h)
            return True
    except Exception:
        pass
    return False


@pytest.fixture(scope="session")
def gpu_ctx():
    import lumen_amd
    print(f"[gpu] creating context", flush=True)
    ctx = lumen_amd.Context(0)
    yield ctx
    ctx.close()
