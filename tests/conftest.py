import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libkite_gpu.so")


def _have_gpu() -> bool:
    """A HIP device the engine can open (through the runtime the engine
    library links, not through torch)."""
    try:
        from kite_amd import GpuContext
        GpuContext(0).close()
        return True
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu_ctx():
    """One engine context for the whole GPU session (tests run in one process)."""
    if not _have_gpu():
        pytest.skip("no HIP device")
    from kite_amd import GpuContext
    ctx = GpuContext(0)
    yield ctx
    ctx.close()


@pytest.fixture
def tmpdir_path(tmp_path):
    return tmp_path
