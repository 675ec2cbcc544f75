"""The C ABI library loads and exports every symbol include/kite_gpu.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def declared_functions() -> list[str]:
    text = (ROOT / "include" / "kite_gpu.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kite_gpu_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from kite_amd import _native
    lib = ctypes.CDLL(str(_native.gpu_lib_path()))
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python mirror binds the same set
    assert set(names) <= set(_native.GPU_SYMBOLS)


def test_abi_version_and_error_string():
    from kite_amd import load_gpu
    lib = load_gpu()
    assert lib.kite_gpu_abi_version() == 5
    assert isinstance(lib.kite_gpu_last_error(), bytes)


def test_invalid_arguments_fail_before_any_device_work():
    from kite_amd import load_gpu
    lib = load_gpu()
    assert lib.kite_gpu_count_begin(None, 31, 1, 256) == -1
    assert b"ctx is NULL" in lib.kite_gpu_last_error()
    assert lib.kite_gpu_open(0, 0, None) == -1
