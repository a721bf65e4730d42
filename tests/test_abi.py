"""The C-ABI library loads and exports every symbol include/ouxsbli.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import re

from ouxsbli_amd import _abi, _native
from conftest import ROOT


def _declared():
    text = (ROOT / "include" / "ouxsbli.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(ox(?:2d)?_\w+)\s*\(", text, re.M)))


def test_header_and_python_agree():
    assert _declared() == sorted(_abi.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(str(_native.LIB_PATH))
    for name in _declared():
        assert hasattr(L, name), name


def test_version_and_errors_without_gpu():
    L = _native.load()
    assert L.ox_version().decode().startswith("ouxsbli-mi355x")
    # argument validation happens before any HIP call
    cfg = _abi.OxConfig()
    cfg.abi_version = 999
    h = ctypes.c_void_p()
    rc = L.ox_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc == _abi.OX_EINVAL
    assert b"abi_version" in L.ox_last_error()
    assert L.ox_create(None, ctypes.byref(h)) == _abi.OX_EINVAL


def test_config_struct_layout():
    # 17 int32 then 9 doubles; doubles 8-aligned
    assert ctypes.sizeof(_abi.OxConfig) == 17 * 4 + 4 + 9 * 8
    assert ctypes.sizeof(_abi.OxDiag) == 6 * 8
