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


def test_grid_limits_refused_without_gpu():
    """ox_create takes grids up to 2^31 cells (64-bit state indices) and xy
    planes up to 2^31 / 15 cells (the z sweep's 32-bit offsets from a column
    pointer span at least three planes); beyond them it refuses before any
    HIP call."""
    from ouxsbli_amd.cases import CaseSpec
    L = _native.load()
    h = ctypes.c_void_p()
    too_many = CaseSpec(nx=1300, ny=1300, nz=1300, order=6, visc="Euler").to_ox()
    assert L.ox_create(ctypes.byref(too_many), ctypes.byref(h)) == _abi.OX_EINVAL
    assert b"2^31 cells" in L.ox_last_error()
    wide = CaseSpec(nx=12000, ny=12000, nz=6, order=2, visc="Euler").to_ox()
    assert L.ox_create(ctypes.byref(wide), ctypes.byref(h)) == _abi.OX_EINVAL
    assert b"xy plane" in L.ox_last_error()


def test_footprint_of_a_grid_beyond_32_bit_state_indices():
    """760^3 with ghosts holds 5 * 760^3 > 2^31 state values; its FP64 NS RK3
    footprint (two states, the residual, two sets of six primitive arrays)
    fits one MI355X's 288 GB."""
    from ouxsbli_amd.cases import CaseSpec
    L = _native.load()
    spec = CaseSpec(nx=760, ny=760, nz=760, order=6, visc="NS", rk=3)
    b = ctypes.c_uint64()
    assert L.ox_footprint(ctypes.byref(spec.to_ox()), ctypes.byref(b)) == _abi.OX_OK
    cells = 760**3
    assert 5 * cells > 2**31
    assert b.value >= cells * (2 * 40 + 40 + 2 * 6 * 8)
    assert b.value < cells * (2 * 40 + 40 + 2 * 6 * 8) + (1 << 20)
    assert b.value < 288e9
