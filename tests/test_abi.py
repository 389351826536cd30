"""The C-ABI library loads on a CPU-only host and exports exactly what
include/mettleq_sv.h declares (no compute calls without a GPU)."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "mettleq_sv.h"


def _declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(msv_[a-z0-9_]+)\s*\(", text)))


def _lib_path():
    from mettleq_amd import _native
    p = _native.library_path()
    if not p.exists():
        pytest.fail(f"{p} is not built (run __graft_entry__.build())")
    return p


def test_library_exports_every_declared_symbol():
    p = _lib_path()
    out = subprocess.run(["nm", "-D", "--defined-only", str(p)], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (msv_[a-z0-9_]+)\b", out))
    declared = set(_declared())
    assert declared, "no declarations parsed"
    assert declared <= exported, sorted(declared - exported)
    assert exported <= declared, sorted(exported - declared)


def test_ctypes_signature_table_matches_header():
    from mettleq_amd import _native
    assert set(_native.SIGNATURES) == set(_declared())


def test_library_loads_and_reports_without_gpu():
    from mettleq_amd import _native
    lib = _native.load_library()
    assert lib.msv_version() >= 10000
    # no GPU here: a device count call must return cleanly (0 devices or an error code)
    n = ctypes.c_int(-1)
    code = lib.msv_device_count(ctypes.byref(n))
    assert code in (0, _native.MSV_E_HIP)


def test_invalid_arguments_are_rejected_before_any_device_call():
    from mettleq_amd import _native
    lib = _native.load_library()
    h = ctypes.c_void_p()
    assert lib.msv_state_create(0, 0, 0, ctypes.byref(h)) == _native.MSV_E_INVALID
    assert "n_qubits" in _native.last_error()
    assert lib.msv_state_create(41, 0, 0, ctypes.byref(h)) == _native.MSV_E_INVALID
    assert lib.msv_state_create(5, 7, 0, ctypes.byref(h)) == _native.MSV_E_INVALID
    assert lib.msv_state_destroy(None) == 0
    assert lib.msv_apply_qft(None, 0, 1, 0) == _native.MSV_E_INVALID
    assert lib.msv_set_tuning(999, 1) == _native.MSV_E_INVALID
    v = ctypes.c_int64()
    assert lib.msv_get_tuning(_native.MSV_TUNE_TILE_LOCAL_BITS, ctypes.byref(v)) == 0
    assert v.value == 13


def test_error_mapping():
    from mettleq_amd import _native
    with pytest.raises(ValueError):
        _native.check(_native.MSV_E_INVALID)
    with pytest.raises(MemoryError):
        _native.check(_native.MSV_E_OOM)
    with pytest.raises(NotImplementedError):
        _native.check(_native.MSV_E_UNSUPPORTED)
    with pytest.raises(_native.NativeError):
        _native.check(_native.MSV_E_HIP)


def test_missing_library_fails_loudly(tmp_path):
    from mettleq_amd import _native
    with pytest.raises(_native.NativeUnavailableError):
        _native.load_library(tmp_path / "absent.so")


def test_kernels_are_built_for_gfx950():
    """The embedded device code object targets gfx950 (MI355X)."""
    data = _lib_path().read_bytes()
    assert b"amdgcn-amd-amdhsa--gfx950" in data
