"""ctypes binding of the HIP statevector library (include/mettleq_sv.h).

This module replaces the reference's MLX import seam (_mlx_compat.py:3-10):
importing the package never touches the GPU, and the first call that needs
the library raises ``NativeUnavailableError`` if ``libmettleq_sv.so`` has not
been built (there is no CPU fallback on the product path).
"""
from __future__ import annotations

import ctypes
import os
import threading
from pathlib import Path

import numpy as np

LIB_NAME = "libmettleq_sv.so"
LIB_PATH = Path(__file__).resolve().parent / LIB_NAME

MSV_OK = 0
MSV_E_INVALID = 1
MSV_E_OOM = 2
MSV_E_HIP = 3
MSV_E_REFUSED = 4
MSV_E_UNSUPPORTED = 5

MSV_C64 = 0
MSV_C128 = 1

MSV_PHASE_ZZ = 0
MSV_PHASE_CPHASE = 1

MSV_ROT_X = 0
MSV_ROT_Y = 1
MSV_ROT_Z = 2

MSV_G_U = 0
MSV_G_DIAG1 = 1
MSV_G_DIAG2 = 2
MSV_G_SWAP = 3

MSV_TUNE_TILE_LOCAL_BITS = 1
MSV_TUNE_TILE_GATE_BITS = 2
MSV_TUNE_PHASE_BLOCKS = 3
MSV_TUNE_SWEEP_SPLIT = 4
MSV_TUNE_TILE_BLOCKS_PER_CU = 5
MSV_TUNE_ELEM_BLOCKS_PER_CU = 6
MSV_TUNE_COL_BITS = 7
MSV_TUNE_ROT_SPLIT_MIN = 8
MSV_TUNE_XCD_GROUPS = 9
MSV_TUNE_TILE_REG_BITS = 10
MSV_TUNE_TILE_PREFETCH = 11
MSV_TUNE_DEFER_DIAG = 12
MSV_TUNE_DEFER_PERM = 13
MSV_TUNE_FUSED_BLOCKS_PER_CU = 14
MSV_TUNE_REORDER_GATES = 15
MSV_TUNE_DEFER_TAIL = 16
MSV_TUNE_LANE_SWAP = 17
MSV_PENDING_BASIS = 1
MSV_PENDING_DIAG = 2
MSV_PENDING_PERM = 4
MSV_PENDING_TAIL = 8

# kernel kinds of the dispatch log (msv_state_dispatch_log)
KERNEL_KINDS = {
    1: "k_tile", 2: "k_rot", 3: "k_rot_fused", 4: "k_rot_relabel", 5: "k_product", 6: "k_qft",
    7: "k_qft_fused", 8: "k_phase", 9: "k_xor_gather", 10: "k_xor_gather_phase", 11: "k_init_basis",
    12: "k_scale", 13: "k_convert", 14: "k_block_probsum", 15: "k_sample_resolve", 16: "k_probs_full",
    17: "k_marginal", 18: "k_pauli", 19: "k_inner", 20: "k_prefix_seq", 21: "k_sample_search",
}
SWEEP_KINDS = frozenset(range(1, 14))

MSV_PF_OUT_OF_PLACE = 0
MSV_PF_IN_PLACE = 1
MSV_PF_ALLOWED = 0
MSV_PF_ALLOWED_UNVERIFIED = 1
MSV_PF_ALLOWED_OVERRIDE = 2
MSV_PF_REFUSED = 3
MSV_PF_FAIL_MAX_BUFFER = 1
MSV_PF_FAIL_WORKING_SET = 2
MSV_PF_FAIL_QUBITS = 4


class NativeUnavailableError(RuntimeError):
    """The HIP library is not built or cannot be loaded."""


class NativeError(RuntimeError):
    """A HIP runtime failure reported by the library."""


class msv_gate(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("ncontrols", ctypes.c_int32),
        ("q", ctypes.c_int32 * 6),
        ("m", ctypes.c_double * 8),
    ]


class msv_device_info(ctypes.Structure):
    _fields_ = [
        ("total_bytes", ctypes.c_uint64),
        ("free_bytes", ctypes.c_uint64),
        ("max_alloc_bytes", ctypes.c_uint64),
        ("compute_units", ctypes.c_int32),
        ("lds_bytes_per_block", ctypes.c_int32),
        ("name", ctypes.c_char * 96),
        ("arch", ctypes.c_char * 32),
    ]


class msv_mps_stats(ctypes.Structure):
    _fields_ = [
        ("svd_calls", ctypes.c_int64),
        ("two_site_calls", ctypes.c_int64),
        ("swap_calls", ctypes.c_int64),
        ("trunc_events", ctypes.c_int64),
        ("truncated", ctypes.c_int64),
        ("max_bond_ever", ctypes.c_int64),
        ("renormalization_count", ctypes.c_int64),
        ("center", ctypes.c_int64),
        ("local_discarded_weight_sum", ctypes.c_double),
        ("local_discarded_weight_max", ctypes.c_double),
        ("relative_discarded_weight_sum", ctypes.c_double),
        ("relative_discarded_weight_max", ctypes.c_double),
    ]


class msv_preflight_limits(ctypes.Structure):
    _fields_ = [
        ("max_buffer_bytes", ctypes.c_uint64),
        ("working_set_bytes", ctypes.c_uint64),
    ]


class msv_preflight_report(ctypes.Structure):
    _fields_ = [
        ("decision", ctypes.c_int32),
        ("failures", ctypes.c_int32),
        ("allowed", ctypes.c_int32),
        ("overridden", ctypes.c_int32),
        ("buffer_check", ctypes.c_int32),
        ("working_check", ctypes.c_int32),
        ("state_bytes", ctypes.c_uint64),
        ("lower_bound_bytes", ctypes.c_uint64),
        ("max_buffer_bytes", ctypes.c_uint64),
        ("working_set_bytes", ctypes.c_uint64),
    ]


_P = ctypes.c_void_p
_I = ctypes.c_int
_I32P = ctypes.POINTER(ctypes.c_int32)
_U64 = ctypes.c_uint64
_U64P = ctypes.POINTER(ctypes.c_uint64)
_DP = ctypes.POINTER(ctypes.c_double)
_U8P = ctypes.POINTER(ctypes.c_uint8)

# name -> (restype, argtypes); the list is also the set of symbols the header
# declares (tests check both directions).
SIGNATURES = {
    "msv_last_error": (ctypes.c_char_p, []),
    "msv_version": (_I, []),
    "msv_device_count": (_I, [ctypes.POINTER(_I)]),
    "msv_get_device_info": (_I, [_I, ctypes.POINTER(msv_device_info)]),
    "msv_preflight": (_I, [_I, _I, _I, _U64, _I, _I, ctypes.POINTER(msv_preflight_limits),
                           ctypes.POINTER(msv_preflight_report)]),
    "msv_state_create": (_I, [_I, _I, _I, ctypes.POINTER(_P)]),
    "msv_state_destroy": (_I, [_P]),
    "msv_state_reset": (_I, [_P]),
    "msv_state_set_basis": (_I, [_P, _U64]),
    "msv_state_info": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I), _U64P]),
    "msv_state_upload": (_I, [_P, _P, _U64, _U64]),
    "msv_state_download": (_I, [_P, _P, _U64, _U64]),
    "msv_state_copy": (_I, [_P, _P]),
    "msv_state_scale": (_I, [_P, ctypes.c_double, ctypes.c_double]),
    "msv_state_sync": (_I, [_P]),
    "msv_state_release_scratch": (_I, [_P]),
    "msv_state_device_ptr": (_P, [_P]),
    "msv_apply_1q_layer": (_I, [_P, _DP, _U8P]),
    "msv_apply_phase_layer": (_I, [_P, _I32P, _I32P, _DP, _I]),
    "msv_apply_diagonal": (_I, [_P, _I32P, _DP, _I, _I32P, _DP, _I]),
    "msv_apply_xor_affine": (_I, [_P, _U64P, _U64]),
    "msv_apply_qft": (_I, [_P, _I, _I, _I]),
    "msv_apply_qft_stage": (_I, [_P, _I, _I, _I]),
    "msv_apply_gates": (_I, [_P, ctypes.POINTER(msv_gate), _I]),
    "msv_apply_circuit": (_I, [_P, _I32P, _I32P, _DP, _I]),
    "msv_probabilities": (_I, [_P, _I32P, _I, _DP]),
    "msv_sample": (_I, [_P, _I32P, _I, _U64, _U64, _U64, _U64P]),
    "msv_sample_noalloc": (_I, [_P, _I32P, _I, _U64, _U64, _U64, _U64P]),
    "msv_expval_pauli": (_I, [_P, _U64P, _U64P, _I, _DP, _DP]),
    "msv_inner": (_I, [_P, _P, _DP, _DP]),
    "msv_norm2": (_I, [_P, _DP]),
    "msv_set_tuning": (_I, [_I, ctypes.c_int64]),
    "msv_get_tuning": (_I, [_I, ctypes.POINTER(ctypes.c_int64)]),
    "msv_state_sweep_count": (_I, [_P, _U64P, _I]),
    "msv_state_create_virtual": (_I, [_I, _I, _I, ctypes.POINTER(_P)]),
    "msv_state_dispatch_log": (_I, [_P, _I32P, _I, _U64P, _I]),
    "msv_state_pending": (_I, [_P, ctypes.POINTER(ctypes.c_int)]),
    "msv_state_timer_start": (_I, [_P]),
    "msv_state_timer_stop": (_I, [_P, _DP]),
    "msv_probe_copy_bandwidth": (_I, [_I, _U64, _I, _DP]),
    "msv_ansatz_create": (_I, [_I, _I, ctypes.POINTER(_P)]),
    "msv_ansatz_destroy": (_I, [_P]),
    "msv_ansatz_rotation": (_I, [_P, _I, _I, _I]),
    "msv_ansatz_phase": (_I, [_P, _I, _I, _I, ctypes.c_double]),
    "msv_ansatz_observable": (_I, [_P, _U64P, _DP, _I]),
    "msv_ansatz_evaluate": (_I, [_P, _DP, _I, _I, _DP, _DP]),
    "msv_ansatz_info": (_I, [_P, _I, _I, ctypes.POINTER(ctypes.c_int64), _I]),
    "msv_ansatz_time_sweeps": (_I, [_P, _I, _I, _I, ctypes.POINTER(ctypes.c_double)]),
    "msv_mps_create": (_I, [_I, _I, ctypes.c_double, _I, _I, ctypes.POINTER(_P)]),
    "msv_mps_destroy": (_I, [_P]),
    "msv_mps_reset": (_I, [_P]),
    "msv_mps_apply_1q": (_I, [_P, _I, _DP]),
    "msv_mps_apply_2q": (_I, [_P, _I, _I, _DP]),
    "msv_mps_bond_dims": (_I, [_P, _I32P]),
    "msv_mps_get_stats": (_I, [_P, ctypes.POINTER(msv_mps_stats)]),
    "msv_mps_site_tensor": (_I, [_P, _I, _P, _I32P, _I32P]),
    "msv_mps_norm": (_I, [_P, _DP]),
    "msv_mps_expval_product": (_I, [_P, _I32P, _DP, _I, _DP, _DP]),
    "msv_mps_to_state": (_I, [_P, _P]),
    "msv_mps_sync": (_I, [_P]),
}

_lib = None
_lock = threading.Lock()


def library_path() -> Path:
    override = os.environ.get("METTLEQ_SV_LIBRARY")
    return Path(override) if override else LIB_PATH


def load_library(path: Path | None = None) -> ctypes.CDLL:
    """Load (once) and declare every entry point; raise if it is missing."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = Path(path) if path is not None else library_path()
        if not p.exists():
            raise NativeUnavailableError(
                f"{p} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
                "(hipcc --offload-arch=gfx950) before using the statevector engine"
            )
        try:
            lib = ctypes.CDLL(str(p))
        except OSError as error:  # pragma: no cover - depends on the host
            raise NativeUnavailableError(f"cannot load {p}: {error}") from error
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if path is None:
            _lib = lib
        return lib


def lib() -> ctypes.CDLL:
    return _lib if _lib is not None else load_library()


def last_error() -> str:
    msg = lib().msv_last_error()
    return msg.decode("utf-8", "replace") if msg else ""


def check(code: int, what: str = "") -> None:
    """Map a status code to the Python exception the reference raises."""
    if code == MSV_OK:
        return
    msg = last_error() or f"{what} failed with status {code}"
    if code == MSV_E_INVALID:
        raise ValueError(msg)
    if code in (MSV_E_OOM, MSV_E_REFUSED):
        raise MemoryError(msg)
    if code == MSV_E_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise NativeError(msg)


def device_count() -> int:
    out = ctypes.c_int(0)
    code = lib().msv_device_count(ctypes.byref(out))
    if code != MSV_OK:
        return 0
    return int(out.value)


def device_info(device: int = 0) -> dict:
    info = msv_device_info()
    check(lib().msv_get_device_info(int(device), ctypes.byref(info)), "device_info")
    return {
        "total_bytes": int(info.total_bytes),
        "free_bytes": int(info.free_bytes),
        "max_alloc_bytes": int(info.max_alloc_bytes),
        "compute_units": int(info.compute_units),
        "lds_bytes_per_block": int(info.lds_bytes_per_block),
        "name": info.name.decode("utf-8", "replace"),
        "arch": info.arch.decode("utf-8", "replace"),
    }


def preflight(n_qubits: int, dtype: int, mode: int, reserved_bytes: int, allow_unsafe: bool,
              limits: tuple | None, device: int = 0) -> dict:
    """msv_preflight: `limits` = (max_buffer_bytes, working_set_bytes) with 0
    or None for unknown, or None for the device's live limits. Returns the
    report fields and, when refused, the library's message."""
    rep = msv_preflight_report()
    lim_p = None
    if limits is not None:
        lim = msv_preflight_limits(int(limits[0] or 0), int(limits[1] or 0))
        lim_p = ctypes.byref(lim)
    code = lib().msv_preflight(int(n_qubits), int(dtype), int(mode), int(reserved_bytes),
                               1 if allow_unsafe else 0, int(device), lim_p, ctypes.byref(rep))
    if code not in (MSV_OK, MSV_E_REFUSED):
        check(code, "preflight")
    out = {name: int(getattr(rep, name)) for name, _ in msv_preflight_report._fields_}
    out["message"] = last_error() if code == MSV_E_REFUSED else None
    return out


def as_ptr(array: np.ndarray, ctype):
    return array.ctypes.data_as(ctypes.POINTER(ctype))


def set_tuning(key: int, value: int) -> None:
    check(lib().msv_set_tuning(int(key), int(value)), "set_tuning")


def get_tuning(key: int) -> int:
    out = ctypes.c_int64(0)
    check(lib().msv_get_tuning(int(key), ctypes.byref(out)), "get_tuning")
    return int(out.value)


def probe_copy_ms(device: int, nbytes: int, iters: int = 10) -> float:
    out = ctypes.c_double(0.0)
    check(lib().msv_probe_copy_bandwidth(int(device), int(nbytes), int(iters), ctypes.byref(out)),
          "probe_copy")
    return float(out.value)
