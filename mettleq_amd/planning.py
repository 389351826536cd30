"""Execution-method and device selection for SDK callers, for one MI355X.

The reference selects one method (statevector or MPS) and one device (CPU or
Apple GPU) per execution with an auditable reason (planning.py:27-204). This
engine has one exact method on one device type: the gfx950 statevector in
complex64 or complex128. Selection here therefore answers two questions:

* method: "statevector" or "matrix_product_state" (mettleq_amd.mps) with
  the reference's rules: explicit requests are kept; an automatic request
  takes MPS only with allow_approximation, for shallow local circuits from
  automatic_mps_min_qubits qubits on, or when the statevector preflight
  refuses; otherwise an exact request the preflight refuses stays exact so
  that the preflight raises its StatevectorMemoryError;
* device: "gpu" at or above the crossover width, "cpu" below it, where "cpu"
  names the caller's own CPU engine (the adaptive SDK backends run Aer or
  Lightning there): this package ships no CPU statevector. The MPS method
  runs on the GPU here (its SVDs run on the device), where the reference's
  is CPU-only.

MI355X calibration of the crossover. Below about 20 qubits a circuit on this
engine costs its fixed per-circuit latency, not HBM time: about 1 ms per
request at 12-20 qubits in the served stream (host translation, fusion and a
handful of sweep launches), against a 2^n-proportional CPU cost. The
gate-by-gate CPU path of the reference, restated in C, moves 19.7 GB/s on one
core of the MI355X box's EPYC host: a 100-gate circuit costs
100 x 2^n x 16 B / 19.7 GB/s, 1.3 ms at 14 qubits and 0.3 ms at 12. The
default crossover is therefore 14 qubits (the reference's M3 Pro value is 16).
Callers with their own measurements pass statevector_gpu_min_qubits.
"""
from __future__ import annotations

from dataclasses import asdict, dataclass
from typing import Optional, Sequence

from .execution import statevector_preflight

METHOD_ALIASES = {
    "auto": "automatic",
    "automatic": "automatic",
    "sv": "statevector",
    "statevector": "statevector",
    "mps": "matrix_product_state",
    "matrix_product_state": "matrix_product_state",
}
DEVICE_ALIASES = {"auto": "auto", "cpu": "cpu", "gpu": "gpu"}

DEFAULT_STATEVECTOR_GPU_MIN_QUBITS = 14
DEFAULT_AUTOMATIC_MPS_MIN_QUBITS = 24


@dataclass(frozen=True)
class ExecutionSelection:
    requested_method: str
    selected_method: str
    requested_device: str
    selected_device: str
    dtype: str
    reason: str
    statevector_preflight: Optional[dict]
    allow_approximation: bool = False
    mps_compatible: bool = False
    mps_compatibility_reasons: tuple = ()
    mps_svd_device: Optional[str] = None
    cpu_gpu_policy: str = (
        "one numerical device is selected per execution; CPU and GPU times "
        "are benchmarked independently and are not combined into a speed claim"
    )

    def to_dict(self) -> dict:
        result = asdict(self)
        result["mps_compatibility_reasons"] = list(self.mps_compatibility_reasons)
        return result


def normalize_method(method: str) -> str:
    try:
        return METHOD_ALIASES[str(method).strip().lower()]
    except KeyError as exc:
        raise ValueError("method must be one of: automatic, statevector, matrix_product_state") from exc


def normalize_device(device: str) -> str:
    try:
        return DEVICE_ALIASES[str(device).strip().lower()]
    except KeyError as exc:
        raise ValueError("device must be one of: auto, cpu, gpu") from exc


def _mps_compatibility(n_qubits: int, operations: Sequence[dict]):
    """Shallow local circuits only (planning.py:77-99): no three-qubit ops,
    two-qubit ops on neighbours, at most 4 (n - 1) entangling ops."""
    reasons = []
    entanglers = 0
    for op in operations:
        wires = list(op.get("wires", []))
        if len(wires) >= 3:
            reasons.append("contains_three_or_more_qubit_operation")
        elif len(wires) == 2:
            entanglers += 1
            if abs(int(wires[0]) - int(wires[1])) != 1:
                reasons.append("contains_non_adjacent_two_qubit_operation")
    if entanglers > max(1, 4 * max(1, n_qubits - 1)):
        reasons.append("entangling_operation_count_exceeds_auto_mps_limit")
    unique = tuple(sorted(set(reasons)))
    return not unique, unique, entanglers


def _gpu_available() -> bool:
    try:
        from . import _native
        return _native.device_count() > 0
    except Exception:
        return False


def select_execution(
    n_qubits: int,
    operations: Sequence[dict],
    *,
    method: str = "automatic",
    device: str = "auto",
    dtype: str = "complex64",
    allow_approximation: bool = False,
    allow_unsafe_statevector: Optional[bool] = None,
    statevector_gpu_min_qubits: int = DEFAULT_STATEVECTOR_GPU_MIN_QUBITS,
    automatic_mps_min_qubits: int = DEFAULT_AUTOMATIC_MPS_MIN_QUBITS,
) -> ExecutionSelection:
    """One method and one device with an auditable reason (planning.py:108-204)."""
    n = int(n_qubits)
    requested_method = normalize_method(method)
    requested_device = normalize_device(device)
    if not isinstance(allow_approximation, bool):
        raise TypeError("allow_approximation must be a boolean")
    if statevector_gpu_min_qubits < 1:
        raise ValueError("statevector GPU crossover threshold must be positive")
    if dtype not in ("complex64", "complex128"):
        raise ValueError("dtype must be complex64 or complex128")
    mps_ok, mps_reasons, _ = _mps_compatibility(n, operations)
    preflight = None
    if requested_method in ("automatic", "statevector"):
        preflight = statevector_preflight(n, dtype=dtype, allow_unsafe=allow_unsafe_statevector)
    if requested_method == "statevector":
        selected_method, reason = "statevector", "explicit_statevector_request"
    elif requested_method == "matrix_product_state":
        selected_method, reason = "matrix_product_state", "explicit_matrix_product_state_request"
    elif allow_approximation and n >= int(automatic_mps_min_qubits) and mps_ok:
        selected_method, reason = "matrix_product_state", "automatic_shallow_local_mps_with_approximation_allowed"
    elif preflight is not None and preflight["allowed"]:
        selected_method, reason = "statevector", "automatic_exact_statevector_preflight_passed"
    elif allow_approximation and mps_ok:
        selected_method, reason = "matrix_product_state", "automatic_statevector_refused_mps_fallback_allowed"
    else:
        # keep the exact request: the preflight raises before allocation
        selected_method, reason = "statevector", "automatic_exact_request_has_no_allowed_approximate_fallback"
    if requested_device != "auto":
        if requested_device == "gpu" and not _gpu_available():
            raise RuntimeError("MI355X execution was requested but no HIP device is visible")
        if selected_method == "matrix_product_state" and requested_device == "cpu":
            raise ValueError("the MPS method runs on the GPU here; use device='gpu' or 'auto'")
        selected_device = requested_device
        device_reason = f"explicit_{requested_device}_request"
    elif not _gpu_available():
        selected_device = "cpu"
        device_reason = "automatic_gpu_unavailable"
    elif selected_method == "matrix_product_state":
        selected_device = "gpu"
        device_reason = "automatic_mps_device_native"
    elif n < int(statevector_gpu_min_qubits):
        selected_device = "cpu"
        device_reason = "automatic_below_measured_statevector_gpu_crossover"
    else:
        selected_device = "gpu"
        device_reason = "automatic_at_or_above_measured_statevector_gpu_crossover"
    return ExecutionSelection(
        requested_method=requested_method,
        selected_method=selected_method,
        requested_device=requested_device,
        selected_device=selected_device,
        dtype=dtype,
        reason=f"{reason}; {device_reason}",
        statevector_preflight=preflight,
        allow_approximation=allow_approximation,
        mps_compatible=mps_ok,
        mps_compatibility_reasons=mps_reasons,
        mps_svd_device="gpu" if selected_method == "matrix_product_state" else None,
    )
