"""Statevector memory preflight, runtime capability report and execution plan.

Restates the reference's preflight contract (execution.py:186-407) for one
MI355X: the same decision strings, report keys, override policy and
``StatevectorMemoryError``, re-sized for 288 GB HBM and for this engine's
kernels, which are in place except the XOR permutation (a gather into a
second buffer). Two lower bounds therefore exist:

  in_place      one state + the library's fixed scratch (staging ring and
                reduction buffers) -- circuits without XOR layers
  out_of_place  two states (the reference's model, execution.py:195-223)

Device limits come from an injectable provider (``_device_info``) so tests
can fake a constrained GPU exactly like the reference's tests do
(tests/mlxQExecutionPlanTest.py:69-81).
"""
from __future__ import annotations

import math
import os
from collections import Counter
from typing import Any, Dict, Iterable, List, Optional

STATE_DTYPE = "complex64"
BYTES_PER_AMPLITUDE = {"complex64": 8, "complex128": 16}
INDEX_BITS = 64
QUBIT_LIMIT = 40
# staging ring (1 MiB) + reduction scratch (up to 1 MiB for the usual
# reductions) + LUTs, rounded up; full probability vectors are separate
FIXED_SCRATCH_BYTES = 8 << 20
STATEVECTOR_UNSAFE_OVERRIDE_ENV = "METTLEQ_ALLOW_UNSAFE_STATEVECTOR"
HIP_KERNELS_ENV = "METTLEQ_HIP_KERNELS"

_TRUE_VALUES = {"1", "true", "on", "yes", "enabled"}
_FALSE_VALUES = {"0", "false", "off", "no", "disabled", ""}


class StatevectorMemoryError(MemoryError):
    """Raised before allocation when reported device limits are insufficient."""

    def __init__(self, message: str, report: Dict[str, Any]):
        super().__init__(message)
        self.report = report


def _device_info() -> Dict[str, Any]:
    """Device memory limits of GPU 0 from the HIP runtime (empty if none)."""
    try:
        from . import _native
        if _native.device_count() < 1:
            return {}
        info = _native.device_info(0)
    except Exception:
        return {}
    # free memory now: it already excludes the states this process holds
    return {
        "device_name": info["name"],
        "architecture": info["arch"],
        "max_buffer_length": info["max_alloc_bytes"] or None,
        "max_recommended_working_set_size": info["free_bytes"],
        "total_bytes": info["total_bytes"],
        "compute_units": info["compute_units"],
    }


def _limit(value) -> Optional[int]:
    if value is None:
        return None
    value = int(value)
    return value if value > 0 else None


def state_memory_estimate(n_qubits: int, dtype: str = STATE_DTYPE) -> Dict[str, Any]:
    """State bytes and the two lower bounds (execution.py:195-223)."""
    n = int(n_qubits)
    if n < 0:
        raise ValueError("n_qubits must be non-negative")
    if dtype not in BYTES_PER_AMPLITUDE:
        raise ValueError(f"dtype must be one of {sorted(BYTES_PER_AMPLITUDE)}")
    bpa = BYTES_PER_AMPLITUDE[dtype]
    state_bytes = (1 << n) * bpa
    return {
        "dtype": dtype,
        "amplitudes": 1 << n,
        "bytes_per_amplitude": bpa,
        "state_bytes": state_bytes,
        "state_mib": state_bytes / (1024.0 * 1024.0),
        "temporary_memory_expectation_bytes": state_bytes,
        "minimum_input_plus_output_bytes": 2 * state_bytes,
        "minimum_in_place_bytes": state_bytes + FIXED_SCRATCH_BYTES,
        "expectation_basis": (
            "in-place layer kernels need one state plus fixed scratch; an XOR "
            "permutation gathers into a second state buffer"
        ),
    }


def statevector_unsafe_override_policy(allow_unsafe: Optional[bool] = None) -> Dict[str, Any]:
    """Constructor argument first, then the environment; invalid values fail
    closed (execution.py:226-267)."""
    raw = os.environ.get(STATEVECTOR_UNSAFE_OVERRIDE_ENV)
    if allow_unsafe is not None:
        if not isinstance(allow_unsafe, bool):
            raise TypeError("allow_unsafe_statevector must be a boolean or None")
        return {"enabled": allow_unsafe, "source": "constructor_argument",
                "environment_variable": STATEVECTOR_UNSAFE_OVERRIDE_ENV, "environment_value": raw}
    if raw is None:
        return {"enabled": False, "source": "default_disabled",
                "environment_variable": STATEVECTOR_UNSAFE_OVERRIDE_ENV, "environment_value": None}
    normalized = raw.strip().lower()
    if normalized in _TRUE_VALUES:
        enabled = True
    elif normalized in _FALSE_VALUES:
        enabled = False
    else:
        raise ValueError(f"{STATEVECTOR_UNSAFE_OVERRIDE_ENV} must be a boolean value")
    return {"enabled": enabled, "source": "environment",
            "environment_variable": STATEVECTOR_UNSAFE_OVERRIDE_ENV, "environment_value": raw}


_DECISIONS = {
    0: ("allowed_within_reported_lower_bounds",
        "statevector lower bounds fit within the reported device limits"),
    1: ("allowed_with_unverified_device_limits",
        "one or more device memory limits are unavailable"),
    2: ("allowed_with_unsafe_override",
        "reported device limits reject the allocation, but the explicit unsafe override permits it"),
    3: ("refused", "reported device limits are below the statevector lower bound"),
}


def statevector_preflight(
    n_qubits: int,
    *,
    dtype: str = STATE_DTYPE,
    allow_unsafe: Optional[bool] = None,
    mode: str = "out_of_place",
    reserved_bytes: int = 0,
) -> Dict[str, Any]:
    """One-sided check before allocating a statevector (execution.py:270-374).

    The decision is the library's ``msv_preflight`` (one implementation for
    every binder of the C ABI); this function supplies the limits and shapes
    the reference's report dict around it.

    ``mode`` selects the lower bound: "out_of_place" (two states, the
    reference model) or "in_place" (one state plus fixed scratch). The
    working set is the device's free memory now, which already excludes
    every state allocated so far; ``reserved_bytes`` is memory admitted but
    not yet allocated (the serving loop's admissions in flight), so no byte
    is counted twice.
    """
    from . import _native

    n = int(n_qubits)
    if n < 0:
        raise ValueError("n_qubits must be non-negative")
    if mode not in ("out_of_place", "in_place"):
        raise ValueError("mode must be 'out_of_place' or 'in_place'")
    memory = state_memory_estimate(n, dtype=dtype)
    override = statevector_unsafe_override_policy(allow_unsafe)
    info = _device_info()
    max_buffer = _limit(info.get("max_buffer_length"))
    working_total = _limit(info.get("max_recommended_working_set_size"))
    rep = _native.preflight(
        n, _native.MSV_C64 if dtype == "complex64" else _native.MSV_C128,
        _native.MSV_PF_OUT_OF_PLACE if mode == "out_of_place" else _native.MSV_PF_IN_PLACE,
        int(reserved_bytes), bool(override["enabled"]), (max_buffer, working_total))
    working_set = rep["working_set_bytes"] if working_total is not None else None
    lower = rep["lower_bound_bytes"]
    failures: List[str] = []
    if rep["failures"] & _native.MSV_PF_FAIL_MAX_BUFFER:
        failures.append("single_state_exceeds_max_buffer_length")
    if rep["failures"] & _native.MSV_PF_FAIL_WORKING_SET:
        failures.append("two_state_lower_bound_exceeds_recommended_working_set"
                        if mode == "out_of_place"
                        else "in_place_lower_bound_exceeds_recommended_working_set")
    if rep["failures"] & _native.MSV_PF_FAIL_QUBITS:
        failures.append("qubit_count_exceeds_engine_limit")
    decision, reason = _DECISIONS[rep["decision"]]

    def _check(v):
        return None if v < 0 else bool(v)

    second_key = ("two_state_lower_bound_within_recommended_working_set" if mode == "out_of_place"
                  else "in_place_lower_bound_within_recommended_working_set")
    return {
        "schema_version": 1,
        "allowed": bool(rep["allowed"]),
        "decision": decision,
        "reason": reason,
        "overridden": bool(rep["overridden"]),
        "override": override,
        "failure_reasons": failures,
        "qubits": n,
        "dtype": dtype,
        "mode": mode,
        "device": {
            "name": info.get("device_name"),
            "architecture": info.get("architecture"),
            "max_buffer_length": max_buffer,
            "max_recommended_working_set_size": working_set,
            "free_bytes_reported": working_total,
            "reserved_bytes": int(reserved_bytes),
        },
        "checks": {
            "single_state_within_max_buffer_length": _check(rep["buffer_check"]),
            second_key: _check(rep["working_check"]),
        },
        "cost_model": {
            **memory,
            "lower_bound_bytes": lower,
            "recommended_working_set_headroom_bytes": (working_set - lower
                                                       if working_set is not None else None),
            "scope": "allocation preflight lower bound",
            "not_included": ["full probability vectors requested by the caller",
                             "memory other processes allocate after the check"],
        },
        "passing_guarantee": (
            "Passing does not guarantee that an arbitrary circuit fits; it only "
            "establishes that the reported lower bounds do not already rule it out."
        ),
    }


def require_statevector_preflight(n_qubits: int, *, dtype: str = STATE_DTYPE,
                                  allow_unsafe: Optional[bool] = None,
                                  mode: str = "out_of_place",
                                  reserved_bytes: int = 0) -> Dict[str, Any]:
    """Return the report or raise StatevectorMemoryError (execution.py:377-407)."""
    report = statevector_preflight(n_qubits, dtype=dtype, allow_unsafe=allow_unsafe, mode=mode,
                                   reserved_bytes=reserved_bytes)
    if report["allowed"]:
        return report
    cost = report["cost_model"]
    device = report["device"]
    gib = 1024.0 ** 3

    def _lim(v):
        return f"{float(v) / gib:.2f} GiB" if v is not None else "unknown"

    bound = "out-of-place" if mode == "out_of_place" else "in-place"
    message = (
        f"Refusing {n_qubits}-qubit {dtype} statevector before allocation: one state is "
        f"{cost['state_bytes'] / gib:.2f} GiB and the {bound} lower bound is "
        f"{cost['lower_bound_bytes'] / gib:.2f} GiB; device max buffer is "
        f"{_lim(device['max_buffer_length'])} and free working set is "
        f"{_lim(device['max_recommended_working_set_size'])}. Reduce qubits, use complex64, "
        f"or set {STATEVECTOR_UNSAFE_OVERRIDE_ENV}=1 to deliberately bypass this preflight."
    )
    raise StatevectorMemoryError(message, report)


# ---------------------------------------------------------------------------
# runtime status and execution plan
# ---------------------------------------------------------------------------


def hip_runtime_status(n_qubits: Optional[int] = None, *, dtype: str = STATE_DTYPE) -> Dict[str, Any]:
    """Whether the HIP engine can run here (the counterpart of
    metal_runtime_status, execution.py:476-591)."""
    raw = os.environ.get(HIP_KERNELS_ENV)
    policy = "auto_default" if raw is None else raw.strip().lower()
    checks = {"library_loaded": False, "device_present": False, "architecture_gfx950": False,
              "index_width": n_qubits is None or int(n_qubits) <= QUBIT_LIMIT,
              "dtype_supported": dtype in BYTES_PER_AMPLITUDE}
    arch = None
    try:
        from . import _native
        _native.lib()
        checks["library_loaded"] = True
        if _native.device_count() > 0:
            checks["device_present"] = True
            arch = _native.device_info(0)["arch"]
            checks["architecture_gfx950"] = arch.split(":")[0] == "gfx950"
    except Exception:
        pass
    if policy not in ("auto_default",) and policy not in _TRUE_VALUES and policy not in _FALSE_VALUES:
        return {"policy": "invalid", "requested": False, "enabled": False, "checks": checks,
                "reason": f"invalid {HIP_KERNELS_ENV} value", "architecture": arch}
    requested = policy == "auto_default" or policy in _TRUE_VALUES
    enabled = requested and all(checks.values())
    reason = ("HIP kernels selected" if enabled else
              ("disabled by policy" if not requested else
               "missing: " + ", ".join(k for k, v in checks.items() if not v)))
    return {"policy": policy, "requested": requested, "enabled": enabled, "checks": checks,
            "reason": reason, "architecture": arch,
            "kernel_constraints": {"index_bits": INDEX_BITS, "index_qubit_limit": QUBIT_LIMIT}}


def _family(op: Dict[str, Any]) -> str:
    return str(op.get("name", "")).lower().lstrip("_") or "gate"


def build_execution_plan(*, n_qubits: int, operations: List[Dict[str, Any]],
                         optimized_operations: List[Dict[str, Any]],
                         statevector_preflight_data: Optional[Dict[str, Any]] = None,
                         dtype: str = STATE_DTYPE, basis: bool = True) -> Dict[str, Any]:
    """Launch-accounting report (execution.py:783-1029).

    The prediction is the native planner's own: the optimized operations run
    on a virtual state of the same width and dtype (msv_state_create_virtual),
    which plans, defers and logs every kernel exactly as the real state will
    but launches nothing. `basis`: the real state starts from a deferred basis
    state (after a reset). selected_custom_kernels lists, per dispatched
    entry, the kernels it launches (deferred layers launch none: their work
    rides in a later entry's sweep); the last entry is the flush of deferred
    work that a read of the state triggers. Observed dispatches are filled
    in as the real state runs (record_observed_dispatches, mark_synchronized).
    """
    from . import _native as nat
    from .dispatch import planned_run

    n = int(n_qubits)
    status = hip_runtime_status(n, dtype=dtype)
    memory = state_memory_estimate(n, dtype)
    per_op: list = []
    run = planned_run(list(optimized_operations), n, dtype, basis=basis, per_op=per_op)
    sweep_names = {nat.KERNEL_KINDS[k] for k in nat.SWEEP_KINDS}
    selected = []
    for idx, kinds in per_op:
        fam = ("deferred_work_flush" if idx is None
               else "+".join(sorted({_family(optimized_operations[i]) for i in idx})))
        selected.append({
            "optimized_operation_indices": idx,
            "kernel_family": fam,
            "expected_launches": sum(1 for k in kinds if k in sweep_names),
            "concrete_kernels": list(kinds),
        })
    matched = [
        {"optimized_operation_index": i, "kernel_family": _family(op)}
        for i, op in enumerate(optimized_operations) if str(op.get("name", "")).startswith("_")
    ]
    launches = sum(item["expected_launches"] for item in selected)
    sweep_bytes = 2 * memory["state_bytes"]
    return {
        "schema_version": 2,
        "engine": "mettleq_amd HIP statevector (gfx950)",
        "qubits": n,
        "dtype": dtype,
        "input_operation_count": len(operations),
        "optimized_operation_count": len(optimized_operations),
        "matched_structured_patterns": matched,
        "pattern_counts": dict(Counter(m["kernel_family"] for m in matched)),
        "selected_custom_kernels": selected,
        "selected_concrete_kernels": sorted({k for item in selected for k in item["concrete_kernels"]}),
        "expected_custom_kernel_launches": launches,
        "capabilities": status,
        "memory": memory,
        "statevector_preflight": statevector_preflight_data,
        "execution_cost_model": {
            "scope": "full-state sweeps planned by the native planner (virtual state)",
            "state_bytes": memory["state_bytes"],
            "bytes_per_sweep": sweep_bytes,
            "predicted_sweeps": run["sweeps"],
            "predicted_input_output_bytes": run["sweeps"] * sweep_bytes,
        },
        "observed_custom_dispatches": [],
        "observed_custom_kernel_launches": None,
        "observed_sweeps": None,
        "execution_status": "planned",
        "synchronized": False,
    }


def record_observed_dispatches(plan: Optional[Dict[str, Any]], per_op: list) -> None:
    """Kernels the real state launched per dispatched entry (its dispatch
    log): the counterpart of record_custom_dispatch (execution.py:1032-1049)."""
    if plan is None:
        return
    for idx, kinds in per_op:
        plan["observed_custom_dispatches"].append(
            {"optimized_operation_indices": idx, "concrete_kernels": list(kinds)})


def mark_synchronized(plan: Optional[Dict[str, Any]], observed_sweeps: Optional[int] = None,
                      flush_kernels: Optional[list] = None) -> None:
    """execution.py:1094-1111: the state has been synchronized; `flush_kernels`
    are the kernels the synchronize launched for deferred work."""
    if plan is None:
        return
    from . import _native as nat
    if flush_kernels is not None:
        plan["observed_custom_dispatches"].append(
            {"optimized_operation_indices": None, "concrete_kernels": list(flush_kernels)})
    sweep_names = {nat.KERNEL_KINDS[k] for k in nat.SWEEP_KINDS}
    plan["observed_custom_kernel_launches"] = sum(
        1 for d in plan["observed_custom_dispatches"] for k in d["concrete_kernels"] if k in sweep_names)
    plan["execution_status"] = "synchronized"
    plan["synchronized"] = True
    if observed_sweeps is not None:
        plan["observed_sweeps"] = int(observed_sweeps)
