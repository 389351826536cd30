"""Statevector preflight on the CPU with injected device limits, mirroring
the reference's fake-limit tests (tests/mlxQExecutionPlanTest.py:57-141)."""
import pytest

import mettleq_amd.execution as execution
from mettleq_amd.execution import (
    STATEVECTOR_UNSAFE_OVERRIDE_ENV,
    StatevectorMemoryError,
    require_statevector_preflight,
    state_memory_estimate,
    statevector_preflight,
)

GIB = 1 << 30
MI355X_HBM = 288 * 10**9


def _limits(monkeypatch, *, max_buffer, working_set):
    monkeypatch.setattr(execution, "_device_info", lambda: {
        "device_name": "Constrained test GPU", "architecture": "test",
        "max_buffer_length": max_buffer, "max_recommended_working_set_size": working_set})


def test_memory_estimate():
    m = state_memory_estimate(10)
    assert m["state_bytes"] == 8 * (1 << 10)
    assert m["minimum_input_plus_output_bytes"] == 16 * (1 << 10)
    assert state_memory_estimate(33, "complex128")["state_bytes"] == 128 * GIB
    assert state_memory_estimate(34)["state_bytes"] == 128 * GIB


def test_refuses_before_allocation(monkeypatch):
    _limits(monkeypatch, max_buffer=4096, working_set=12000)
    report = statevector_preflight(10)
    assert not report["allowed"]
    assert report["decision"] == "refused"
    assert report["failure_reasons"] == [
        "single_state_exceeds_max_buffer_length",
        "two_state_lower_bound_exceeds_recommended_working_set",
    ]
    assert report["cost_model"]["state_bytes"] == 8192
    assert report["cost_model"]["minimum_input_plus_output_bytes"] == 16384
    with pytest.raises(StatevectorMemoryError) as caught:
        require_statevector_preflight(10)
    assert caught.value.report == report
    assert "before allocation" in str(caught.value)
    assert STATEVECTOR_UNSAFE_OVERRIDE_ENV in str(caught.value)
    assert isinstance(caught.value, MemoryError)


def test_simulator_refuses_without_touching_the_gpu(monkeypatch):
    """The preflight raises before any native call."""
    _limits(monkeypatch, max_buffer=4096, working_set=12000)
    from mettleq_amd.sim import StateVectorSimulator
    import mettleq_amd.native_state as ns

    def boom(*a, **k):
        raise AssertionError("native allocation attempted")

    monkeypatch.setattr(ns.NativeState, "__init__", boom)
    with pytest.raises(StatevectorMemoryError):
        StateVectorSimulator(10)


def test_override_is_explicit_and_observable(monkeypatch):
    _limits(monkeypatch, max_buffer=4096, working_set=12000)
    report = statevector_preflight(10, allow_unsafe=True)
    assert report["allowed"] and report["overridden"]
    assert report["decision"] == "allowed_with_unsafe_override"
    assert report["override"]["source"] == "constructor_argument"
    assert report["failure_reasons"]


def test_environment_fails_closed(monkeypatch):
    _limits(monkeypatch, max_buffer=4096, working_set=12000)
    monkeypatch.setenv(STATEVECTOR_UNSAFE_OVERRIDE_ENV, "typo")
    with pytest.raises(ValueError, match=STATEVECTOR_UNSAFE_OVERRIDE_ENV):
        statevector_preflight(10)
    monkeypatch.setenv(STATEVECTOR_UNSAFE_OVERRIDE_ENV, "1")
    assert statevector_preflight(10)["overridden"]
    assert not statevector_preflight(10, allow_unsafe=False)["allowed"]


def test_missing_limits_are_unverified(monkeypatch):
    _limits(monkeypatch, max_buffer=0, working_set=0)
    report = statevector_preflight(20)
    assert report["allowed"]
    assert report["decision"] == "allowed_with_unverified_device_limits"
    assert report["checks"] == {
        "single_state_within_max_buffer_length": None,
        "two_state_lower_bound_within_recommended_working_set": None,
    }
    with pytest.raises(ValueError):
        statevector_preflight(4, dtype="complex256")


def test_mi355x_sized_decisions(monkeypatch):
    """288 GB (268 GiB) HBM: 34q c64 and 33q c128 (128 GiB each) fit in place
    and, just, out of place (256 GiB); 35q c64 (256 GiB) fits only in place
    and is refused out of place (512 GiB); 36q is refused either way."""
    _limits(monkeypatch, max_buffer=MI355X_HBM, working_set=MI355X_HBM - 2 * GIB)
    for n, dt in ((34, "complex64"), (33, "complex128")):
        assert statevector_preflight(n, dtype=dt, mode="in_place")["allowed"]
        assert statevector_preflight(n, dtype=dt, mode="out_of_place")["allowed"]
    assert statevector_preflight(35, mode="in_place")["allowed"]
    r = statevector_preflight(35, mode="out_of_place")
    assert r["decision"] == "refused"
    assert not statevector_preflight(36, mode="in_place")["allowed"]
    assert statevector_preflight(30)["decision"] == "allowed_within_reported_lower_bounds"
    # both config-3 states resident together: 33q c64 (64 GiB) + 33q c128 (128 GiB)
    assert statevector_preflight(33, dtype="complex128", mode="in_place",
                                 reserved_bytes=64 * GIB)["allowed"]


def test_reserved_bytes_admission(monkeypatch):
    _limits(monkeypatch, max_buffer=MI355X_HBM, working_set=100 * GIB)
    assert statevector_preflight(32, mode="in_place")["allowed"]  # 32 GiB
    r = statevector_preflight(32, mode="in_place", reserved_bytes=80 * GIB)
    assert not r["allowed"]
    assert r["failure_reasons"] == ["in_place_lower_bound_exceeds_recommended_working_set"]


def test_qubit_limit(monkeypatch):
    _limits(monkeypatch, max_buffer=0, working_set=0)
    r = statevector_preflight(41)
    assert not r["allowed"]
    assert "qubit_count_exceeds_engine_limit" in r["failure_reasons"]
