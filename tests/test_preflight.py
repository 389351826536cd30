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


def test_live_free_memory_counts_resident_states_once(monkeypatch):
    """The working set is the device's free memory now (total - resident):
    a 33q complex64 state (64 GiB) already allocated lowers it by 64 GiB, and
    a 33q complex128 state (128 GiB) still fits beside it in place on 288 GB;
    a third 64 GiB state does not fit out of place."""
    resident = 64 * GIB
    _limits(monkeypatch, max_buffer=MI355X_HBM, working_set=MI355X_HBM - resident)
    r = statevector_preflight(33, dtype="complex128", mode="in_place")
    assert r["decision"] == "allowed_within_reported_lower_bounds"
    assert r["device"]["max_recommended_working_set_size"] == MI355X_HBM - resident
    resident += 128 * GIB
    _limits(monkeypatch, max_buffer=MI355X_HBM, working_set=MI355X_HBM - resident)
    assert statevector_preflight(33, mode="in_place")["allowed"]  # 64 GiB + 8 MiB of ~76 GiB
    r = statevector_preflight(33, mode="out_of_place")
    assert r["decision"] == "refused"
    assert r["failure_reasons"] == ["two_state_lower_bound_exceeds_recommended_working_set"]


def test_pool_admission_matches_device_capacity(monkeypatch):
    """StatePool on a simulated 100 GiB device whose free memory drops with
    every allocated state: 8 GiB states are admitted until free memory runs
    out (12, each needing 8 GiB + 8 MiB of scratch), not half as many as a
    double count of resident bytes would give."""
    from mettleq_amd.serving import StatePool
    total = 100 * GIB
    held = []

    class Big:
        def __init__(self, n):
            self.n = n
            held.append(self)

        def close(self):
            held.remove(self)

    def dev():
        free = total - sum((1 << s.n) * 8 for s in held)
        return {"device_name": "sim", "architecture": "test", "max_buffer_length": total,
                "max_recommended_working_set_size": free}

    monkeypatch.setattr(execution, "_device_info", dev)
    pool = StatePool(factory=Big)  # the default in-place complex64 preflight
    n = 30  # 8 GiB complex64
    states = [pool.acquire(n)]
    while True:
        r = statevector_preflight(n, mode="in_place")
        if not r["allowed"]:
            assert r["decision"] == "refused"
            break
        states.append(pool.acquire(n))
    assert len(states) == 12  # 12 x 8 GiB = 96 GiB; the 13th needs 8 GiB + 8 MiB of 4 GiB
    assert pool.resident_bytes == 96 * GIB


def test_native_preflight_through_ctypes():
    """msv_preflight itself, as a non-Python binder calls it: injected limits,
    decision codes, failure bits and the refusal message."""
    import ctypes
    from mettleq_amd import _native as nat
    lib = nat.lib()
    rep = nat.msv_preflight_report()
    lim = nat.msv_preflight_limits(4096, 12000)
    rc = lib.msv_preflight(10, nat.MSV_C64, nat.MSV_PF_OUT_OF_PLACE, 0, 0, 0,
                           ctypes.byref(lim), ctypes.byref(rep))
    assert rc == nat.MSV_E_REFUSED and rep.decision == nat.MSV_PF_REFUSED and not rep.allowed
    assert rep.failures == nat.MSV_PF_FAIL_MAX_BUFFER | nat.MSV_PF_FAIL_WORKING_SET
    assert rep.state_bytes == 8192 and rep.lower_bound_bytes == 16384
    msg = nat.last_error()
    assert "before allocation" in msg and STATEVECTOR_UNSAFE_OVERRIDE_ENV in msg
    rc = lib.msv_preflight(10, nat.MSV_C64, nat.MSV_PF_OUT_OF_PLACE, 0, 1, 0,
                           ctypes.byref(lim), ctypes.byref(rep))
    assert rc == nat.MSV_OK and rep.decision == nat.MSV_PF_ALLOWED_OVERRIDE and rep.overridden
    lim = nat.msv_preflight_limits(0, 0)
    rc = lib.msv_preflight(20, nat.MSV_C128, nat.MSV_PF_IN_PLACE, 0, 0, 0,
                           ctypes.byref(lim), ctypes.byref(rep))
    assert rc == nat.MSV_OK and rep.decision == nat.MSV_PF_ALLOWED_UNVERIFIED
    assert rep.buffer_check == -1 and rep.working_check == -1
    assert rep.lower_bound_bytes == (1 << 20) * 16 + (8 << 20)
    lim = nat.msv_preflight_limits(MI355X_HBM, 100 * GIB)
    rc = lib.msv_preflight(32, nat.MSV_C64, nat.MSV_PF_IN_PLACE, 80 * GIB, 0, 0,
                           ctypes.byref(lim), ctypes.byref(rep))
    assert rc == nat.MSV_E_REFUSED and rep.failures == nat.MSV_PF_FAIL_WORKING_SET
    assert rep.working_set_bytes == 20 * GIB
    rc = lib.msv_preflight(41, nat.MSV_C64, nat.MSV_PF_IN_PLACE, 0, 1, 0,
                           ctypes.byref(nat.msv_preflight_limits(0, 0)), ctypes.byref(rep))
    assert rc == nat.MSV_E_REFUSED and rep.failures == nat.MSV_PF_FAIL_QUBITS and not rep.overridden
    assert lib.msv_preflight(-1, 0, 0, 0, 0, 0, None, ctypes.byref(rep)) == nat.MSV_E_INVALID


def test_reserved_bytes_admission(monkeypatch):
    _limits(monkeypatch, max_buffer=MI355X_HBM, working_set=100 * GIB)
    assert statevector_preflight(32, mode="in_place")["allowed"]  # 32 GiB
    # 80 GiB admitted by other workers but not yet allocated
    r = statevector_preflight(32, mode="in_place", reserved_bytes=80 * GIB)
    assert not r["allowed"]
    assert r["failure_reasons"] == ["in_place_lower_bound_exceeds_recommended_working_set"]


def test_qubit_limit(monkeypatch):
    _limits(monkeypatch, max_buffer=0, working_set=0)
    r = statevector_preflight(41)
    assert not r["allowed"]
    assert "qubit_count_exceeds_engine_limit" in r["failure_reasons"]
