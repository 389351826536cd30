"""GPU tests of two benchmark configurations end to end.

* 20 qubits complex64, 10 brick layers (U3 on every qubit, then CZ on the
  even/odd brick pairs), through the drop-in `Device` (fusion + dispatch):
  the FP64 probability vector against the complex128 oracle within 5e-6;
  1000 shots (seed base + 1) bit-exact against the oracle's sampler on the
  GPU's amplitudes; and the same circuit in complex128 end to end, whose
  1000 shots equal the oracle's own complex128 run shot for shot.
* The memory-limit run (brick layers, depth 20): 33 qubits complex64 and
  33 qubits complex128 (64 + 128 GiB) resident together, 1 - F <= 1e-6;
  34 qubits complex64 (128 GiB) in place, norm kept and a circuit followed
  by its inverse returning |0...0>; preflight decisions from the device's
  own limits.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GIB = 1 << 30


def _brick_ops(n, depth):
    from mettleq_amd.workloads import brick_20q, circuit_ops
    return circuit_ops(brick_20q(n, depth), n)


_ORACLE = {}


def _oracle_brick20():
    """The complex128 oracle's state of the 20q brick circuit (computed once)."""
    if "brick20" not in _ORACLE:
        from oracle import sv_oracle as O
        _ORACLE["brick20"] = O.run_ops(20, _brick_ops(20, 10))
    return _ORACLE["brick20"]


def _probs_fp64(amps):
    a = np.asarray(amps, dtype=np.complex128)
    return a.real * a.real + a.imag * a.imag


def test_config1_brick20_probabilities_and_shots(gpu):
    from mettleq_amd.device import Device
    from mettleq_amd.workloads import BASE_SEED
    from oracle import sv_oracle as O
    n = 20
    ops = _brick_ops(n, 10)
    want = _oracle_brick20()
    dev = Device(n)
    try:
        dev.execute(ops)
        st = dev.native
        probs = st.probabilities()
        shots = st.sample(1000, seed=BASE_SEED + 1)
        got = st.to_numpy()
    finally:
        dev.close()
    assert np.max(np.abs(probs - np.abs(want) ** 2)) <= 5e-6
    assert abs(probs.sum() - 1.0) < 1e-5
    # bit-exact sampling: the oracle's sampler on the GPU's own amplitudes
    ref = O.sample_indices(O.amp_probs_c64(got), 1000, BASE_SEED + 1, 0)
    assert np.array_equal(shots, ref)


def test_config1_brick20_complex128_end_to_end_shots(gpu):
    """complex128 on the GPU against the oracle's complex128 run: amplitudes
    within 1e-12 and the 1000 shots identical shot for shot (both sample the
    same seeded Philox stream over FP64 probabilities that agree to ~1e-16)."""
    from mettleq_amd.device import Device
    from mettleq_amd.workloads import BASE_SEED
    from oracle import sv_oracle as O
    n = 20
    ops = _brick_ops(n, 10)
    want = _oracle_brick20()
    dev = Device(n, dtype="complex128")
    try:
        dev.execute(ops)
        st = dev.native
        got = st.to_numpy()
        shots = st.sample(1000, seed=BASE_SEED + 1)
    finally:
        dev.close()
    assert np.max(np.abs(got - want)) < 1e-12
    ref = O.sample_indices(_probs_fp64(want), 1000, BASE_SEED + 1, 0)
    assert np.array_equal(shots, ref)


def _run_brick(st, n, depth, inverse_after=False):
    from mettleq_amd import _native as nat
    from mettleq_amd import gates as G
    from mettleq_amd.workloads import brick_20q
    layers = brick_20q(n, depth)
    st.reset()

    def apply(layer, adjoint):
        if layer["type"] == "u3":
            mats = [G.single_qubit_matrix("U3", a) for a in layer["angles"]]
            st.apply_1q_layer([m.conj().T for m in mats] if adjoint else mats)
        else:
            b = layer["bonds"]
            st.apply_phase_layer([nat.MSV_PHASE_CPHASE] * len(b), b, [math.pi] * len(b))

    for layer in layers:
        apply(layer, False)
    if inverse_after:
        for layer in reversed(layers):
            apply(layer, True)


def test_config3_preflight_from_device_limits(gpu):
    from mettleq_amd import _native as nat
    from mettleq_amd.execution import statevector_preflight
    info = nat.device_info(0)
    if info["free_bytes"] < 200 * GIB:
        pytest.skip(f"needs a 288 GB device, {info['free_bytes'] / GIB:.0f} GiB free")
    assert info["max_alloc_bytes"] >= 128 * GIB
    for n, dt, mode in ((34, "complex64", "in_place"), (33, "complex128", "in_place"),
                        (33, "complex64", "out_of_place")):
        r = statevector_preflight(n, dtype=dt, mode=mode)
        assert r["decision"] == "allowed_within_reported_lower_bounds", (n, dt, mode, r)
    for n, mode in ((35, "out_of_place"), (36, "in_place")):
        r = statevector_preflight(n, mode=mode)
        assert r["decision"] == "refused" and not r["allowed"], (n, mode)


def test_config3_33q_complex64_vs_complex128_fidelity(gpu):
    from mettleq_amd import _native as nat
    from mettleq_amd.execution import require_statevector_preflight
    from mettleq_amd.native_state import NativeState
    if nat.device_info(0)["free_bytes"] < 200 * GIB:
        pytest.skip("needs a 288 GB device")
    n, depth = 33, 20
    require_statevector_preflight(n, dtype="complex128", mode="in_place")
    with NativeState(n, "complex128") as s128:
        _run_brick(s128, n, depth)
        require_statevector_preflight(n, dtype="complex64", mode="in_place")
        with NativeState(n, "complex64") as s64:
            _run_brick(s64, n, depth)
            ip = s64.inner(s128)
            f = abs(ip) ** 2 / (s64.norm2() * s128.norm2())
            print(f"33q complex64 vs complex128: 1-F = {1 - f:.3e}", flush=True)
            assert 1.0 - f <= 1e-6
            assert abs(s128.norm2() - 1.0) < 1e-10


def test_config3_34q_complex64_in_place(gpu):
    from mettleq_amd import _native as nat
    from mettleq_amd.execution import require_statevector_preflight
    from mettleq_amd.native_state import NativeState
    if nat.device_info(0)["free_bytes"] < 200 * GIB:
        pytest.skip("needs a 288 GB device")
    n = 34
    require_statevector_preflight(n, dtype="complex64", mode="in_place")
    with NativeState(n, "complex64") as st:
        _run_brick(st, n, 20)
        assert abs(st.norm2() - 1.0) < 1e-5
        # a shorter circuit and its exact inverse return |0...0>
        _run_brick(st, n, 4, inverse_after=True)
        a0 = st.to_numpy(0, 4)
        assert abs(abs(a0[0]) ** 2 - 1.0) < 1e-5
        assert np.max(np.abs(a0[1:])) < 1e-3
