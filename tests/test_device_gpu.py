"""The host layers above the C ABI on the GPU: Device (fusion + dispatch),
StateVectorSimulator, and the SDK-facing helpers of integrations._common,
against the oracle (reference semantics)."""
import math

import numpy as np
import pytest

from oracle import sv_oracle as O

pytestmark = pytest.mark.gpu

TOL = 5e-6


def _random_ops(n, count, seed):
    rng = np.random.default_rng(seed)
    names1 = ["H", "X", "Y", "Z", "S", "SDG", "T", "TDG", "SX", "RX", "RY", "RZ", "U1", "U2", "U3"]
    names2 = ["CNOT", "CZ", "CPHASE", "CRX", "CRY", "CRZ", "CH", "SWAP", "ISWAP", "ZZPHASE",
              "XXPHASE", "YYPHASE"]
    npar = {"RX": 1, "RY": 1, "RZ": 1, "U1": 1, "U2": 2, "U3": 3, "CPHASE": 1, "CRX": 1, "CRY": 1,
            "CRZ": 1, "ZZPHASE": 1, "XXPHASE": 1, "YYPHASE": 1}
    ops = []
    for _ in range(count):
        if n >= 2 and rng.random() < 0.4:
            name = names2[int(rng.integers(len(names2)))]
            w = [int(x) for x in rng.choice(n, 2, replace=False)]
        else:
            name = names1[int(rng.integers(len(names1)))]
            w = [int(rng.integers(n))]
        ops.append({"name": name, "wires": w,
                    "parameters": [float(x) for x in rng.uniform(-3, 3, npar.get(name, 0))]})
    return ops


@pytest.mark.parametrize("n,count,seed", [(2, 40, 0), (5, 120, 1), (9, 200, 2), (14, 160, 3)])
def test_device_execute_random_circuits(gpu, n, count, seed):
    from mettleq_amd.device import Device
    ops = _random_ops(n, count, seed)
    want = O.run_ops(n, ops)
    dev = Device(n)
    try:
        dev.execute(ops)
        dev.synchronize()
        got = dev.native.to_numpy()
    finally:
        dev.close()
    assert np.max(np.abs(got - want)) < TOL


def test_device_structured_workload_and_plan(gpu):
    """A small layered circuit (U3-all / RZZ ring / CNOT chain / QFT) fuses
    into structured layers; the plan reports them and the result matches."""
    from mettleq_amd.device import Device
    from mettleq_amd.workloads import circuit_ops, layered_30q
    n = 12
    ops = circuit_ops(layered_30q(n, 10), n)
    dev = Device(n)
    try:
        dev.execute(ops, report=True)
        dev.synchronize()
        got = dev.native.to_numpy()
        plan = dev.last_execution_plan
    finally:
        dev.close()
    assert np.max(np.abs(got - O.run_ops(n, ops))) < TOL
    assert isinstance(plan, dict)


def test_simulator_api(gpu):
    from mettleq_amd.sim import StateVectorSimulator, iqft, qft
    n = 4
    sim = StateVectorSimulator(n)
    try:
        sim.h(0)
        sim.cnot(0, 1)
        p = sim.probabilities_array([0, 1])
        assert np.allclose(p, [0.5, 0, 0, 0.5], atol=1e-6)
        shots = sim.sample_array(200, [0, 1], rng=np.random.default_rng(1))
        assert set(map(tuple, shots.tolist())) <= {(0, 0), (1, 1)}
        sim.reset()
        rng = np.random.default_rng(4)
        m4 = O.gate_matrix("XXPHASE", [0.7]) @ O.gate_matrix("CRY", [1.1])
        psi = O.zero_state(n)
        for q in range(n):
            u = O.U3(*rng.uniform(0, 6, 3))
            sim.apply_single(u, q)
            psi = O.apply_dense(psi, n, u, [q])
        sim.apply_dense_gate(m4, [2, 0])
        psi = O.apply_dense(psi, n, m4, [2, 0])
        sim.apply_diagonal([1, 1j, -1, np.exp(0.3j)], [3, 1])
        psi = O.apply_dense(psi, n, np.diag([1, 1j, -1, np.exp(0.3j)]), [3, 1])
        sim.apply_controlled_single(O.RY(0.4), 1, 3)
        psi = O.apply_controlled(psi, n, O.RY(0.4), [1], 3)
        sim.apply_zz_layer(0.25, [(0, 1), (2, 3)])
        psi = O.apply_dense(psi, n, O.gate_matrix("ZZPHASE", [0.25]), [0, 1])
        psi = O.apply_dense(psi, n, O.gate_matrix("ZZPHASE", [0.25]), [2, 3])
        qft(sim, [0, 1, 2, 3])
        psi = O.qft(psi, n, [0, 1, 2, 3])
        assert np.max(np.abs(sim.state - psi)) < TOL
        iqft(sim, [0, 1, 2, 3])
        psi = O.iqft(psi, n, [0, 1, 2, 3])
        assert np.max(np.abs(sim.state - psi)) < TOL
        with pytest.raises(ValueError):
            sim.apply_single(O.H(), 4)
    finally:
        sim.close()


def test_common_helpers_bell_values(gpu):
    from mettleq_amd.integrations._common import (apply_global_phase, core_operation, execute_operations,
                                                  marginal_probabilities, pauli_product_expectations,
                                                  sample_bits, statevector_numpy)
    ops = [core_operation("H", [0]), core_operation("CNOT", [0, 1])]
    cache = {}
    dev = execute_operations(2, ops, execution_cache=cache)
    vals = pauli_product_expectations(dev, [{0: "X", 1: "X"}, {0: "Z", 1: "Z"}, {0: "Y", 1: "Y"}])
    assert np.allclose(np.real(vals), [1, 1, -1], atol=1e-6)
    assert np.allclose(marginal_probabilities(dev, [1, 0]), [0.5, 0, 0, 0.5], atol=1e-6)
    bits = sample_bits(dev, 300, [0, 1], np.random.default_rng(3))
    assert set(map(tuple, bits.tolist())) <= {(0, 0), (1, 1)}
    apply_global_phase(dev, 0.5)
    sv = statevector_numpy(dev)
    assert np.allclose(sv, np.exp(0.5j) * np.array([1, 0, 0, 1]) / math.sqrt(2), atol=1e-6)
    # the cached device is reused and reset for the next circuit of that width
    dev2 = execute_operations(2, [core_operation("X", [1])], execution_cache=cache)
    assert dev2 is dev
    assert np.allclose(statevector_numpy(dev2), [0, 1, 0, 0], atol=1e-6)
    with pytest.raises(ValueError):
        core_operation("NOPE", [0])


@pytest.mark.parametrize("phase_name", ["CPHASE", "ZZPHASE"])
@pytest.mark.parametrize("n", [8, 20])
def test_chain_phase_rx_amortization_matches_unfused(gpu, monkeypatch, phase_name, n):
    """The reference's chain-phase + RX amortization case
    (mlxQMetalKernelsTest.py:153-191): fused and unfused runs agree within
    5e-6 and with the oracle; the plan reports one amortized layer and one
    sweep fewer."""
    from mettleq_amd.device import Device
    bonds = [(q, q + 1) for q in range(n - 1)] + ([(n - 1, 0)] if phase_name == "CPHASE" else [])
    ops = ([{"name": phase_name, "wires": list(b), "parameters": [0.23]} for b in bonds]
           + [{"name": "RX", "wires": [w], "parameters": [-0.31]} for w in range(n)])
    outs, plans = [], []
    for flag in ("0", "1"):
        monkeypatch.setenv("METTLEQ_CHAINPHASE_RX_FUSION", flag)
        dev = Device(n)
        try:
            dev.execute(ops, report=True)
            dev.synchronize()
            outs.append(dev.native.to_numpy())
            plans.append(dev.last_execution_plan)
        finally:
            dev.close()
    assert np.max(np.abs(outs[0] - outs[1])) < 5e-6
    assert np.max(np.abs(outs[1] - O.run_ops(n, ops))) < TOL
    assert plans[1]["pattern_counts"] == {"chainphase_rxlayer": 1}
    # both forms reach the engine as the same two calls (the diagonal stays
    # deferred and the RX layer's first sweep reads it): equal launches, as
    # predicted by the planner
    assert plans[1]["expected_custom_kernel_launches"] == plans[0]["expected_custom_kernel_launches"]
    for p in plans:
        assert p["observed_custom_kernel_launches"] == p["expected_custom_kernel_launches"]


_PLAN_ORACLE = {}


@pytest.mark.parametrize("reset", [False, True])
@pytest.mark.parametrize("workload", ["layered", "brick"])
def test_execution_plan_predicted_equals_observed(gpu, workload, reset):
    """The execution plan's prediction (the native planner on a virtual
    state) equals what the real state launches: per dispatched entry the same
    kernels in the same order, the same sweeps, the same final flush
    (reference execution plan and observed dispatches, execution.py:783-1111)."""
    from mettleq_amd.device import Device
    from mettleq_amd.workloads import brick_20q, circuit_ops, layered_30q
    n = 20
    ops = circuit_ops(layered_30q(n, 20) if workload == "layered" else brick_20q(n, 10), n)
    dev = Device(n)
    try:
        if reset:
            dev.reset()
        dev.execute(ops, report=True)
        dev.synchronize()
        got = dev.native.to_numpy()
        plan = dev.last_execution_plan
    finally:
        dev.close()
    if workload not in _PLAN_ORACLE:
        _PLAN_ORACLE[workload] = O.run_ops(n, ops)
    assert np.max(np.abs(got - _PLAN_ORACLE[workload])) < TOL
    pred = [e["concrete_kernels"] for e in plan["selected_custom_kernels"]]
    obs = [e["concrete_kernels"] for e in plan["observed_custom_dispatches"]]
    assert pred == obs
    assert plan["observed_custom_kernel_launches"] == plan["expected_custom_kernel_launches"]
    assert plan["observed_sweeps"] == plan["execution_cost_model"]["predicted_sweeps"]
    assert plan["observed_sweeps"] > 0
