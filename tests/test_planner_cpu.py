"""The native planner on a virtual state (msv_state_create_virtual), on the
CPU: the sweeps and kernels a real state would launch, with no device
memory. These are the execution plan's predictions (reference
execution.py:783-1029); tests/test_device_gpu.py checks them against the
dispatch log of real runs."""
import math
from collections import Counter

import numpy as np
import pytest

from mettleq_amd import _native as nat
from mettleq_amd.native_state import NativeState


def _layered(n, layers):
    import bench
    from mettleq_amd.layers import compose_inverse_affine
    from mettleq_amd.workloads import layered_30q
    lay = {"chain": compose_inverse_affine([{"name": "CNOT", "wires": [i, i + 1]} for i in range(n - 1)], n)}
    with NativeState.planner(n) as st:
        bench.run_circuit(st, layered_30q(n, layers), n, lay)
        st.sync()
        return st.sweep_count(), Counter(st.dispatch_log())


def test_layered_30q_plan_matches_the_measured_step():
    """40 layers at 30 qubits: 38 sweeps, the kernel mix of the rocprofv3
    trace of the same step (profiles/k_rot_30q_layered_circuit.md)."""
    sweeps, kinds = _layered(30, 40)
    assert sweeps == 38
    assert kinds == Counter({"k_rot": 18, "k_qft": 8, "k_rot_relabel": 7, "k_qft_fused": 4,
                             "k_product": 1, "k_rot_fused": 1})


def test_pending_tail_sweep_counts_at_30q_without_a_gpu():
    """U3 / RZZ ring / CNOT chain: two sweeps per U3 layer with pending tails
    (the last tail runs at the read), three without."""
    from mettleq_amd import gates as G
    from mettleq_amd.layers import compose_inverse_affine
    from mettleq_amd.workloads import ring_bonds
    n = 30
    rng = np.random.default_rng(1)
    rows, c = compose_inverse_affine([{"name": "CNOT", "wires": [i, i + 1]} for i in range(n - 1)], n)
    bonds = ring_bonds(n)
    for tail, want in ((1, [2, 2, 2, 1]), (0, [3, 3, 3, 0])):
        old = nat.get_tuning(nat.MSV_TUNE_DEFER_TAIL)
        nat.set_tuning(nat.MSV_TUNE_DEFER_TAIL, tail)
        try:
            with NativeState.planner(n) as st:
                counts = []
                for i in range(3):
                    if i:
                        st.apply_phase_layer([nat.MSV_PHASE_ZZ] * len(bonds), bonds, list(rng.uniform(0, 3, len(bonds))))
                        st.apply_xor_affine(rows, c)
                    s0 = st.sweep_count()
                    st.apply_1q_layer([G.single_qubit_matrix("U3", rng.uniform(0, 6, 3)) for _ in range(n)])
                    counts.append(st.sweep_count() - s0)
                s0 = st.sweep_count()
                st.sync()
                counts.append(st.sweep_count() - s0)
        finally:
            nat.set_tuning(nat.MSV_TUNE_DEFER_TAIL, old)
        assert counts == want


def test_virtual_state_holds_no_amplitudes():
    with NativeState.planner(34) as st:  # 128 GiB: nothing is allocated
        st.apply_1q_layer([np.eye(2)] * 34)
        assert st.sweep_count() >= 1
    with NativeState.planner(12) as st:
        for fn in (st.probabilities, st.norm2, lambda: st.sample(4), lambda: st.to_numpy(0, 4)):
            with pytest.raises(NotImplementedError):
                fn()
        assert st.device_ptr() == 0


def test_execution_plan_from_the_planner():
    """build_execution_plan: per dispatched entry the kernels it launches
    (deferred layers none), a final flush entry, and totals that add up."""
    from mettleq_amd.execution import build_execution_plan
    from mettleq_amd.fusion import fuse_operations
    from mettleq_amd.workloads import brick_20q, circuit_ops
    n = 20
    ops = circuit_ops(brick_20q(n, 10), n)
    fused = fuse_operations(ops, n)
    plan = build_execution_plan(n_qubits=n, operations=ops, optimized_operations=fused, basis=True)
    sel = plan["selected_custom_kernels"]
    assert sel[-1]["kernel_family"] == "deferred_work_flush"
    assert plan["expected_custom_kernel_launches"] == sum(e["expected_launches"] for e in sel)
    # the reset is deferred: the first 1q layer writes its product state
    assert sel[0]["concrete_kernels"] == ["k_product"]
    # CZ bricks are deferred into the next layer's first sweep (no launch of their own)
    cz = [e for e in sel if e["kernel_family"] == "diaglayer"]
    assert cz and all(e["expected_launches"] == 0 for e in cz[:-1])
    sweeps = plan["execution_cost_model"]["predicted_sweeps"]
    kinds = [k for e in sel for k in e["concrete_kernels"]]
    assert sweeps == sum(1 for k in kinds if k != "k_product")
    assert plan["observed_custom_dispatches"] == [] and plan["observed_sweeps"] is None
