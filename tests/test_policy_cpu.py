"""Execution selection and SDK engine recommendation for one MI355X
(mettleq_amd.planning, mettleq_amd.integrations.policy), mirroring the
reference's planning and policy tests (tests/mlxQExecutionPlanningTest.py:
19-113, tests/mlxQSdkPolicyTest.py) with this engine's calibration."""
import pytest

import mettleq_amd.execution as execution
import mettleq_amd.planning as planning
from mettleq_amd.integrations.policy import profile_operations, recommend_sdk_engine

GIB = 1 << 30


@pytest.fixture(autouse=True)
def _limits(monkeypatch):
    monkeypatch.setattr(execution, "_device_info", lambda: {
        "device_name": "MI355X (test limits)", "architecture": "gfx950",
        "max_buffer_length": 288 * 10**9, "max_recommended_working_set_size": 280 * GIB})


def _layer(n):
    return ([{"name": "U3", "wires": [q], "parameters": [0.1, 0.2, 0.3]} for q in range(n)]
            + [{"name": "CNOT", "wires": [q, q + 1], "parameters": []} for q in range(0, n - 1, 2)])


def test_cpu_below_and_gpu_at_the_crossover(monkeypatch):
    monkeypatch.setattr(planning, "_gpu_available", lambda: True)
    small = planning.select_execution(5, [], statevector_gpu_min_qubits=8)
    large = planning.select_execution(8, [], statevector_gpu_min_qubits=8)
    assert small.selected_method == "statevector" and small.selected_device == "cpu"
    assert large.selected_device == "gpu"
    assert "not combined into a speed claim" in large.cpu_gpu_policy


def test_default_crossover_is_the_mi355x_calibration(monkeypatch):
    monkeypatch.setattr(planning, "_gpu_available", lambda: True)
    assert planning.DEFAULT_STATEVECTOR_GPU_MIN_QUBITS == 14
    assert planning.select_execution(13, []).selected_device == "cpu"
    assert planning.select_execution(14, []).selected_device == "gpu"


def test_no_gpu_and_explicit_requests(monkeypatch):
    monkeypatch.setattr(planning, "_gpu_available", lambda: False)
    assert planning.select_execution(20, []).selected_device == "cpu"
    assert "automatic_gpu_unavailable" in planning.select_execution(20, []).reason
    with pytest.raises(RuntimeError):
        planning.select_execution(20, [], device="gpu")
    with pytest.raises(ValueError, match="method must be"):
        planning.select_execution(4, [], method="hybrid")
    with pytest.raises(ValueError, match="device must be"):
        planning.select_execution(4, [], device="neural-engine")
    sel = planning.select_execution(4, [], method="mps")
    assert sel.selected_method == "matrix_product_state"
    assert sel.reason.startswith("explicit_matrix_product_state_request")


def _local_chain(n):
    return [{"name": "CNOT", "wires": [w, w + 1], "parameters": []} for w in range(n - 1)]


def test_automatic_mps_requires_opt_in_and_local_circuits(monkeypatch):
    """The reference's rules (tests/mlxQExecutionPlanningTest.py:42-70), with
    the MPS selected on the GPU, where this engine runs its SVDs."""
    monkeypatch.setattr(planning, "_gpu_available", lambda: True)
    compatible = planning.select_execution(12, _local_chain(12), allow_approximation=True,
                                           automatic_mps_min_qubits=12)
    exact = planning.select_execution(12, _local_chain(12), allow_approximation=False,
                                      automatic_mps_min_qubits=12)
    nonlocal_circuit = planning.select_execution(12, [{"name": "CNOT", "wires": [0, 11], "parameters": []}],
                                                 allow_approximation=True, automatic_mps_min_qubits=12)
    assert compatible.selected_method == "matrix_product_state"
    assert compatible.selected_device == "gpu" and compatible.mps_svd_device == "gpu"
    assert exact.selected_method == "statevector"
    assert nonlocal_circuit.selected_method == "statevector"
    assert "contains_non_adjacent_two_qubit_operation" in nonlocal_circuit.mps_compatibility_reasons
    deep = planning.select_execution(4, _local_chain(4) * 5, allow_approximation=True, automatic_mps_min_qubits=2)
    assert "entangling_operation_count_exceeds_auto_mps_limit" in deep.mps_compatibility_reasons
    # a statevector the preflight refuses falls back to MPS only when allowed
    assert planning.select_execution(40, _local_chain(40), allow_approximation=True,
                                     automatic_mps_min_qubits=64).selected_method == "matrix_product_state"
    with pytest.raises(ValueError, match="runs on the GPU"):
        planning.select_execution(3, [], method="mps", device="cpu")
    with pytest.raises(TypeError):
        planning.select_execution(3, [], allow_approximation="yes")


def test_refused_statevector_keeps_the_exact_request(monkeypatch):
    monkeypatch.setattr(planning, "_gpu_available", lambda: True)
    sel = planning.select_execution(38, [])
    assert sel.selected_method == "statevector"
    assert sel.reason.startswith("automatic_exact_request_has_no_allowed_approximate_fallback")
    assert sel.statevector_preflight["decision"] == "refused"


def test_profile_reports_gate_mix_topology_and_planned_launches():
    profile = profile_operations(8, _layer(8))
    assert profile.operations == 12
    assert profile.single_qubit_operations == 8
    assert profile.topology_class == "nearest_neighbor"
    assert profile.estimated_fusion_coverage == pytest.approx(1.0)
    # the planner: one 1q-layer product write from |0...0>, the CNOTs as one
    # deferred permutation flushed at the end
    assert profile.planned_kernel_launches >= 1


def test_small_circuits_stay_on_the_sdk_cpu_engine():
    assert recommend_sdk_engine("qiskit", 12, _layer(12)).engine == "qiskit_aer_cpu_statevector"
    assert recommend_sdk_engine("pennylane", 12, _layer(12)).engine == "pennylane_lightning_cpu"


def test_gpu_above_the_crossover_and_double_precision_stays_on_gpu():
    d = recommend_sdk_engine("qiskit", 22, _layer(22), output_contract="expectation")
    assert d.engine == "mettleq_gpu_statevector"
    assert d.adjusted_gpu_crossover_qubits <= 16
    dd = recommend_sdk_engine("qiskit", 24, _layer(24), precision="double")
    assert dd.engine == "mettleq_gpu_statevector_complex128"


def test_mps_requests():
    assert recommend_sdk_engine("qiskit", 40, _layer(40), method="mps").engine == "qiskit_aer_cpu_mps"
    assert recommend_sdk_engine("pennylane", 40, _layer(40), method="mps").engine == "mettleq_gpu_mps"


def test_refusal_and_bounded_local_fallback():
    assert recommend_sdk_engine("qiskit", 38, _layer(38)).engine == "refused"
    assert recommend_sdk_engine("qiskit", 38, _layer(38), allow_approximation=True).engine == "qiskit_aer_cpu_mps"


def test_engine_targets():
    from mettleq_amd.integrations.policy import engine_target
    assert engine_target("qiskit_aer_cpu_statevector").sdk_cpu
    assert engine_target("pennylane_lightning_cpu").sdk_cpu
    assert engine_target("qiskit_aer_cpu_mps").method == "matrix_product_state"
    g = engine_target("mettleq_gpu_statevector")
    assert not g.sdk_cpu and (g.method, g.dtype) == ("statevector", "complex64")
    assert engine_target("mettleq_gpu_statevector_complex128").dtype == "complex128"
    m = engine_target("mettleq_gpu_mps")
    assert not m.sdk_cpu and m.method == "matrix_product_state"
    with pytest.raises(ValueError):
        engine_target("refused")


def test_adaptive_plan_groups_circuits_by_engine():
    """The per-batch part of the reference's adaptive front ends
    (qiskit.py:244-316, pennylane.py:690-738): one decision per circuit, the
    indices grouped by engine in order of first use."""
    from mettleq_amd.integrations.policy import adaptive_plan
    batch = [(22, _layer(22), "samples"), (10, _layer(10), "samples"), (24, _layer(24), "expectation"),
             (12, _layer(12), "statevector")]
    decisions, groups = adaptive_plan("qiskit", batch)
    engines = [d.engine for d in decisions]
    assert engines == ["mettleq_gpu_statevector", "qiskit_aer_cpu_statevector",
                       "mettleq_gpu_statevector", "qiskit_aer_cpu_statevector"]
    assert list(groups) == ["mettleq_gpu_statevector", "qiskit_aer_cpu_statevector"]
    assert groups["mettleq_gpu_statevector"] == [0, 2] and groups["qiskit_aer_cpu_statevector"] == [1, 3]
    # without the SDK's CPU engine everything stays on this one
    _, only = adaptive_plan("pennylane", batch, fallback_available=False)
    assert list(only) == ["mettleq_gpu_statevector"] and only["mettleq_gpu_statevector"] == [0, 1, 2, 3]
    # double precision keeps complex128 on the GPU
    d2, _ = adaptive_plan("qiskit", batch[:1], precision="double")
    assert d2[0].engine == "mettleq_gpu_statevector_complex128"


def test_adaptive_plan_refuses_before_running_anything():
    from mettleq_amd.integrations.policy import adaptive_plan
    with pytest.raises(ValueError, match="dense preflight refused"):
        adaptive_plan("qiskit", [(10, _layer(10), "samples"), (38, _layer(38), "samples")])
    decisions, groups = adaptive_plan("qiskit", [(38, _layer(38), "samples")], allow_approximation=True)
    assert list(groups) == ["qiskit_aer_cpu_mps"]
