"""Circuit-aware SDK engine recommendation, calibrated for one MI355X.

Restates the reference's inspectable policy (integrations/policy.py:28-248):
every decision carries the calibrated crossover, the circuit profile, the
output contract, the precision and the dense-memory preflight. What changes
on this engine:

* the launch estimate is the native planner's own count (a virtual state
  plans, defers and logs the kernels without device memory), not a
  structural guess;
* double precision stays on the GPU (complex128 is a native mode of every
  kernel here), where the reference must fall back to a CPU engine;
* MPS requests go to Aer's MPS for Qiskit (as in the reference, no GPU
  crossover is established for MPS) and to this engine's device MPS
  (mettleq_amd.mps) for PennyLane, where the reference uses its CPU MPS;
* crossovers come from this engine's measured per-circuit latency (see
  mettleq_amd.planning): 16 qubits through Qiskit and 14 through PennyLane,
  where the SDK translation and fusion add less host time per gate.
"""
from __future__ import annotations

from dataclasses import asdict, dataclass
from typing import Optional, Sequence

from ..execution import statevector_preflight

_SINGLE = {"H", "X", "Y", "Z", "S", "SDG", "T", "TDG", "SX", "RX", "RY", "RZ", "U1", "U2", "U3"}
_AFFINE = {"X", "CNOT", "CX", "SWAP"}
_DIAGONAL = {"Z", "S", "SDG", "T", "TDG", "RZ", "U1", "CZ", "CPHASE"}
_PAIR_PHASE = {"XXPHASE", "YYPHASE", "ZZPHASE"}
_CONTROLLED = {"CH", "CRX", "CRY", "CRZ"}
_SUPPORTED = _SINGLE | _AFFINE | _DIAGONAL | _PAIR_PHASE | _CONTROLLED | {"ISWAP", "TOFFOLI", "FREDKIN"}

QISKIT_GPU_CROSSOVER_QUBITS = 16
PENNYLANE_GPU_CROSSOVER_QUBITS = 14


@dataclass(frozen=True)
class CircuitProfile:
    qubits: int
    operations: int
    single_qubit_operations: int
    two_qubit_operations: int
    nonlocal_two_qubit_operations: int
    maximum_two_qubit_span: int
    affine_operations: int
    diagonal_operations: int
    pair_phase_operations: int
    controlled_single_operations: int
    unsupported_operations: int
    estimated_native_operations: int
    estimated_fusion_coverage: float
    planned_sweeps: int
    planned_kernel_launches: int
    estimated_operations_per_launch: float
    topology_class: str

    def to_dict(self) -> dict:
        return asdict(self)


@dataclass(frozen=True)
class SDKEngineDecision:
    sdk: str
    engine: str
    reason: str
    precision: str
    output_contract: str
    calibrated_gpu_crossover_qubits: int
    adjusted_gpu_crossover_qubits: int
    statevector_preflight: dict
    profile: CircuitProfile
    fallback_available: bool

    def to_dict(self) -> dict:
        result = asdict(self)
        result["profile"] = self.profile.to_dict()
        return result


def profile_operations(n_qubits: int, operations: Sequence[dict]) -> CircuitProfile:
    """Gate mix, topology and the planner's launch count, with no statevector
    allocated (policy.py:75-148)."""
    from .. import _native as nat
    from ..dispatch import planned_run
    from ..fusion import fuse_operations
    n = int(n_qubits)
    ops = list(operations)
    names = [str(op.get("name", "")).upper() for op in ops]
    pairs = [tuple(int(w) for w in op.get("wires", [])) for op in ops if len(op.get("wires", [])) == 2]
    spans = [abs(a - b) for a, b in pairs]
    nonlocal_pairs = sum(s != 1 for s in spans)
    native = sum(name in _SUPPORTED for name in names)
    sweeps, launches = 0, 0
    if ops and native == len(ops) and n >= 1:
        run = planned_run(fuse_operations(ops, n), n)
        sweeps = int(run["sweeps"])
        launches = sum(1 for k in run["kernels"] if k in {nat.KERNEL_KINDS[i] for i in nat.SWEEP_KINDS})
    if not pairs:
        topology = "single_qubit_only"
    elif nonlocal_pairs == 0:
        topology = "nearest_neighbor"
    elif nonlocal_pairs <= max(1, len(pairs) // 5):
        topology = "mostly_local"
    else:
        topology = "long_range"
    return CircuitProfile(
        qubits=n, operations=len(ops),
        single_qubit_operations=sum(len(op.get("wires", [])) == 1 for op in ops),
        two_qubit_operations=len(pairs), nonlocal_two_qubit_operations=int(nonlocal_pairs),
        maximum_two_qubit_span=max(spans, default=0),
        affine_operations=sum(nm in _AFFINE for nm in names),
        diagonal_operations=sum(nm in _DIAGONAL for nm in names),
        pair_phase_operations=sum(nm in _PAIR_PHASE for nm in names),
        controlled_single_operations=sum(nm in _CONTROLLED for nm in names),
        unsupported_operations=sum(nm not in _SUPPORTED for nm in names),
        estimated_native_operations=native,
        estimated_fusion_coverage=float(native / len(ops)) if ops else 1.0,
        planned_sweeps=sweeps, planned_kernel_launches=launches,
        estimated_operations_per_launch=float(len(ops) / launches) if launches else 0.0,
        topology_class=topology,
    )


def recommend_sdk_engine(
    sdk: str,
    n_qubits: int,
    operations: Sequence[dict],
    *,
    precision: str = "single",
    output_contract: str = "samples",
    method: str = "statevector",
    allow_approximation: bool = False,
    fallback_available: bool = True,
    allow_unsafe_statevector: Optional[bool] = None,
    qiskit_gpu_crossover_qubits: int = QISKIT_GPU_CROSSOVER_QUBITS,
    pennylane_gpu_crossover_qubits: int = PENNYLANE_GPU_CROSSOVER_QUBITS,
) -> SDKEngineDecision:
    """Aer / Lightning CPU or this engine, with the evidence (policy.py:151-248)."""
    sdk = str(sdk).strip().lower()
    if sdk not in {"qiskit", "pennylane"}:
        raise ValueError("sdk must be 'qiskit' or 'pennylane'")
    precision = str(precision).strip().lower()
    if precision not in {"single", "double"}:
        raise ValueError("precision must be 'single' or 'double'")
    output_contract = str(output_contract).strip().lower()
    if output_contract not in {"statevector", "expectation", "probabilities", "samples"}:
        raise ValueError("unsupported output contract")
    method = str(method).strip().lower()
    profile = profile_operations(n_qubits, operations)
    dtype = "complex128" if precision == "double" else "complex64"
    preflight = statevector_preflight(int(n_qubits), dtype=dtype, allow_unsafe=allow_unsafe_statevector)
    base = int(qiskit_gpu_crossover_qubits) if sdk == "qiskit" else int(pennylane_gpu_crossover_qubits)
    adjusted = base
    reasons = []
    if profile.estimated_fusion_coverage < 0.5:
        adjusted += 2
        reasons.append("low_native_fusion_coverage")
    if output_contract in {"expectation", "probabilities"}:
        adjusted = max(1, adjusted - 1)
        reasons.append("gpu_native_reduction_contract")
    if profile.unsupported_operations:
        adjusted += 2
        reasons.append("unsupported_or_generic_operations_present")
    if profile.operations and profile.planned_kernel_launches and profile.estimated_operations_per_launch < 1.5:
        adjusted += 1
        reasons.append("high_planned_launch_density")
    cpu_engine = "qiskit_aer_cpu_statevector" if sdk == "qiskit" else "pennylane_lightning_cpu"
    gpu_engine = "mettleq_gpu_statevector" + ("_complex128" if precision == "double" else "")
    mps_engine = "qiskit_aer_cpu_mps" if sdk == "qiskit" else "mettleq_gpu_mps"
    if method in {"matrix_product_state", "mps"}:
        engine = mps_engine
        reason = ("MPS GPU crossover is not established; Aer's MPS is the Qiskit path" if sdk == "qiskit"
                  else "this engine's MPS (device SVDs) is the PennyLane path")
    elif not preflight["allowed"]:
        if allow_approximation and profile.topology_class in {
                "single_qubit_only", "nearest_neighbor", "mostly_local"}:
            engine = mps_engine
            reason = "dense preflight refused; bounded local MPS fallback was explicitly allowed"
        else:
            engine = "refused"
            reason = "dense preflight refused and no trustworthy approximate fallback was allowed"
    elif int(n_qubits) >= adjusted or not fallback_available:
        engine = gpu_engine
        reason = ("circuit reached the adjusted MI355X crossover" if int(n_qubits) >= adjusted
                  else "no CPU engine available")
    else:
        engine = cpu_engine
        reason = "circuit remains below the adjusted MI355X crossover"
    if reasons:
        reason += "; " + ", ".join(reasons)
    return SDKEngineDecision(
        sdk=sdk, engine=engine, reason=reason, precision=precision, output_contract=output_contract,
        calibrated_gpu_crossover_qubits=base, adjusted_gpu_crossover_qubits=adjusted,
        statevector_preflight=preflight, profile=profile, fallback_available=bool(fallback_available),
    )


@dataclass(frozen=True)
class EngineTarget:
    """Where an engine name of recommend_sdk_engine runs: the SDK's own CPU
    simulator (Aer / Lightning) or this package's engine, with the method
    and state type to run it with."""
    sdk_cpu: bool
    method: str
    dtype: str


def engine_target(engine: str) -> EngineTarget:
    if engine in ("qiskit_aer_cpu_statevector", "pennylane_lightning_cpu"):
        return EngineTarget(True, "statevector", "complex64")
    if engine == "qiskit_aer_cpu_mps":
        return EngineTarget(True, "matrix_product_state", "complex64")
    if engine == "mettleq_gpu_statevector":
        return EngineTarget(False, "statevector", "complex64")
    if engine == "mettleq_gpu_statevector_complex128":
        return EngineTarget(False, "statevector", "complex128")
    if engine == "mettleq_gpu_mps":
        return EngineTarget(False, "matrix_product_state", "complex64")
    raise ValueError(f"engine {engine!r} does not run anything")


def adaptive_plan(sdk: str, circuits: Sequence[tuple], **policy) -> tuple:
    """The per-batch part of the reference's adaptive SDK front ends
    (qiskit.py:244-316 AdaptiveQiskitBackend.run, pennylane.py:690-738
    AdaptivePennyLaneDevice.execute): one recommend_sdk_engine decision per
    circuit ((n_qubits, operations, output_contract) each, the remaining
    keywords passed through), and the circuit indices grouped by engine in
    order of first use. A refused circuit raises ValueError with its reason
    before anything runs. Returns (decisions, groups)."""
    decisions = [recommend_sdk_engine(sdk, n, ops, output_contract=contract, **policy)
                 for n, ops, contract in circuits]
    groups: dict = {}
    for index, decision in enumerate(decisions):
        if decision.engine == "refused":
            raise ValueError(decision.reason)
        groups.setdefault(decision.engine, []).append(index)
    return decisions, groups
