"""PennyLane device ``mettleq`` over the HIP statevector engine.

Mirrors the reference's MettleQDevice (integrations/pennylane.py:95-643):
wire k of the device is core wire k (MSB first, pennylane.py:397), Ising
angles are halved on entry (pennylane.py:394-396), GlobalPhase(phi) is
exp(-i phi) I, and State / Probability / Expectation / Variance / Sample /
Counts measurements are served from the device-resident statevector.
By default gradients stay with PennyLane's parameter shift
(use_device_gradient=False, pennylane.py:216-223). With
diff_method="adjoint" or "device", a tape that is a rotation ansatz (RX /
RY / RZ on trainable or fixed angles, fixed CZ / ControlledPhaseShift /
IsingZZ) measured by one expectation of a Z-string sum is differentiated by
the batched engine's adjoint pass (mettleq_amd.batched: energies and every
trainable parameter's derivative in one launch); other tapes are refused
for device derivatives, as PennyLane's own adjoint device refuses what it
cannot differentiate.
"""
from __future__ import annotations

from collections import Counter
from dataclasses import replace
from os import path
from typing import Optional

import numpy as np

try:
    import pennylane as qml
    from pennylane.devices import Device as PennyLaneDevice
    from pennylane.devices import ExecutionConfig
    from pennylane.devices.modifiers import simulator_tracking, single_tape_support
    from pennylane.devices.preprocess import (decompose, validate_device_wires,
                                              validate_measurements, validate_observables)
    try:
        from pennylane.exceptions import DeviceError
    except ImportError:  # older PennyLane
        from pennylane import DeviceError
    from pennylane.measurements import (CountsMP, ExpectationMP, ProbabilityMP, SampleMP, StateMP,
                                        VarianceMP)
    from pennylane.transforms import broadcast_expand, convert_to_numpy_parameters
    try:
        from pennylane.transforms.core import CompilePipeline as _Program
    except ImportError:  # older PennyLane
        from pennylane.transforms.core import TransformProgram as _Program
except ImportError as exc:  # pragma: no cover - exercised without the extra
    raise ImportError("The MettleQ PennyLane device requires the optional PennyLane dependency.") from exc

from ._common import (apply_global_phase, bit_counts, core_operation, execute_operations,
                      local_expectation, marginal_probabilities, pauli_product_expectation,
                      pauli_product_expectations, sample_bits, statevector_numpy)
from .policy import PENNYLANE_GPU_CROSSOVER_QUBITS, adaptive_plan, engine_target

_PENNYLANE_TO_CORE = {
    "Hadamard": "H", "PauliX": "X", "PauliY": "Y", "PauliZ": "Z", "S": "S", "T": "T", "SX": "SX",
    "RX": "RX", "RY": "RY", "RZ": "RZ", "PhaseShift": "U1", "CNOT": "CNOT", "CH": "CH", "CZ": "CZ",
    "ControlledPhaseShift": "CPHASE", "CRX": "CRX", "CRY": "CRY", "CRZ": "CRZ", "SWAP": "SWAP",
    "ISWAP": "ISWAP", "Toffoli": "TOFFOLI", "CSWAP": "FREDKIN", "IsingXX": "XXPHASE",
    "IsingYY": "YYPHASE", "IsingZZ": "ZZPHASE",
}
_ANALYTIC = (StateMP, ProbabilityMP, ExpectationMP, VarianceMP)
_SHOT = (ProbabilityMP, SampleMP, CountsMP, ExpectationMP, VarianceMP)


@simulator_tracking
@single_tape_support
class MettleQDevice(PennyLaneDevice):
    """PennyLane device running the exact statevector on one MI355X."""

    config_filepath = path.join(path.dirname(__file__), "pennylane_capabilities.toml")

    def __init__(self, wires, shots=None, *, seed=None, method: str = "automatic", device: str = "auto",
                 execution_report: bool = False, allow_unsafe_statevector: Optional[bool] = None,
                 dtype: str = "complex64", allow_approximation: bool = False,
                 mps_max_bond_dimension: int = 64, mps_truncation_threshold: float = 1e-10,
                 **_ignored) -> None:
        if wires is None:
            raise DeviceError("MettleQDevice requires an explicit wire set")
        super().__init__(wires=wires, shots=shots)
        if self.wires is None or len(self.wires) <= 0:
            raise DeviceError("MettleQDevice requires at least one wire")
        method = str(method).lower()
        method = {"mps": "matrix_product_state", "sv": "statevector"}.get(method, method)
        if method not in ("automatic", "statevector", "matrix_product_state"):
            raise DeviceError("method must be one of: automatic, statevector, matrix_product_state")
        self._method = method
        self._mps_options = {"allow_approximation": bool(allow_approximation),
                             "mps_max_bond_dimension": int(mps_max_bond_dimension),
                             "mps_truncation_threshold": float(mps_truncation_threshold)}
        if str(device).lower() not in ("auto", "gpu"):
            raise DeviceError("device must be 'auto' or 'gpu' for the HIP engine")
        self._rng = np.random.default_rng(seed)
        self._execution_report = bool(execution_report)
        self._allow_unsafe_statevector = allow_unsafe_statevector
        self._dtype = dtype
        self.last_execution_plan = None
        self.statevector_preflight = None
        self.last_execution_selection = None

    @property
    def name(self) -> str:
        return "mettleq"

    _DEVICE_GRADIENTS = ("adjoint", "device")

    def setup_execution_config(self, config: Optional[ExecutionConfig] = None, circuit=None):
        config = ExecutionConfig() if config is None else config
        device_grad = getattr(config, "gradient_method", None) in self._DEVICE_GRADIENTS
        return replace(config, convert_to_numpy=True, use_device_gradient=device_grad,
                       use_device_jacobian_product=False, grad_on_execution=False)

    def supports_derivatives(self, execution_config=None, circuit=None) -> bool:
        if execution_config is not None and getattr(execution_config, "gradient_method", None) \
                not in self._DEVICE_GRADIENTS:
            return False
        return circuit is None or self._batched_tape(circuit) is not None

    def _batched_tape(self, tape):
        """(ansatz structure, Z-string terms, constant, trainable values) of a
        tape the batched engine differentiates, or None."""
        from .. import batched
        if len(tape.measurements) != 1 or not isinstance(tape.measurements[0], ExpectationMP):
            return None
        sentence = self._sentence(tape.measurements[0].obs)
        if sentence is None:
            return None
        words, coeffs = [], []
        for word, c in sentence.items():
            words.append({self.wires.index(w): p for w, p in word.items()})
            coeffs.append(complex(c))
        tz = batched.z_terms(words, coeffs)
        if tz is None:
            return None
        trainable = list(tape.trainable_params)
        slot_of = {t: i for i, t in enumerate(trainable)}
        values = np.zeros(len(trainable))
        ops, counter = [], 0
        for op in tape.operations:
            params = []
            for v in op.parameters:
                val = self._parameter(v, op.name)
                if counter in slot_of:
                    values[slot_of[counter]] = val
                    params.append(batched.ParamRef(slot_of[counter]))
                else:
                    params.append(val)
                counter += 1
            name = {"RX": "RX", "RY": "RY", "RZ": "RZ", "CZ": "CZ", "ControlledPhaseShift": "CPHASE",
                    "IsingZZ": "ZZPHASE", "Identity": "I"}.get(op.name)
            if name is None:
                return None
            if name in ("CZ", "CPHASE", "ZZPHASE") and any(isinstance(p, batched.ParamRef) for p in params):
                return None
            if name == "ZZPHASE":
                params = [params[0] / 2.0]
            ops.append({"name": name, "wires": [self.wires.index(w) for w in op.wires],
                        "parameters": params})
        st = batched.ansatz_structure(len(self.wires), ops, n_params=len(trainable))
        if st is None:
            return None
        return st, tz[0], tz[1], values

    def compute_derivatives(self, circuits, execution_config=None):
        from .. import batched
        jacs = []
        for tape in circuits:
            match = self._batched_tape(tape)
            if match is None:
                raise DeviceError("device derivatives need a rotation ansatz with one Z-string expectation")
            st, terms, const, values = match
            _, g = batched.shared_estimator().evaluate(st, terms, const, values[None, :], grads=True)
            row = g[0]
            jacs.append(float(row[0]) if len(row) == 1 else tuple(float(x) for x in row))
        return tuple(jacs)

    def execute_and_compute_derivatives(self, circuits, execution_config=None):
        return self.execute(circuits, execution_config), self.compute_derivatives(circuits, execution_config)

    def _supports_operation(self, op) -> bool:
        return op.name in _PENNYLANE_TO_CORE or op.name in {"Identity", "GlobalPhase"}

    def preprocess_transforms(self, execution_config: Optional[ExecutionConfig] = None):
        program = _Program()
        program.add_transform(convert_to_numpy_parameters)
        program.add_transform(decompose, stopping_condition=self._supports_operation,
                              device_wires=self.wires, name=self.name)
        program.add_transform(broadcast_expand)
        program.add_transform(validate_device_wires, self.wires, name=self.name)
        program.add_transform(validate_measurements,
                              analytic_measurements=lambda m: isinstance(m, _ANALYTIC),
                              sample_measurements=lambda m: isinstance(m, _SHOT), name=self.name)
        program.add_transform(validate_observables,
                              stopping_condition=lambda o: bool(o.has_matrix), name=self.name)
        return program

    def execute(self, circuits, execution_config=None):
        cache: dict = {}
        try:
            return tuple(self._execute_circuit(c, cache) for c in circuits)
        finally:
            for dev in cache.values():
                dev.close()

    def _execute_circuit(self, circuit, cache, *, method: Optional[str] = None, dtype: Optional[str] = None):
        operations, phase = self._canonical_operations(circuit)
        shots = circuit.shots if circuit.shots else self.shots
        default_shots = shots.total_shots or 1000
        device = execute_operations(len(self.wires), operations, shots=default_shots,
                                    report=self._execution_report,
                                    allow_unsafe_statevector=self._allow_unsafe_statevector,
                                    execution_cache=cache, dtype=dtype or self._dtype,
                                    method=method or self._method, **self._mps_options)
        apply_global_phase(device, phase)
        self.last_execution_plan = device.last_execution_plan
        self.statevector_preflight = device.statevector_preflight
        self.last_execution_selection = device.execution_selection
        if shots.has_partitioned_shots:
            return tuple(self._measure_all(device, circuit.measurements, int(s)) for s in shots)
        return self._measure_all(device, circuit.measurements, shots.total_shots)

    def _canonical_operations(self, circuit):
        operations, phase = [], 0.0
        for op in circuit.operations:
            if op.name == "Identity":
                continue
            if op.name == "GlobalPhase":
                phase -= self._parameter(op.parameters[0], op.name)
                continue
            if op.name not in _PENNYLANE_TO_CORE:
                raise DeviceError(f"Operation {op.name!r} reached MettleQ without a supported decomposition")
            params = [self._parameter(v, op.name) for v in op.parameters]
            if op.name in ("IsingXX", "IsingYY", "IsingZZ"):
                params[0] /= 2.0
            wires = [self.wires.index(w) for w in op.wires]
            try:
                operations.append(core_operation(_PENNYLANE_TO_CORE[op.name], wires, params))
            except ValueError as exc:
                raise DeviceError(str(exc)) from exc
        return operations, phase

    @staticmethod
    def _parameter(value, name: str) -> float:
        arr = np.asarray(value)
        if arr.ndim != 0:
            raise DeviceError(f"Operation {name!r} still has broadcast parameters")
        try:
            return float(arr)
        except (TypeError, ValueError) as exc:
            raise DeviceError(f"Operation {name!r} has a non-numeric parameter") from exc

    def _measure_all(self, device, measurements, shots):
        device.synchronize()
        vals = tuple(self._measure(device, m, shots) for m in measurements)
        return vals[0] if len(vals) == 1 else vals

    def _measure(self, device, m, shots):
        if isinstance(m, StateMP):
            if shots is not None:
                raise DeviceError("State measurements require analytic execution")
            return statevector_numpy(device)
        if m.obs is None:
            wires = ([self.wires.index(w) for w in m.wires] if len(m.wires)
                     else list(range(len(self.wires))))
            if isinstance(m, ProbabilityMP):
                if shots is None:
                    return marginal_probabilities(device, wires)
                samples = sample_bits(device, shots, wires, self._rng)
                idx = samples.dot(1 << np.arange(len(wires) - 1, -1, -1))
                return np.bincount(idx, minlength=1 << len(wires)) / float(shots)
            if isinstance(m, SampleMP):
                if shots is None:
                    raise DeviceError("Sample measurements require finite shots")
                return sample_bits(device, shots, wires, self._rng)
            if isinstance(m, CountsMP):
                if shots is None:
                    raise DeviceError("Counts measurements require finite shots")
                counts = bit_counts(sample_bits(device, shots, wires, self._rng))
                if m.all_outcomes:
                    for v in range(1 << len(wires)):
                        counts.setdefault(format(v, f"0{len(wires)}b"), 0)
                return counts
        if isinstance(m, (ExpectationMP, VarianceMP)):
            if shots is None:
                e, e2 = self._moments(device, m.obs)
                if isinstance(m, ExpectationMP):
                    return self._real(e, "expectation")
                return self._real(e2 - e * e, "variance")
            samples = self._sample_observable(device, m.obs, shots)
            if isinstance(m, ExpectationMP):
                return self._real(np.mean(samples), "expectation")
            return self._real(np.var(samples), "variance")
        if isinstance(m, (SampleMP, CountsMP)):
            if shots is None:
                raise DeviceError("Observable sampling requires finite shots")
            samples = self._sample_observable(device, m.obs, shots)
            if isinstance(m, SampleMP):
                return samples
            counts = dict(Counter(np.real_if_close(v).item() for v in samples))
            if m.all_outcomes:
                for v in np.unique(np.real_if_close(m.eigvals())):
                    counts.setdefault(v.item(), 0)
            return counts
        raise DeviceError(f"Measurement {type(m).__name__!r} is not supported")

    def _moments(self, device, obs):
        sentence = self._sentence(obs)
        if sentence is not None:
            return self._sentence_value(device, sentence), self._sentence_value(device, sentence @ sentence)
        wires = [self.wires.index(w) for w in obs.wires]
        if len(wires) > 8:
            raise DeviceError("Non-Pauli observables spanning more than 8 wires are not supported")
        mat = np.asarray(qml.matrix(obs, wire_order=obs.wires))
        return local_expectation(device, wires, mat), local_expectation(device, wires, mat @ mat)

    def _sample_observable(self, device, obs, shots):
        sentence = self._sentence(obs)
        if sentence is not None and len(sentence) == 1:
            word, coeff = next(iter(sentence.items()))
            if not word:
                return np.full(int(shots), np.real_if_close(coeff))
            e = pauli_product_expectation(device, {self.wires.index(w): p for w, p in word.items()})
            e = float(np.clip(e.real, -1.0, 1.0))
            signs = self._rng.choice(np.array([-1.0, 1.0]), size=int(shots),
                                     p=((1.0 - e) / 2.0, (1.0 + e) / 2.0))
            return np.real_if_close(complex(coeff) * signs)
        wires = [self.wires.index(w) for w in obs.wires]
        mat = np.asarray(qml.matrix(obs, wire_order=obs.wires))
        evals, evecs = np.linalg.eigh(mat)
        probs = np.array([local_expectation(device, wires, np.outer(v, v.conj())).real for v in evecs.T])
        probs = np.clip(probs, 0.0, None)
        probs /= probs.sum()
        return np.real_if_close(evals[self._rng.choice(len(evals), size=int(shots), p=probs)])

    def _sentence_value(self, device, sentence) -> complex:
        coeffs, words = [], []
        for word, c in sentence.items():
            coeffs.append(complex(c))
            words.append({self.wires.index(w): p for w, p in word.items()})
        vals = pauli_product_expectations(device, words)
        return sum(c * v for c, v in zip(coeffs, vals))

    @staticmethod
    def _sentence(obs):
        try:
            return qml.pauli.pauli_sentence(obs)
        except (TypeError, ValueError):
            return None

    @staticmethod
    def _real(value, label: str) -> float:
        value = complex(value)
        tol = 5e-5 * max(1.0, abs(value.real))
        if abs(value.imag) > tol:
            raise DeviceError(f"Hermitian {label} developed an imaginary component {value.imag:.3e}")
        return float(value.real)


@simulator_tracking
@single_tape_support
class AdaptivePennyLaneDevice(MettleQDevice):
    """Each tape to Lightning's CPU simulator or to the MI355X engine
    (pennylane.py:646-741): recommend_sdk_engine on the tape's gate stream,
    the precision, the output contract its measurements imply and the dense
    preflight, with the crossover calibrated on MI355X. Without
    pennylane_lightning every tape stays on this engine; the decisions are
    kept in last_adaptive_decisions."""

    def __init__(self, wires, shots=None, *, precision: str = "single",
                 pennylane_gpu_crossover_qubits: int = PENNYLANE_GPU_CROSSOVER_QUBITS, **kwargs) -> None:
        super().__init__(wires=wires, shots=shots, **kwargs)
        precision = str(precision).strip().lower()
        if precision not in ("single", "double"):
            raise DeviceError("precision must be 'single' or 'double'")
        self._adaptive_precision = precision
        self._adaptive_gpu_crossover = int(pennylane_gpu_crossover_qubits)
        self.last_adaptive_decisions = []

    @property
    def name(self) -> str:
        return "mettleq.adaptive"

    @staticmethod
    def _lightning_available() -> bool:
        try:
            import pennylane_lightning  # noqa: F401
        except ImportError:
            return False
        return True

    @staticmethod
    def _output_contract(circuit) -> str:
        kinds = [type(m) for m in circuit.measurements]
        for mp, contract in ((StateMP, "statevector"), (ExpectationMP, "expectation"),
                             (ProbabilityMP, "probabilities")):
            if any(issubclass(k, mp) for k in kinds):
                return contract
        return "samples"

    def execute(self, circuits, execution_config=None):
        config = self.setup_execution_config(execution_config)
        circuits = list(circuits)
        batch = [(len(self.wires), self._canonical_operations(c)[0], self._output_contract(c)) for c in circuits]
        try:
            decisions, _ = adaptive_plan(
                "pennylane", batch, precision=self._adaptive_precision, method=self._method,
                allow_approximation=self._mps_options["allow_approximation"],
                fallback_available=self._lightning_available(),
                allow_unsafe_statevector=self._allow_unsafe_statevector,
                pennylane_gpu_crossover_qubits=self._adaptive_gpu_crossover)
        except ValueError as exc:
            raise DeviceError(str(exc)) from exc
        self.last_adaptive_decisions = [d.to_dict() for d in decisions]
        lightning: dict = {}
        cache: dict = {}
        results = []
        try:
            for circuit, decision in zip(circuits, decisions):
                target = engine_target(decision.engine)
                if target.sdk_cpu:
                    key = repr(circuit.shots)
                    if key not in lightning:
                        lightning[key] = qml.device("lightning.qubit", wires=self.wires, shots=circuit.shots)
                    results.append(lightning[key].execute(circuit, config))
                    self.last_execution_plan = None
                    self.statevector_preflight = decision.statevector_preflight
                    self.last_execution_selection = decision.to_dict()
                else:
                    results.append(self._execute_circuit(circuit, cache, method=target.method, dtype=target.dtype))
        finally:
            for dev in cache.values():
                dev.close()
        return tuple(results)

