"""Shared execution and measurement helpers of the SDK adapters.

Restates the reference's adapter seam (integrations/_common.py:29-427):
canonical op validation, ``execute_operations`` with a per-width Device
cache, global phase, statevector / marginal / sample read-back and Pauli
expectations, all over the HIP engine: the exact statevector, or the MPS
method (mettleq_amd.mps) when the caller asks for it or allows approximation
(mettleq_amd.planning.select_execution).
"""
from __future__ import annotations

from collections import Counter
from typing import Iterable, Optional, Sequence

import numpy as np

from ..device import Device

_ARITIES = {
    "H": 1, "I": 1, "X": 1, "Y": 1, "Z": 1, "S": 1, "SDG": 1, "T": 1, "TDG": 1, "SX": 1,
    "RX": 1, "RY": 1, "RZ": 1, "U1": 1, "U2": 1, "U3": 1,
    "CNOT": 2, "CH": 2, "CZ": 2, "CPHASE": 2, "CRX": 2, "CRY": 2, "CRZ": 2, "SWAP": 2,
    "ISWAP": 2, "XXPHASE": 2, "YYPHASE": 2, "ZZPHASE": 2,
    "TOFFOLI": 3, "FREDKIN": 3,
}
_PARAMETER_COUNTS = {
    "RX": 1, "RY": 1, "RZ": 1, "U1": 1, "U2": 2, "U3": 3, "CPHASE": 1, "CRX": 1, "CRY": 1,
    "CRZ": 1, "XXPHASE": 1, "YYPHASE": 1, "ZZPHASE": 1,
}
_ALIASES = {"CX": "CNOT", "CP": "CPHASE", "CCX": "TOFFOLI", "CSWAP": "FREDKIN"}


def core_operation(name: str, wires: Iterable[int], parameters: Sequence[float] = ()) -> dict:
    """Build and validate one canonical op (_common.py:86-115)."""
    canonical = _ALIASES.get(str(name).upper(), str(name).upper())
    w = [int(x) for x in wires]
    p = [float(x) for x in parameters]
    if canonical == "I":
        return {"name": "I", "wires": w, "parameters": []}
    if canonical not in _ARITIES:
        raise ValueError(f"Unsupported operation: {name}")
    if len(w) != _ARITIES[canonical]:
        raise ValueError(f"{canonical} expects {_ARITIES[canonical]} wire(s), got {len(w)}")
    expected = _PARAMETER_COUNTS.get(canonical, 0)
    if len(p) != expected:
        raise ValueError(f"{canonical} expects {expected} parameter(s), got {len(p)}")
    return {"name": canonical, "wires": w, "parameters": p}


def execute_operations(n_qubits: int, operations: Sequence[dict], *, shots: int = 1000,
                       report: bool = False, allow_unsafe_statevector: Optional[bool] = None,
                       method: str = "automatic", execution_device: str = "auto",
                       execution_cache: Optional[dict] = None, dtype: str = "complex64",
                       allow_approximation: bool = False, mps_max_bond_dimension: int = 64,
                       mps_truncation_threshold: float = 1e-10,
                       automatic_mps_min_qubits: Optional[int] = None,
                       **_ignored) -> Device:
    """Execute canonical ops on a (cached per width) Device (_common.py:118-228)."""
    from ..planning import DEFAULT_AUTOMATIC_MPS_MIN_QUBITS, normalize_method, select_execution
    method = normalize_method(method)
    device_key = str(execution_device).strip().lower()
    if device_key not in ("auto", "gpu"):
        raise ValueError("execution_device must be 'auto' or 'gpu' for the HIP engine")
    ops = list(operations)
    if method == "matrix_product_state" or allow_approximation:
        sel = select_execution(int(n_qubits), ops, method=method, device="auto", dtype=dtype,
                               allow_approximation=bool(allow_approximation),
                               allow_unsafe_statevector=allow_unsafe_statevector,
                               automatic_mps_min_qubits=(DEFAULT_AUTOMATIC_MPS_MIN_QUBITS
                                                         if automatic_mps_min_qubits is None
                                                         else int(automatic_mps_min_qubits)))
        if sel.selected_method == "matrix_product_state":
            from ..mps import MPSOptions
            opts = MPSOptions(dmax=int(mps_max_bond_dimension), eps=float(mps_truncation_threshold))
            device = Device(n_qubits, shots=shots, backend="mps", mps_opts=opts)
            device.execute(ops)
            device.execution_selection = {**sel.to_dict(), "engine": "hip_mps"}
            return device
    needs_second_buffer = any(str(op.get("name", "")).upper() in ("CNOT", "CX", "X", "SWAP")
                              for op in ops)
    mode = "out_of_place" if needs_second_buffer else "in_place"
    cache_key = (int(n_qubits), "sv", "gpu", dtype)
    device = execution_cache.get(cache_key) if execution_cache is not None else None
    if device is None:
        device = Device(n_qubits, shots=shots, allow_unsafe_statevector=allow_unsafe_statevector,
                        dtype=dtype, preflight_mode=mode)
        if execution_cache is not None:
            execution_cache[cache_key] = device
    else:
        device.shots = int(shots)
        device.reset()
    device.execute(ops, report=report)
    device.execution_selection = {"selected_method": "statevector", "selected_device": "gpu",
                                  "engine": "hip", "preflight_mode": mode}
    if device.last_execution_plan is not None:
        device.last_execution_plan["sdk_execution_selection"] = device.execution_selection
    return device


def apply_global_phase(device: Device, phase: float) -> None:
    """_common.py:231-238, as one device scale pass."""
    if phase:
        f = complex(np.cos(float(phase)), np.sin(float(phase)))
        if device.backend == "mps":
            device.sim.apply_single(np.eye(2) * f, 0)
        else:
            device.native.scale(f)


def statevector_numpy(device: Device) -> np.ndarray:
    device.synchronize()
    if device.backend == "mps":
        return device.sim.to_statevector()
    return device.native.to_numpy().astype(np.complex64, copy=False)


def marginal_probabilities(device: Device, wires: Sequence[int]) -> np.ndarray:
    """FP64 marginal, renormalized on the host (_common.py:256-266)."""
    device.synchronize()
    result = np.asarray(device.sim.probabilities_array(list(wires)), dtype=np.float64)
    total = float(result.sum())
    if total <= 0.0:
        raise RuntimeError("Simulator returned a zero-probability state")
    return result / total


def sample_bits(device: Device, shots: int, wires: Sequence[int], rng: np.random.Generator) -> np.ndarray:
    """_common.py:269-288."""
    wires = list(wires)
    if not wires:
        return np.empty((int(shots), 0), dtype=np.int64)
    device.synchronize()
    return np.asarray(device.sim.sample_array(int(shots), wires, rng=rng), dtype=np.int64)


def bit_counts(samples: np.ndarray) -> dict:
    return dict(Counter("".join(str(int(b)) for b in row) for row in samples))


def pauli_product_expectation(device: Device, paulis: dict) -> complex:
    return pauli_product_expectations(device, [paulis])[0]


def pauli_product_expectations(device: Device, words: Sequence[dict]) -> list:
    """All words in one device pass per 16 words (_common.py:384-427)."""
    normalized = [{int(w): str(p).upper() for w, p in word.items() if str(p).upper() != "I"}
                  for word in words]
    for word in normalized:
        for p in word.values():
            if p not in ("X", "Y", "Z"):
                raise ValueError("Pauli products support only I, X, Y, and Z")
    device.synchronize()
    if device.backend == "mps":
        mats = {"X": np.array([[0, 1], [1, 0]], dtype=complex), "Y": np.array([[0, -1j], [1j, 0]]),
                "Z": np.diag([1.0, -1.0]).astype(complex)}
        return [device.sim.expectation_product({w: mats[p] for w, p in word.items()}) for word in normalized]
    vals = device.native.expval_pauli(normalized)
    return [complex(v) for v in vals]


def local_expectation(device: Device, wires: Sequence[int], matrix: np.ndarray) -> complex:
    """<psi| O_wires |psi> for a small Hermitian O, via its Pauli expansion
    (4^k words evaluated in one device pass per 16 words)."""
    wires = [int(w) for w in wires]
    k = len(wires)
    m = np.asarray(matrix, dtype=np.complex128)
    if m.shape != (1 << k, 1 << k):
        raise ValueError(f"Observable matrix must have shape {(1 << k, 1 << k)}, got {m.shape}")
    if len(set(wires)) != k:
        raise ValueError("Observable wires must be unique")
    if any(w < 0 or w >= device.wires for w in wires):
        raise ValueError("Observable wire is outside the device")
    paulis = {"I": np.eye(2), "X": np.array([[0, 1], [1, 0]]), "Y": np.array([[0, -1j], [1j, 0]]),
              "Z": np.array([[1, 0], [0, -1]])}
    labels = ["I", "X", "Y", "Z"]
    words, coeffs = [], []
    for code in range(4 ** k):
        lab = [labels[(code >> (2 * (k - 1 - j))) & 3] for j in range(k)]
        p = np.array([[1.0]])
        for l in lab:
            p = np.kron(p, paulis[l])
        c = np.trace(p.conj().T @ m) / (1 << k)
        if abs(c) > 1e-15:
            words.append({w: l for w, l in zip(wires, lab) if l != "I"})
            coeffs.append(c)
    if not words:
        return 0j
    vals = pauli_product_expectations(device, words)
    return complex(sum(c * v for c, v in zip(coeffs, vals)))
