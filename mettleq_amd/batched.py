"""Batched estimator path of the SDK adapters.

The reference's EstimatorV2 runs one bound circuit per parameter binding
(integrations/qiskit.py:906-1090), and its PennyLane device differentiates
by parameter shift, two circuits per parameter (integrations/pennylane.py:
216-223, 558-625). When a parameterized circuit is a rotation ansatz and the
observable a sum of Z-strings, this module hands all bindings to the batched
engine instead (msv_ansatz_*: every binding in one graph launch, gradients
by one adjoint pass):

* rotations RX / RY / RZ whose angle is a constant or an affine function
  a * p + b of one circuit parameter p (ParamRef), each given its own engine
  parameter slot whose per-binding value the host computes; the gradient of
  p sums a times the slots' gradients;
* fixed two-qubit phases CZ, CPHASE(theta), ZZPHASE(theta);
* an observable sum_t w_t Z^{S_t} with real weights (the identity term adds
  a constant).

Anything else returns None from `ansatz_structure` / `z_terms`, and the
adapter keeps its per-binding path.
"""
from __future__ import annotations

import hashlib
import math
import threading
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np


@dataclass(frozen=True)
class ParamRef:
    """Angle = coeff * (circuit parameter `index`) + offset."""
    index: int
    coeff: float = 1.0
    offset: float = 0.0


@dataclass
class AnsatzStructure:
    n: int
    gates: List[tuple]           # ("RX"|"RY"|"RZ", qubit, slot) or ("CPHASE"|"ZZPHASE", a, b, theta)
    slot_index: np.ndarray       # circuit parameter of each slot (-1: constant angle)
    slot_coeff: np.ndarray
    slot_offset: np.ndarray
    n_params: int                # circuit parameters

    @property
    def key(self) -> bytes:
        h = hashlib.sha1(repr((self.n, [g if g[0] in ("RX", "RY", "RZ") else (g[0], g[1], g[2], float(g[3]))
                                         for g in self.gates])).encode())
        return h.digest()

    def slot_values(self, values: np.ndarray) -> np.ndarray:
        """[B, n_params] circuit parameter values -> [B, slots] engine parameters."""
        v = np.asarray(values, dtype=np.float64)
        if v.ndim == 1:
            v = v[None, :]
        if v.shape[1] != self.n_params:
            raise ValueError(f"expected {self.n_params} parameter values per binding, got {v.shape[1]}")
        idx = np.where(self.slot_index >= 0, self.slot_index, 0)
        base = v[:, idx] if self.n_params else np.zeros((v.shape[0], len(idx)))
        return np.where(self.slot_index >= 0, base * self.slot_coeff, 0.0) + self.slot_offset

    def param_grads(self, slot_grads: np.ndarray) -> np.ndarray:
        """[B, slots] engine gradients -> [B, n_params] circuit parameter gradients."""
        g = np.zeros((slot_grads.shape[0], self.n_params))
        for s, (p, a) in enumerate(zip(self.slot_index, self.slot_coeff)):
            if p >= 0:
                g[:, p] += a * slot_grads[:, s]
        return g


_ROT = ("RX", "RY", "RZ")


def ansatz_structure(n: int, ops: Sequence[dict], n_params: Optional[int] = None) -> Optional[AnsatzStructure]:
    """Canonical ops (wire 0 = most significant qubit; rotation angles floats
    or ParamRef) -> the batched engine's gate list, or None when an op is
    outside the rotation + phase form."""
    gates: List[tuple] = []
    idx, coeff, off = [], [], []
    max_p = -1
    for op in ops:
        name = str(op.get("name", "")).upper()
        w = [int(x) for x in op.get("wires", [])]
        params = list(op.get("parameters", []))
        if name == "I":
            continue
        if name in _ROT:
            if len(w) != 1 or len(params) != 1:
                return None
            p = params[0]
            slot = len(idx)
            if isinstance(p, ParamRef):
                idx.append(int(p.index))
                coeff.append(float(p.coeff))
                off.append(float(p.offset))
                max_p = max(max_p, int(p.index))
            else:
                idx.append(-1)
                coeff.append(0.0)
                off.append(float(p))
            gates.append((name, w[0], slot))
            continue
        if any(isinstance(p, ParamRef) for p in params):
            return None
        if name == "CZ" and len(w) == 2:
            gates.append(("CPHASE", w[0], w[1], math.pi))
        elif name in ("CPHASE", "CP") and len(w) == 2:
            gates.append(("CPHASE", w[0], w[1], float(params[0])))
        elif name == "ZZPHASE" and len(w) == 2:
            gates.append(("ZZPHASE", w[0], w[1], float(params[0])))
        else:
            return None
    if not idx:
        return None
    return AnsatzStructure(n=int(n), gates=gates, slot_index=np.asarray(idx, dtype=np.int64),
                           slot_coeff=np.asarray(coeff), slot_offset=np.asarray(off),
                           n_params=int(n_params) if n_params is not None else max_p + 1)


def z_terms(words: Sequence[dict], coeffs: Sequence[complex]) -> Optional[Tuple[List[tuple], float]]:
    """Pauli words {wire: 'X'|'Y'|'Z'|'I'} with coefficients -> (Z-string
    terms (wires, weight), constant), or None when a word has X or Y or a
    coefficient is not real."""
    terms, const = [], 0.0
    for word, c in zip(words, coeffs):
        c = complex(c)
        if abs(c.imag) > 1e-12 * max(1.0, abs(c.real)):
            return None
        qs = []
        for wire, ch in word.items():
            ch = str(ch).upper()
            if ch == "I":
                continue
            if ch != "Z":
                return None
            qs.append(int(wire))
        if qs:
            terms.append((tuple(sorted(qs)), c.real))
        else:
            const += c.real
    return terms, const


class BatchedEstimator:
    """LayeredAnsatz objects cached per circuit structure (each holds its
    buffers, autotuned plans and captured graphs per batch size)."""

    def __init__(self, device: int = 0, max_cached: int = 8):
        self.device = int(device)
        self.max_cached = int(max_cached)
        self._cache: Dict[bytes, object] = {}
        self._lock = threading.Lock()

    def _ansatz(self, st: AnsatzStructure):
        from .ansatz import LayeredAnsatz
        a = self._cache.get(st.key)
        if a is None:
            if len(self._cache) >= self.max_cached:
                old = self._cache.pop(next(iter(self._cache)))
                old.close()
            a = LayeredAnsatz(st.n, self.device)
            for g in st.gates:
                if g[0] in _ROT:
                    a.rotation(g[0], g[1], g[2])
                else:
                    a.phase(g[0], g[1], g[2], g[3])
            self._cache[st.key] = a
        return a

    def evaluate(self, st: AnsatzStructure, terms: List[tuple], const: float, values, grads: bool = False):
        """Energies [B] (and [B, n_params] gradients) of every binding."""
        slots = st.slot_values(values)
        with self._lock:
            a = self._ansatz(st)
            if a.n_params != slots.shape[1]:
                raise ValueError("internal: ansatz slot count mismatch")
            a.observable(terms if terms else [((), 0.0)])
            out = a.evaluate(slots, grads=grads)
        if grads:
            e, g = out
            return e + const, st.param_grads(g)
        return out + const

    def close(self) -> None:
        with self._lock:
            for a in self._cache.values():
                a.close()
            self._cache.clear()


_SHARED: Optional[BatchedEstimator] = None


def shared_estimator() -> BatchedEstimator:
    global _SHARED
    if _SHARED is None:
        _SHARED = BatchedEstimator()
    return _SHARED
