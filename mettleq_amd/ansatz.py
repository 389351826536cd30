"""Batched parameter bindings of a layered rotation ansatz with adjoint
gradients, on the HIP engine (msv_ansatz_* in include/mettleq_sv.h).

This is the batched form of the reference's per-binding estimator loop
(integrations/qiskit.py:906-1090, one circuit execution per parameter
binding) and of PennyLane parameter-shift differentiation through the
`mettleq` device (integrations/pennylane.py:558-625): all bindings run as
one graph launch, and gradients come from one adjoint pass instead of two
shifted circuits per parameter.

Conventions follow the reference's gates (gates.py:5-234): RX/RY/RZ(theta)
= exp(-i theta P / 2); ZZPHASE(t) = exp(-i t ZZ) (the SDK's rzz(theta) /
IsingZZ(theta) is ZZPHASE(theta / 2)); CPHASE(phi) = diag(1, 1, 1, e^{i phi}),
CZ = CPHASE(pi). Observables are sums of weighted Z-strings.
"""
from __future__ import annotations

import ctypes
from typing import Iterable, List, Optional, Sequence, Tuple

import numpy as np

from . import _native as nat

ROT = {"RX": nat.MSV_ROT_X, "RY": nat.MSV_ROT_Y, "RZ": nat.MSV_ROT_Z}


class LayeredAnsatz:
    """Gate list of parameterized rotations and fixed two-qubit phases over
    `n_qubits`, evaluated for many parameter bindings at once."""

    def __init__(self, n_qubits: int, device: int = 0):
        lib = nat.load_library()
        h = ctypes.c_void_p()
        nat.check(lib.msv_ansatz_create(int(n_qubits), int(device), ctypes.byref(h)), "ansatz create")
        self._h = h
        self.n_qubits = int(n_qubits)
        self.n_params = 0
        self.gates: List[Tuple] = []
        self.terms: List[Tuple[int, float]] = []

    # -- construction -------------------------------------------------------
    def rotation(self, name: str, qubit: int, param: int) -> "LayeredAnsatz":
        kind = ROT[name.upper()]
        nat.check(nat.lib().msv_ansatz_rotation(self._h, kind, int(qubit), int(param)), name)
        self.gates.append((name.upper(), int(qubit), int(param)))
        self.n_params = max(self.n_params, int(param) + 1)
        return self

    def cz(self, a: int, b: int) -> "LayeredAnsatz":
        return self.phase("CPHASE", a, b, np.pi)

    def phase(self, name: str, a: int, b: int, theta: float) -> "LayeredAnsatz":
        kind = {"ZZPHASE": nat.MSV_PHASE_ZZ, "CPHASE": nat.MSV_PHASE_CPHASE}[name.upper()]
        nat.check(nat.lib().msv_ansatz_phase(self._h, kind, int(a), int(b), float(theta)), name)
        self.gates.append((name.upper(), int(a), int(b), float(theta)))
        return self

    def observable(self, terms: Iterable[Tuple[Sequence[int], float]]) -> "LayeredAnsatz":
        """terms: (qubits of the Z-string, weight) pairs."""
        masks, weights = [], []
        for qubits, w in terms:
            m = 0
            for q in qubits:
                m |= 1 << int(q)
            masks.append(m)
            weights.append(float(w))
        zm = np.asarray(masks, dtype=np.uint64)
        wt = np.asarray(weights, dtype=np.float64)
        nat.check(nat.lib().msv_ansatz_observable(self._h, nat.as_ptr(zm, ctypes.c_uint64),
                                                  nat.as_ptr(wt, ctypes.c_double), len(masks)),
                  "observable")
        self.terms = list(zip(masks, weights))
        return self

    # -- evaluation ---------------------------------------------------------
    def evaluate(self, params, grads: bool = True):
        """params [batch, n_params] -> energies [batch] (and grads [batch, n_params])."""
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64))
        if p.ndim == 1:
            p = p[None, :]
        if p.ndim != 2 or p.shape[1] != self.n_params:
            raise ValueError(f"params must have shape [batch, {self.n_params}]")
        b = p.shape[0]
        e = np.empty(b, dtype=np.float64)
        g = np.empty((b, self.n_params), dtype=np.float64) if grads else None
        nat.check(nat.lib().msv_ansatz_evaluate(
            self._h, nat.as_ptr(p, ctypes.c_double), b, self.n_params,
            nat.as_ptr(e, ctypes.c_double), nat.as_ptr(g, ctypes.c_double) if grads else None),
            "ansatz evaluate")
        return (e, g) if grads else e

    def info(self, batch: int, grads: bool = True) -> dict:
        out = (ctypes.c_int64 * 7)()
        nat.check(nat.lib().msv_ansatz_info(self._h, int(batch), int(bool(grads)), out, 7), "ansatz info")
        return {"tile_local_bits": out[0], "blocks_per_cu": out[1], "forward_sweeps": out[2],
                "backward_sweeps": out[3], "tile_ops": out[4], "autotuned_ms": out[5] / 1000.0,
                "compiled_program_sweeps": out[6]}

    def sweep_times(self, batch: int, grads: bool = True, reps: int = 3) -> dict:
        """Forward and backward sweeps of the cached plan for `batch`, timed
        alone with HIP events (ms per evaluation), with their algorithmic
        HBM bytes: a forward sweep reads and writes every binding's state
        once (2 x batch x 2^n x 8 B), a backward sweep both phi and lambda.
        Leaves the states without meaningful values: evaluate again before
        reading results."""
        out = (ctypes.c_double * 2)()
        nat.check(nat.lib().msv_ansatz_time_sweeps(self._h, int(batch), int(bool(grads)), int(reps), out),
                  "ansatz sweep timing")
        info = self.info(batch, grads)
        state = float(batch) * 8.0 * 2.0 ** max(self.n_qubits, 10)
        return {"forward_ms": out[0], "backward_ms": out[1],
                "forward_bytes": 2.0 * state * info["forward_sweeps"],
                "backward_bytes": 4.0 * state * info["backward_sweeps"]}

    def close(self) -> None:
        if getattr(self, "_h", None):
            nat.lib().msv_ansatz_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def hardware_efficient(n_qubits: int, n_layers: int, device: int = 0) -> LayeredAnsatz:
    """Per layer: RY(theta[l, q, 0]) then RZ(theta[l, q, 1]) on every qubit,
    then CZ on the ring (q, q + 1 mod n); parameter index 2 (l n + q) + k."""
    a = LayeredAnsatz(n_qubits, device)
    for l in range(n_layers):
        for q in range(n_qubits):
            a.rotation("RY", q, 2 * (l * n_qubits + q))
            a.rotation("RZ", q, 2 * (l * n_qubits + q) + 1)
        for q in range(n_qubits):
            if n_qubits > 1 and not (n_qubits == 2 and q == 1):
                a.cz(q, (q + 1) % n_qubits)
    return a


def zz_ring_terms(n_qubits: int):
    """H = sum_i Z_i Z_{i+1 mod n}."""
    return [((i, (i + 1) % n_qubits), 1.0) for i in range(n_qubits)]
