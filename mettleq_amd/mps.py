"""Matrix product states on the MI355X engine.

The reference's MPS method (src/mettleq/mps_state.py:57-1549, native core
_mps_native.cpp) with the same option names, gate conventions and
diagnostics. Site tensors (dl, 2, dr) live in device memory; each two-site
update merges the pair with the gate on the device, takes the SVD on the
device (a one-sided Jacobi kernel) and applies the reference's truncation rule
(mps_state.py:193-247) before splitting, with the orthogonality center moved
to the pair first (mps_state.py:347-399). Non-adjacent two-qubit gates go
through the reference's restoring swap network (mps_state.py:744-768).

Readout: expectation values of tensor products of 2x2 operators run on the
MPS itself at any width (transfer matrices on the device); the dense state,
probabilities and samples contract the MPS into a statevector of the engine
first (widths up to what a complex64 state fits).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _native as nat

_NUMERICAL = ("singular spectrum", "did not converge", "non-finite", "zero norm", "zero statevector")


@dataclass
class MPSOptions:
    """mps_state.py:57-64. svd_driver and the routing options of the
    reference select CPU LAPACK drivers and its lookahead planner; this
    engine has one device SVD and the restoring swap network."""
    dmax: int = 64
    eps: float = 1e-10
    renormalize_splits: bool = True


class MPSNumericalError(RuntimeError):
    """A split retained a zero or non-finite spectrum (mps_state.py:66-86)."""


def _c128(a, shape) -> np.ndarray:
    arr = np.asarray(a, dtype=np.complex128)
    if arr.shape != shape:
        raise ValueError("Gate dimension does not match target qubits")
    return np.ascontiguousarray(arr.reshape(-1).view(np.float64))


class MPSState:
    """MPSState (mps_state.py:250-1549) on the device."""

    def __init__(self, n_qubits: int, opts: Optional[MPSOptions] = None, device: int = 0):
        self.n = int(n_qubits)
        if self.n <= 0:
            raise ValueError("n_qubits must be positive")
        self.opts = opts or MPSOptions()
        if self.opts.dmax < 1:
            raise ValueError("MPS dmax must be at least 1")
        if self.opts.eps < 0.0:
            raise ValueError("MPS eps must be non-negative")
        self.tensor_device = "gpu"
        self.svd_device = "gpu"
        h = ctypes.c_void_p()
        nat.check(nat.lib().msv_mps_create(self.n, int(self.opts.dmax), float(self.opts.eps),
                                           1 if self.opts.renormalize_splits else 0, int(device),
                                           ctypes.byref(h)), "mps_create")
        self._h = h

    # -------------- lifetime --------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            nat.lib().msv_mps_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - interpreter shutdown ordering
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _call(self, fn, *args, what=""):
        try:
            nat.check(fn(self._h, *args), what)
        except ValueError as exc:
            if any(k in str(exc) for k in _NUMERICAL):
                raise MPSNumericalError(str(exc)) from exc
            raise

    def reset(self) -> None:
        self._call(nat.lib().msv_mps_reset, what="mps_reset")

    # -------------- gates --------------
    def apply_single(self, U, q: int) -> None:
        if not 0 <= int(q) < self.n:
            raise ValueError("Qubit index out of range")
        m = _c128(U, (2, 2))
        self._call(nat.lib().msv_mps_apply_1q, int(q), nat.as_ptr(m, ctypes.c_double), what="mps_apply_1q")

    def apply_two(self, U4, c: int, t: int) -> None:
        c, t = int(c), int(t)
        if c == t:
            raise ValueError("Control and target must differ")
        if not (0 <= c < self.n and 0 <= t < self.n):
            raise ValueError("Qubit index out of range")
        m = _c128(U4, (4, 4))
        self._call(nat.lib().msv_mps_apply_2q, c, t, nat.as_ptr(m, ctypes.c_double), what="mps_apply_2q")

    def apply_diagonal(self, diagonal, wires: List[int]) -> None:
        """One- or two-qubit diagonal gates (mps_state.py:568-586)."""
        wires = [int(w) for w in wires]
        values = np.asarray(diagonal, dtype=np.complex128).reshape(-1)
        if len(wires) == 1:
            if values.size != 2:
                raise ValueError("one-qubit diagonal gates need two entries")
            self.apply_single(np.diag(values), wires[0])
            return
        if len(wires) != 2:
            raise ValueError("MPS diagonal gates support one or two qubits")
        if values.size != 4:
            raise ValueError("two-qubit diagonal gates need four entries")
        self.apply_two(np.diag(values), wires[0], wires[1])

    def apply_zz_phase(self, theta: float, first: int, second: int) -> None:
        """exp(-i theta Z x Z) (mps_state.py:680-709, 770-785)."""
        if int(first) == int(second):
            raise ValueError("ZZ phase operands must differ")
        even = complex(math.cos(theta), -math.sin(theta))
        odd = complex(math.cos(theta), math.sin(theta))
        self.apply_two(np.diag([even, odd, odd, even]), first, second)

    def apply_dense_gate(self, gate, qubits) -> None:
        qubits = [int(q) for q in qubits]
        if len(qubits) == 1:
            self.apply_single(gate, qubits[0])
        elif len(qubits) == 2:
            self.apply_two(gate, qubits[0], qubits[1])
        else:
            raise ValueError("MPS dense gates support one or two qubits")

    # -------------- diagnostics --------------
    def _stats(self) -> nat.msv_mps_stats:
        st = nat.msv_mps_stats()
        self._call(nat.lib().msv_mps_get_stats, ctypes.byref(st), what="mps_get_stats")
        return st

    def bond_dims(self) -> List[int]:
        out = np.zeros(max(0, self.n - 1), dtype=np.int32)
        if self.n > 1:
            self._call(nat.lib().msv_mps_bond_dims, nat.as_ptr(out, ctypes.c_int32), what="mps_bond_dims")
        return [int(x) for x in out]

    def bond_max(self) -> int:
        return int(self._stats().max_bond_ever)

    def bond_mean(self) -> float:
        b = self.bond_dims()
        return float(sum(b)) / len(b) if b else 1.0

    def truncated(self) -> bool:
        return bool(self._stats().truncated)

    def trunc_count(self) -> int:
        return int(self._stats().trunc_events)

    @property
    def canonical_center(self) -> int:
        return int(self._stats().center)

    def truncation_diagnostics(self) -> dict:
        """The reference's counters (mps_state.py:1137-1230) that this
        engine keeps: events, discarded weights, SVD and two-site counts."""
        st = self._stats()
        bonds = self.bond_dims()
        return {
            "events": int(st.trunc_events),
            "truncated": bool(st.truncated),
            "local_discarded_weight_sum": float(st.local_discarded_weight_sum),
            "local_discarded_weight_max": float(st.local_discarded_weight_max),
            "relative_discarded_weight_sum": float(st.relative_discarded_weight_sum),
            "relative_discarded_weight_max": float(st.relative_discarded_weight_max),
            "tensor_device": self.tensor_device,
            "svd_device": self.svd_device,
            "svd_calls": int(st.svd_calls),
            "two_site_calls": int(st.two_site_calls),
            "swap_fast_path_calls": int(st.swap_calls),
            "configured_max_bond_dimension": int(self.opts.dmax),
            "configured_truncation_threshold": float(self.opts.eps),
            "current_bond_dimension_max": int(max(bonds, default=1)),
            "current_bond_dimension_mean": float(self.bond_mean()),
            "maximum_bond_dimension_reached": int(st.max_bond_ever),
            "canonical_center": int(st.center),
            "renormalization_count": int(st.renormalization_count),
            "state_norm": self.norm(),
            "approximation_warning": (
                "local discarded weights are telemetry, not a global fidelity bound"
                if st.truncated else None),
        }

    def tensors(self) -> List[np.ndarray]:
        """Host copies of the site tensors, (dl, 2, dr) complex64 each."""
        out = []
        for q in range(self.n):
            dl, dr = ctypes.c_int32(), ctypes.c_int32()
            self._call(nat.lib().msv_mps_site_tensor, q, None, ctypes.byref(dl), ctypes.byref(dr),
                       what="mps_site_tensor")
            t = np.empty((dl.value, 2, dr.value), dtype=np.complex64)
            self._call(nat.lib().msv_mps_site_tensor, q, ctypes.c_void_p(t.ctypes.data), None, None,
                       what="mps_site_tensor")
            out.append(t)
        return out

    # -------------- readout --------------
    def norm(self) -> float:
        out = ctypes.c_double()
        self._call(nat.lib().msv_mps_norm, ctypes.byref(out), what="mps_norm")
        return float(out.value)

    def normalize(self) -> float:
        """Divides the center tensor by the norm (mps_state.py:401-416)."""
        nrm = self.norm()
        if not math.isfinite(nrm) or nrm <= 0.0:
            raise MPSNumericalError("Cannot renormalize an MPS with zero or non-finite norm")
        self.apply_single(np.eye(2) / nrm, self.canonical_center)
        return nrm

    def expectation_product(self, operators: Dict[int, np.ndarray]) -> complex:
        """<psi| x_w O_w |psi> / <psi|psi> (mps_state.py:1301-1330)."""
        wires = [int(w) for w in operators]
        if len(set(wires)) != len(wires):
            raise ValueError("Duplicate qubit indices")
        if any(w < 0 or w >= self.n for w in wires):
            raise ValueError("Qubit index out of range")
        ops = np.concatenate([_c128(operators[w], (2, 2)) for w in operators]) if wires else np.zeros(8)
        w = np.asarray(wires, dtype=np.int32)
        re, im = ctypes.c_double(), ctypes.c_double()
        self._call(nat.lib().msv_mps_expval_product, nat.as_ptr(w, ctypes.c_int32),
                   nat.as_ptr(np.ascontiguousarray(ops), ctypes.c_double), len(wires), ctypes.byref(re),
                   ctypes.byref(im), what="mps_expval_product")
        return complex(re.value, im.value)

    def to_native_state(self):
        """The normalized dense state as an engine state (NativeState)."""
        from .native_state import NativeState
        st = NativeState(self.n, "complex64")
        try:
            self._call(nat.lib().msv_mps_to_state, st.handle, what="mps_to_state")
        except Exception:
            st.close()
            raise
        return st

    def to_statevector(self) -> np.ndarray:
        """mps_state.py:1417-1440: the normalized dense state, complex64."""
        with self.to_native_state() as st:
            return st.to_numpy()

    def probabilities(self, wires: Optional[List[int]] = None) -> List[float]:
        with self.to_native_state() as st:
            return [float(p) for p in st.probabilities(wires)]

    def probabilities_array(self, wires: Optional[List[int]] = None) -> np.ndarray:
        with self.to_native_state() as st:
            return np.asarray(st.probabilities(wires), dtype=np.float64)

    def sample_array(self, shots: int, wires: Optional[List[int]] = None, *, rng=None) -> np.ndarray:
        """Shot bits, first wire = MSB (mps_state.py:1442-1540): the per-call
        seed is drawn from ``rng`` like the statevector simulator's; the
        engine's counter-based sampler draws from the contracted state."""
        from .sim import unpack_bits
        shots = int(shots)
        if shots < 0:
            raise ValueError("shots must be non-negative")
        wires = list(range(self.n)) if wires is None else [int(w) for w in wires]
        if len(set(wires)) != len(wires) or any(w < 0 or w >= self.n for w in wires):
            raise ValueError("Qubit index out of range or duplicated")
        rng = np.random.default_rng() if rng is None else rng
        seed = int(rng.integers(0, np.iinfo(np.uint32).max, dtype=np.uint32))
        with self.to_native_state() as st:
            packed = st.sample(shots, wires, seed=seed)
        return unpack_bits(packed, len(wires))

    def sample(self, shots: int, wires: Optional[List[int]] = None) -> List[List[int]]:
        return self.sample_array(shots, wires).tolist()

    def sample_counts(self, shots: int, wires: Optional[List[int]] = None) -> Dict[str, int]:
        counts: Dict[str, int] = {}
        for bits in self.sample(shots, wires):
            key = "".join(str(b) for b in bits)
            counts[key] = counts.get(key, 0) + 1
        return counts


def apply_device_gate(mps: MPSState, g) -> None:
    """One device-gate record (msv_gate, from the reference's op routing in
    mettleq_amd.fusion.gates_for_op) on the MPS: a 2x2 with at most one
    control, a one- or two-qubit diagonal, or an uncontrolled SWAP. Gates on
    three qubits (Toffoli, Fredkin) are not MPS operations here."""
    kind, nc = int(g.kind), int(g.ncontrols)
    m = [complex(g.m[2 * i], g.m[2 * i + 1]) for i in range(4)]
    if kind == nat.MSV_G_U:
        u = np.array(m, dtype=np.complex128).reshape(2, 2)
        if nc == 0:
            mps.apply_single(u, g.q[0])
        elif nc == 1:
            cu = np.eye(4, dtype=np.complex128)
            cu[2:, 2:] = u
            mps.apply_two(cu, g.q[1], g.q[0])
        else:
            raise ValueError("the MPS method applies gates on at most two qubits")
    elif kind == nat.MSV_G_DIAG1:
        mps.apply_single(np.diag(m[:2]), g.q[0])
    elif kind == nat.MSV_G_DIAG2:
        mps.apply_two(np.diag(m), g.q[0], g.q[1])
    elif kind == nat.MSV_G_SWAP and nc == 0:
        sw = np.eye(4, dtype=np.complex128)[[0, 2, 1, 3]]
        mps.apply_two(sw, g.q[0], g.q[1])
    else:
        raise ValueError("the MPS method applies gates on at most two qubits")
