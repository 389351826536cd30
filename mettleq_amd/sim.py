"""StateVectorSimulator on a device-resident HIP statevector.

Mirrors the reference's per-gate simulator API (sim.py:17-563) -- same
method names, argument meaning, wire conventions (qubit q = index bit n-1-q)
and error types -- but every operation runs in place on the GPU through the
C ABI instead of building MLX expressions. ``state`` is a host NumPy copy
(download on read, upload on write); the amplitudes otherwise stay on the
device.
"""
from __future__ import annotations

import math
from typing import Iterable, List, Optional, Sequence

import numpy as np

from . import gates as G
from .execution import require_statevector_preflight
from .native_state import NativeState, gate_diag1, gate_diag2, gate_swap, gate_u


def canonical_axis_index(qubit: int, total: int) -> int:
    """sim.py:10-14."""
    if not (0 <= qubit < total):
        raise ValueError("Qubit index out of range")
    return qubit


class StateVectorSimulator:
    def __init__(self, n_qubits: int, *, allow_unsafe_statevector: Optional[bool] = None,
                 dtype: str = "complex64", device: int = 0, preflight_mode: str = "out_of_place",
                 reserved_bytes: int = 0):
        self.n = int(n_qubits)
        if self.n <= 0:
            raise ValueError("n_qubits must be positive")
        self.dtype = dtype
        self.preflight = require_statevector_preflight(
            self.n, dtype=dtype, allow_unsafe=allow_unsafe_statevector, mode=preflight_mode,
            reserved_bytes=reserved_bytes)
        self.native = NativeState(self.n, dtype, device)

    # -- state ---------------------------------------------------------------
    def reset(self) -> None:
        self.native.reset()

    @property
    def state(self) -> np.ndarray:
        return self.native.to_numpy()

    @state.setter
    def state(self, value) -> None:
        arr = np.asarray(value).reshape(-1)
        if arr.size != (1 << self.n):
            raise ValueError("state length does not match the register")
        self.native.from_numpy(arr)

    def close(self) -> None:
        self.native.close()

    # -- gates ---------------------------------------------------------------
    def _gates(self, gates) -> None:
        self.native.apply_gates(list(gates))

    def apply_dense_gate(self, gate, qubits: Iterable[int]) -> None:
        """Dense 2^k x 2^k gate, first listed qubit = MSB (sim.py:47-73).
        k = 1 and diagonal or controlled k >= 2 gates map to device kernels;
        other k = 2 gates are decomposed exactly into device gates."""
        qs = list(qubits)
        k = len(qs)
        if k == 0:
            return
        m = np.asarray(gate, dtype=np.complex128)
        if m.shape != (1 << k, 1 << k):
            raise ValueError("Gate dimension does not match target qubits")
        if len(set(qs)) != k:
            raise ValueError("Duplicate qubit indices")
        for q in qs:
            canonical_axis_index(q, self.n)
        if k == 1:
            self._gates([gate_u(qs[0], m)])
            return
        off = m - np.diag(np.diag(m))
        if not np.any(off):
            if k == 2:
                self._gates([gate_diag2(qs[0], qs[1], np.diag(m))])
                return
        # controlled-U structure: identity except the last 2x2 block
        eye = np.eye(1 << k, dtype=np.complex128)
        upper = m.copy()
        upper[-2:, -2:] = eye[-2:, -2:]
        if np.allclose(upper, eye, atol=0, rtol=0) and np.all(m[:-2, -2:] == 0) and np.all(m[-2:, :-2] == 0):
            self._gates([gate_u(qs[-1], m[-2:, -2:], qs[:-1])])
            return
        if k == 2:
            self._gates(_two_qubit_gates(m, qs[0], qs[1]))
            return
        raise NotImplementedError(
            f"dense {k}-qubit gates other than diagonal or multi-controlled 2x2 are not supported")

    def apply_single(self, gate, q: int) -> None:
        g = np.asarray(gate, dtype=np.complex128)
        if g.shape != (2, 2):
            raise ValueError("Gate dimension does not match target qubits")
        canonical_axis_index(q, self.n)
        self._gates([gate_u(q, g)])

    def apply_two(self, gate, c: int, t: int) -> None:
        if c == t:
            raise ValueError("Control and target must differ")
        self.apply_dense_gate(gate, [c, t])

    def apply_diagonal(self, diag, qubits: Iterable[int]) -> None:
        """Diagonal of length 2^k, first qubit = MSB (sim.py:94-138)."""
        qs = list(qubits)
        k = len(qs)
        if k == 0:
            return
        if len(set(qs)) != k:
            raise ValueError("Duplicate qubit indices")
        d = np.asarray(diag, dtype=np.complex128).reshape(-1)
        if d.shape[0] != (1 << k):
            raise ValueError("Diagonal dimension does not match target qubits")
        for q in qs:
            canonical_axis_index(q, self.n)
        if k == 1:
            self._gates([gate_diag1(qs[0], d[0], d[1])])
        elif k == 2:
            self._gates([gate_diag2(qs[0], qs[1], d)])
        else:
            # factor into one two-qubit diagonal per (first qubit, other)? not
            # generally possible; apply as controlled phases of the basis states
            gates = []
            for idx in range(1 << k):
                if d[idx] == 1:
                    continue
                bits = [(idx >> (k - 1 - j)) & 1 for j in range(k)]
                flips = [qs[j] for j in range(k) if bits[j] == 0]
                for f in flips:
                    gates.append(gate_u(f, G.CONST_1Q["X"]))
                gates.append(gate_u(qs[-1], np.diag([1.0, d[idx]]), qs[:-1]))
                for f in flips:
                    gates.append(gate_u(f, G.CONST_1Q["X"]))
            self._gates(gates)

    def apply_controlled_single(self, gate, c: int, t: int) -> None:
        if c == t:
            raise ValueError("Control and target must differ")
        g = np.asarray(gate, dtype=np.complex128)
        if g.shape != (2, 2):
            raise ValueError("Gate dimension does not match target qubit")
        canonical_axis_index(c, self.n)
        canonical_axis_index(t, self.n)
        self._gates([gate_u(t, g, [c])])

    def apply_multi_controlled_single(self, gate, controls: Iterable[int], target: int) -> None:
        cs = list(controls)
        if target in cs or len(set(cs)) != len(cs):
            raise ValueError("Control/target wires must be distinct")
        for q in cs + [target]:
            canonical_axis_index(q, self.n)
        self._gates([gate_u(target, np.asarray(gate, dtype=np.complex128), cs)])

    def apply_controlled_swap(self, c: int, a: int, b: int) -> None:
        if len({c, a, b}) != 3:
            raise ValueError("Control/target wires must be distinct")
        self._gates([gate_swap(a, b, [c])])

    def apply_swap_wires(self, a: int, b: int) -> None:
        if a == b:
            return
        canonical_axis_index(a, self.n)
        canonical_axis_index(b, self.n)
        self._gates([gate_swap(a, b)])

    def apply_xx_phase(self, theta: float, a: int, b: int) -> None:
        self._gates(_pauli_pair_gates("XXPHASE", theta, a, b))

    def apply_yy_phase(self, theta: float, a: int, b: int) -> None:
        self._gates(_pauli_pair_gates("YYPHASE", theta, a, b))

    def apply_zz_phase(self, theta: float, a: int, b: int) -> None:
        self._gates([gate_diag2(a, b, G.two_qubit_diagonal("ZZPHASE", [theta]))])

    def apply_zz_layer(self, theta: float, bonds: List[tuple]) -> None:
        """Fused exp(-i*theta*sum ZZ) in one pass (sim.py:271-300)."""
        bonds = [tuple(b) for b in bonds]
        if bonds:
            from . import layers
            layers.zz_chain_layer(self.native, theta, self.n, bonds)

    def h(self, q):
        self.apply_single(G.CONST_1Q["H"], q)

    def x(self, q):
        self.apply_single(G.CONST_1Q["X"], q)

    def z(self, q):
        self.apply_diagonal([1, -1], [q])

    def cnot(self, c, t):
        self.apply_controlled_single(G.CONST_1Q["X"], c, t)

    def cz(self, c, t):
        self.apply_diagonal([1, 1, 1, -1], [c, t])

    def cphase(self, c, t, phi):
        self.apply_diagonal([1, 1, 1, complex(math.cos(phi), math.sin(phi))], [c, t])

    # -- measurement ---------------------------------------------------------
    def _check_wires(self, wires):
        sel = list(wires)
        if len(set(sel)) != len(sel):
            raise ValueError("Duplicate qubit indices")
        for w in sel:
            canonical_axis_index(w, self.n)
        return sel

    def probabilities_array(self, wires: Optional[List[int]] = None) -> np.ndarray:
        """FP64 probabilities; first requested wire = MSB (sim.py:322-343)."""
        if wires is None:
            return self.native.probabilities(None)
        return self.native.probabilities(self._check_wires(wires))

    def probabilities(self, wires: Optional[List[int]] = None) -> List[float]:
        return self.probabilities_array(wires).tolist()

    def project_measure(self, wires: List[int], outcome_bits: List[int]) -> None:
        """Project onto the given outcome and renormalize (sim.py:350-381)."""
        if len(wires) != len(outcome_bits):
            raise ValueError("wires and outcome_bits must match length")
        self._check_wires(wires)
        self._gates([gate_diag1(w, 1.0 if (b & 1) == 0 else 0.0, 1.0 if (b & 1) else 0.0)
                     for w, b in zip(wires, outcome_bits)])
        norm2 = self.native.norm2()
        if norm2 > 0:
            self.native.scale(1.0 / math.sqrt(norm2))

    def sample_array(self, shots: int, wires: Optional[List[int]] = None, *, rng=None) -> np.ndarray:
        """Shot bits, first wire = MSB (sim.py:383-419). The per-call seed is
        drawn from ``rng`` exactly like the reference; the draw itself uses
        this engine's counter-based generator."""
        shots = int(shots)
        if shots < 0:
            raise ValueError("shots must be non-negative")
        wires = list(range(self.n)) if wires is None else self._check_wires(wires)
        rng = np.random.default_rng() if rng is None else rng
        seed = int(rng.integers(0, np.iinfo(np.uint32).max, dtype=np.uint32))
        packed = self.native.sample(shots, wires, seed=seed, stream_id=0)
        return unpack_bits(packed, len(wires))

    def sample(self, shots: int, wires: Optional[List[int]] = None) -> List[List[int]]:
        return self.sample_array(shots, wires).tolist()

    def sample_counts(self, shots: int, wires: Optional[List[int]] = None):
        counts: dict = {}
        for bits in self.sample(shots, wires):
            key = "".join(str(b) for b in bits)
            counts[key] = counts.get(key, 0) + 1
        return counts


def unpack_bits(packed: np.ndarray, nw: int) -> np.ndarray:
    shifts = np.arange(nw - 1, -1, -1, dtype=np.uint64)
    return ((packed[:, None] >> shifts) & np.uint64(1)).astype(np.int64)


def _pauli_pair_gates(name: str, theta: float, a: int, b: int):
    """exp(-i theta PP) = (V x V) exp(-i theta ZZ) (V^dag x V^dag) with V Z V^dag = P."""
    zz = gate_diag2(a, b, G.two_qubit_diagonal("ZZPHASE", [theta]))
    h = G.CONST_1Q["H"]
    if name == "XXPHASE":
        v, vd = h, h
    else:
        s = G.CONST_1Q["S"]
        v = s @ h
        vd = v.conj().T
    return [gate_u(a, vd), gate_u(b, vd), zz, gate_u(a, v), gate_u(b, v)]


def _two_qubit_gates(m: np.ndarray, a: int, b: int):
    """Exact decomposition of an arbitrary 4x4 (first wire a = MSB) into
    device gates: the columns are mapped by four controlled rotations.
    Implemented as a sequence of multi-controlled 2x2 blocks via the
    standard two-level (Givens) factorisation."""
    m = np.array(m, dtype=np.complex128)
    dim = 4
    ops = []  # list of (i, j, 2x2 acting on basis pair (i, j)) applied in order
    work = m.copy()
    # reduce to diagonal with two-level unitaries acting on rows
    for col in range(dim - 1):
        for row in range(dim - 1, col, -1):
            x, y = work[row - 1, col], work[row, col]
            if abs(y) < 1e-300:
                continue
            r = math.hypot(abs(x), abs(y))
            g = np.array([[np.conj(x) / r, np.conj(y) / r], [-y / r, x / r]], dtype=np.complex128)
            work[[row - 1, row], :] = g @ work[[row - 1, row], :]
            ops.append((row - 1, row, g))
    d = np.diag(work)
    # m = G1^dag ... Gk^dag D ; apply D first then the inverse Givens in reverse
    gates = [gate_diag2(a, b, d)]
    for i, j, g in reversed(ops):
        gates.extend(_two_level(g.conj().T, i, j, a, b))
    return gates


def _two_level(u: np.ndarray, i: int, j: int, a: int, b: int):
    """2x2 u acting on basis states |i>, |j> of the (a, b) pair (a = MSB)."""
    bi = ((i >> 1) & 1, i & 1)
    bj = ((j >> 1) & 1, j & 1)
    x = G.CONST_1Q["X"]
    if bi[0] == bj[0]:
        # differ in b: controlled on a == bi[0]
        pre = [] if bi[0] == 1 else [gate_u(a, x)]
        return pre + [gate_u(b, _order(u, bi[1]), [a])] + pre
    if bi[1] == bj[1]:
        pre = [] if bi[1] == 1 else [gate_u(b, x)]
        return pre + [gate_u(a, _order(u, bi[0]), [b])] + pre
    # differ in both bits: conjugate by CNOT(a -> b) to make them differ in a only
    cx = gate_u(b, x, [a])
    # |i>,|j> -> after CX: bit b ^= bit a
    ti = (bi[0], bi[1] ^ bi[0])
    tj = (bj[0], bj[1] ^ bj[0])
    inner = _two_level(u, ti[0] * 2 + ti[1], tj[0] * 2 + tj[1], a, b)
    return [cx] + inner + [cx]


def _order(u: np.ndarray, first_bit: int) -> np.ndarray:
    """u acts on (|first_bit>, |1-first_bit>); return it in (|0>, |1>) order."""
    if first_bit == 0:
        return u
    return np.array([[u[1, 1], u[1, 0]], [u[0, 1], u[0, 0]]], dtype=np.complex128)


def qft(sim: StateVectorSimulator, wires: List[int]) -> None:
    """Ordered-wire QFT without terminal swaps (sim.py:522-546)."""
    wires = _validate_qft_wires(sim, wires)
    n = len(wires)
    if n == 0:
        return
    if wires == list(range(wires[0], wires[0] + n)):
        sim.native.apply_qft(wires[0], wires[0] + n, inverse=False)
        return
    gates = []
    for j in range(n):
        gates.append(gate_u(wires[j], G.CONST_1Q["H"]))
        for k in range(j + 1, n):
            gates.append(gate_diag2(wires[k], wires[j],
                                    G.two_qubit_diagonal("CPHASE", [math.pi / (2 ** (k - j))])))
    sim.native.apply_gates(gates)


def iqft(sim: StateVectorSimulator, wires: List[int]) -> None:
    """Exact inverse of :func:`qft` on the same ordered wires (sim.py:549-563)."""
    wires = _validate_qft_wires(sim, wires)
    n = len(wires)
    if n == 0:
        return
    if wires == list(range(wires[0], wires[0] + n)):
        sim.native.apply_qft(wires[0], wires[0] + n, inverse=True)
        return
    gates = []
    for j in reversed(range(n)):
        for k in reversed(range(j + 1, n)):
            gates.append(gate_diag2(wires[k], wires[j],
                                    G.two_qubit_diagonal("CPHASE", [-math.pi / (2 ** (k - j))])))
        gates.append(gate_u(wires[j], G.CONST_1Q["H"]))
    sim.native.apply_gates(gates)


def _validate_qft_wires(sim, wires) -> List[int]:
    ordered = list(wires)
    if len(set(ordered)) != len(ordered):
        raise ValueError("Duplicate QFT wire indices")
    for w in ordered:
        canonical_axis_index(w, sim.n)
    return ordered
