"""CPU restatement of the reference statevector semantics (TEST INFRASTRUCTURE).

This module is the checker, not the product: only tests, the smoke check and
the CPU-baseline leg of bench.py may import it. It restates, in NumPy
complex128, what the reference engine computes on its statevector path, with
the reference file:line each function follows (paths relative to the
reference checkout's ``src/mettleq``):

- gate matrices ................ gates.py:5-234, device.py:53-70
- dense k-qubit apply .......... sim.py:47-73 (the METTLEQ_DENSE_ONLY path,
                                 device.py:1146-1151, 1099-1136)
- per-gate structured applies .. sim.py:75-264
- QFT without final swaps ...... sim.py:522-563; closed form of
                                 tests/mlxQQmlQftSubsetTest.py:34-65
- GF(2) affine inverse ......... shaders/xor_affine.py:46-76, gather 22-32
- phase-layer LUT semantics .... shaders/zz.py:78-84, shaders/diag.py:89-98
- probabilities / marginals .... sim.py:322-343
- Pauli-word expectation ....... integrations/_common.py:384-427
- layered-ansatz energies ...... integrations/qiskit.py:952-1018 (EstimatorV2:
                                 one circuit per binding, <psi|H|psi>);
                                 gradients by parameter shift (the PennyLane
                                 device's default differentiation,
                                 integrations/pennylane.py:558-625) and by
                                 FP64 central differences
- sampling ..................... this engine's own counter-based algorithm
                                 (Philox4x32-10 + fixed-order FP64 block CDF);
                                 the reference draws with mx.random.categorical
                                 (sim.py:405-413), whose stream is not
                                 reproducible without MLX.

Pinning (see tests/test_oracle.py): closed forms (QFT matrix, Gray-code
inverse of the CNOT chain, ZZ/CZ diagonals), gate identities and truth tables
from the reference tests, brute-force dense products, and the published
Philox4x32-10 known-answer vectors (Salmon et al., SC'11, Random123
kat_vectors).
"""
from __future__ import annotations

import math
from typing import Iterable, List, Sequence, Tuple

import numpy as np

C128 = np.complex128

# ---------------------------------------------------------------------------
# gate matrices (gates.py:5-234)
# ---------------------------------------------------------------------------


def I():
    return np.eye(2, dtype=C128)


def X():
    return np.array([[0, 1], [1, 0]], dtype=C128)


def Y():
    return np.array([[0, -1j], [1j, 0]], dtype=C128)


def Z():
    return np.array([[1, 0], [0, -1]], dtype=C128)


def H():
    s = 1.0 / math.sqrt(2.0)
    return np.array([[s, s], [s, -s]], dtype=C128)


def S():
    return np.array([[1, 0], [0, 1j]], dtype=C128)


def SDG():
    return np.array([[1, 0], [0, -1j]], dtype=C128)


def T():
    return np.array([[1, 0], [0, complex(math.cos(math.pi / 4), math.sin(math.pi / 4))]], dtype=C128)


def TDG():
    return np.array([[1, 0], [0, complex(math.cos(-math.pi / 4), math.sin(-math.pi / 4))]], dtype=C128)


def SX():
    a, b = complex(0.5, 0.5), complex(0.5, -0.5)
    return np.array([[a, b], [b, a]], dtype=C128)


def RX(theta):
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    return np.array([[c, -1j * s], [-1j * s, c]], dtype=C128)


def RY(theta):
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    return np.array([[c, -s], [s, c]], dtype=C128)


def RZ(theta):
    h = theta / 2
    return np.array([[complex(math.cos(-h), math.sin(-h)), 0], [0, complex(math.cos(h), math.sin(h))]], dtype=C128)


def PhaseShift(phi):
    return np.array([[1, 0], [0, complex(math.cos(phi), math.sin(phi))]], dtype=C128)


U1 = PhaseShift


def U2(phi, lam):
    """gates.py:80-82: RZ(phi) . RY(pi/2) . RZ(lam)."""
    return RZ(phi) @ RY(math.pi / 2) @ RZ(lam)


def U3(theta, phi, lam):
    """gates.py:85-87: RZ(phi) . RY(theta) . RZ(lam) (differs from Qiskit's U by
    the global phase exp(-i(phi+lam)/2))."""
    return RZ(phi) @ RY(theta) @ RZ(lam)


def _controlled(u):
    m = np.eye(4, dtype=C128)
    m[2:, 2:] = u
    return m


def CNOT():
    return _controlled(X())


def CZ():
    return np.diag([1, 1, 1, -1]).astype(C128)


def CPHASE(phi):
    return np.diag([1, 1, 1, complex(math.cos(phi), math.sin(phi))]).astype(C128)


def CRX(t):
    return _controlled(RX(t))


def CRY(t):
    return _controlled(RY(t))


def CRZ(t):
    return _controlled(RZ(t))


def CH():
    return _controlled(H())


def SWAP():
    m = np.zeros((4, 4), dtype=C128)
    m[0, 0] = m[1, 2] = m[2, 1] = m[3, 3] = 1
    return m


def ISWAP():
    m = np.zeros((4, 4), dtype=C128)
    m[0, 0] = m[3, 3] = 1
    m[1, 2] = m[2, 1] = 1j
    return m


def TOFFOLI():
    m = np.eye(8, dtype=C128)
    m[6, 6] = m[7, 7] = 0
    m[6, 7] = m[7, 6] = 1
    return m


def FREDKIN():
    m = np.eye(8, dtype=C128)
    m[5, 5] = m[6, 6] = 0
    m[5, 6] = m[6, 5] = 1
    return m


def pauli_pair_phase(name, theta):
    """device.py:53-70: exp(-i*theta*P(x)P) for ZZ/XX/YY."""
    c = math.cos(theta)
    ms = -1j * math.sin(theta)
    if name == "ZZPHASE":
        em, ep = complex(c, -math.sin(theta)), complex(c, math.sin(theta))
        return np.diag([em, ep, ep, em]).astype(C128)
    if name == "XXPHASE":
        return np.array([[c, 0, 0, ms], [0, c, ms, 0], [0, ms, c, 0], [ms, 0, 0, c]], dtype=C128)
    if name == "YYPHASE":
        return np.array([[c, 0, 0, -ms], [0, c, ms, 0], [0, ms, c, 0], [-ms, 0, 0, c]], dtype=C128)
    raise ValueError(name)


_CONST = {
    "I": I, "H": H, "X": X, "Y": Y, "Z": Z, "S": S, "SDG": SDG, "T": T, "TDG": TDG,
    "SX": SX, "CNOT": CNOT, "CX": CNOT, "CZ": CZ, "CH": CH, "SWAP": SWAP, "ISWAP": ISWAP,
    "TOFFOLI": TOFFOLI, "CCX": TOFFOLI, "FREDKIN": FREDKIN, "CSWAP": FREDKIN,
}


def gate_matrix(name: str, params: Sequence[float] = ()) -> np.ndarray:
    """device.py:1099-1136 (_dense_gate_for)."""
    name = str(name).upper()
    if name in ("SDAG", "S†"):
        name = "SDG"
    if name in ("TDAG", "T†"):
        name = "TDG"
    if name in _CONST:
        return _CONST[name]()
    p = list(params)
    table = {
        "RX": lambda: RX(p[0]), "RY": lambda: RY(p[0]), "RZ": lambda: RZ(p[0]),
        "U1": lambda: U1(p[0]), "U2": lambda: U2(p[0], p[1]), "U3": lambda: U3(p[0], p[1], p[2]),
        "CP": lambda: CPHASE(p[0]), "CPHASE": lambda: CPHASE(p[0]),
        "CRX": lambda: CRX(p[0]), "CRY": lambda: CRY(p[0]), "CRZ": lambda: CRZ(p[0]),
        "ZZPHASE": lambda: pauli_pair_phase("ZZPHASE", p[0]),
        "XXPHASE": lambda: pauli_pair_phase("XXPHASE", p[0]),
        "YYPHASE": lambda: pauli_pair_phase("YYPHASE", p[0]),
    }
    if name not in table:
        raise ValueError(f"Unsupported op for dense path: {name}")
    return table[name]()


# ---------------------------------------------------------------------------
# state application (qubit q = index bit n-1-q, sim.py:10-14)
# ---------------------------------------------------------------------------


def zero_state(n: int) -> np.ndarray:
    """sim.py:32-45."""
    s = np.zeros(1 << n, dtype=C128)
    s[0] = 1.0
    return s


def apply_dense(state: np.ndarray, n: int, gate: np.ndarray, qubits: Sequence[int]) -> np.ndarray:
    """sim.py:47-73: the first listed qubit is the most significant bit of the
    gate's local index."""
    qs = list(qubits)
    k = len(qs)
    if k == 0:
        return state
    if len(set(qs)) != k:
        raise ValueError("Duplicate qubit indices")
    gate = np.asarray(gate, dtype=C128)
    tensor = state.reshape([2] * n)
    perm = [i for i in range(n) if i not in qs] + qs
    t = tensor.transpose(perm).reshape(1 << (n - k), 1 << k)
    t = t @ gate.T
    t = t.reshape([2] * n).transpose(np.argsort(perm))
    return np.ascontiguousarray(t).reshape(1 << n)


def apply_controlled(state, n, u, controls: Sequence[int], target: int) -> np.ndarray:
    """Multi-controlled 2x2 (sim.py:140-201): U on target where all controls are 1."""
    cs = list(controls)
    k = len(cs)
    m = np.eye(1 << (k + 1), dtype=C128)
    m[-2:, -2:] = u
    return apply_dense(state, n, m, cs + [target])


def apply_op(state: np.ndarray, n: int, op: dict) -> np.ndarray:
    """One canonical op dict {"name", "wires", "parameters"} through the dense
    path (device.py:1146-1151)."""
    name = str(op["name"]).upper()
    if name == "I":
        return state
    return apply_dense(state, n, gate_matrix(name, op.get("parameters", [])), op["wires"])


def run_ops(n: int, ops: Iterable[dict], state: np.ndarray | None = None) -> np.ndarray:
    s = zero_state(n) if state is None else np.array(state, dtype=C128)
    for op in ops:
        s = apply_op(s, n, op)
    return s


def apply_1q_layer(state, n, mats: Sequence[np.ndarray]) -> np.ndarray:
    """u2_list_layer_all semantics (shaders/single_qubit.py:512-615): mats[q]
    acts on qubit q; gates on distinct qubits commute."""
    s = state
    for q, m in enumerate(mats):
        s = apply_dense(s, n, m, [q])
    return s


# ---------------------------------------------------------------------------
# QFT (sim.py:485-563) and its closed form
# ---------------------------------------------------------------------------


def qft(state, n, wires: Sequence[int]) -> np.ndarray:
    """sim.py:522-546: for j: H(w_j), then CPHASE(w_k -> w_j, pi/2^(k-j))."""
    w = list(wires)
    s = state
    for j in range(len(w)):
        s = apply_dense(s, n, H(), [w[j]])
        for k in range(j + 1, len(w)):
            s = apply_dense(s, n, CPHASE(math.pi / (2 ** (k - j))), [w[k], w[j]])
    return s


def iqft(state, n, wires: Sequence[int]) -> np.ndarray:
    """sim.py:549-563."""
    w = list(wires)
    s = state
    for j in reversed(range(len(w))):
        for k in reversed(range(j + 1, len(w))):
            s = apply_dense(s, n, CPHASE(-math.pi / (2 ** (k - j))), [w[k], w[j]])
        s = apply_dense(s, n, H(), [w[j]])
    return s


def bit_reverse(value: int, width: int) -> int:
    r = 0
    for b in range(width):
        r |= ((value >> b) & 1) << (width - 1 - b)
    return r


def qft_matrix(m: int) -> np.ndarray:
    """Closed form F[bitrev(row), :] with F_jk = exp(2 pi i jk/N)/sqrt(N)
    (tests/mlxQQmlQftSubsetTest.py:41-47)."""
    size = 1 << m
    idx = np.arange(size)
    f = np.exp(2j * np.pi * np.outer(idx, idx) / size) / math.sqrt(size)
    return f[[bit_reverse(i, m) for i in range(size)], :]


def qft_fft(psi: np.ndarray, m: int) -> np.ndarray:
    """The same transform as qft_matrix(m) @ psi in O(N log N): sqrt(N) * ifft
    followed by the bit-reversal permutation (sim.py:466-482 docstring)."""
    size = 1 << m
    f = np.fft.ifft(np.asarray(psi, dtype=C128)) * math.sqrt(size)
    idx = np.arange(size)
    rev = np.zeros_like(idx)
    for b in range(m):
        rev |= ((idx >> b) & 1) << (m - 1 - b)
    return f[rev]


# ---------------------------------------------------------------------------
# GF(2) affine permutations (shaders/xor_affine.py)
# ---------------------------------------------------------------------------


def compose_inverse_affine(ops: Iterable[dict], n: int) -> Tuple[List[int], int]:
    """xor_affine.py:46-76: replay the X/CNOT/SWAP run in reverse onto the
    identity map; rows[k] is the y-mask whose parity is source qubit k."""
    rows = [1 << (n - 1 - k) for k in range(n)]
    c = 0
    for op in reversed(list(ops)):
        name = str(op["name"]).upper()
        w = list(op["wires"])
        if name == "X":
            c ^= 1 << (n - 1 - w[0])
        elif name in ("CNOT", "CX"):
            a, b = w
            rows[b] ^= rows[a]
            if (c >> (n - 1 - a)) & 1:
                c ^= 1 << (n - 1 - b)
        elif name == "SWAP":
            a, b = w
            rows[a], rows[b] = rows[b], rows[a]
            ca, cb = (c >> (n - 1 - a)) & 1, (c >> (n - 1 - b)) & 1
            if ca != cb:
                c ^= (1 << (n - 1 - a)) | (1 << (n - 1 - b))
        else:
            raise ValueError(f"Not a GF(2) affine gate: {name}")
    return rows, c


def affine_source_indices(n: int, rows: Sequence[int], c: int) -> np.ndarray:
    """src(y) for every y (xor_affine.py:22-32), vectorized."""
    y = np.arange(1 << n, dtype=np.uint64)
    src = np.zeros_like(y)
    for k in range(n):
        par = np.zeros_like(y)
        m = np.uint64(rows[k])
        v = y & m
        while True:
            nz = v != 0
            if not nz.any():
                break
            par ^= (v & np.uint64(1))
            v >>= np.uint64(1)
        src |= par << np.uint64(n - 1 - k)
    return src ^ np.uint64(c)


def xor_affine_gather(state, n, rows, c) -> np.ndarray:
    return state[affine_source_indices(n, rows, c).astype(np.int64)]


# ---------------------------------------------------------------------------
# diagonal layers (shaders/zz.py, shaders/diag.py)
# ---------------------------------------------------------------------------


def zz_layer_diag(n, bonds, thetas) -> np.ndarray:
    """Product over bonds of exp(-i*theta_b*Z_a Z_b): phase_lut[m] =
    exp(-i*theta*(B-2m)) for m mismatched bonds (zz.py:78-84)."""
    idx = np.arange(1 << n, dtype=np.int64)
    ang = np.zeros(1 << n)
    for (a, b), t in zip(bonds, thetas):
        ba = (idx >> (n - 1 - a)) & 1
        bb = (idx >> (n - 1 - b)) & 1
        ang += -t * (1 - 2 * (ba ^ bb))
    return np.exp(1j * ang)


def cphase_layer_diag(n, bonds, thetas) -> np.ndarray:
    """Product over bonds of exp(i*theta_b) when both bits are set, with exact
    -1 for theta = pi (diag.py:89-98)."""
    idx = np.arange(1 << n, dtype=np.int64)
    d = np.ones(1 << n, dtype=C128)
    for (a, b), t in zip(bonds, thetas):
        both = ((idx >> (n - 1 - a)) & 1) & ((idx >> (n - 1 - b)) & 1)
        f = -1.0 + 0j if abs(t - math.pi) < 1e-12 else complex(math.cos(t), math.sin(t))
        d = np.where(both == 1, d * f, d)
    return d


# ---------------------------------------------------------------------------
# reductions
# ---------------------------------------------------------------------------


def probabilities(state, n, wires: Sequence[int] | None = None) -> np.ndarray:
    """sim.py:322-343: the first requested wire is the MSB of the output."""
    p = np.abs(np.asarray(state, dtype=C128)) ** 2
    if wires is None:
        return p
    sel = list(wires)
    if len(set(sel)) != len(sel):
        raise ValueError("Duplicate qubit indices")
    rest = [w for w in range(n) if w not in sel]
    t = p.reshape([2] * n).transpose(sel + rest).reshape(1 << len(sel), -1)
    return t.sum(axis=1)


def pauli_expectation(state, n, word: dict) -> complex:
    """<psi|P|psi> for {wire: 'X'|'Y'|'Z'} (_common.py:351-381)."""
    acted = np.array(state, dtype=C128)
    for w, p in word.items():
        m = {"X": X(), "Y": Y(), "Z": Z(), "I": I()}[p.upper()]
        acted = apply_dense(acted, n, m, [w])
    return complex(np.vdot(state, acted))


# ---------------------------------------------------------------------------
# sampling: Philox4x32-10 + fixed-order FP64 block CDF
# ---------------------------------------------------------------------------

_M0, _M1 = 0xD2511F53, 0xCD9E8D57
_W0, _W1 = 0x9E3779B9, 0xBB67AE85
_U32 = 0xFFFFFFFF


def philox4x32_10(ctr: Sequence[int], key: Sequence[int]) -> Tuple[int, int, int, int]:
    """Salmon et al. (SC'11) Philox4x32 with 10 rounds."""
    c0, c1, c2, c3 = (int(v) & _U32 for v in ctr)
    k0, k1 = (int(v) & _U32 for v in key)
    for _ in range(10):
        p0 = _M0 * c0
        p1 = _M1 * c2
        hi0, lo0 = p0 >> 32, p0 & _U32
        hi1, lo1 = p1 >> 32, p1 & _U32
        c0, c1, c2, c3 = (hi1 ^ c1 ^ k0) & _U32, lo1, (hi0 ^ c3 ^ k1) & _U32, lo0
        k0 = (k0 + _W0) & _U32
        k1 = (k1 + _W1) & _U32
    return c0, c1, c2, c3


def philox_u01(seed: int, stream_id: int, counter: int) -> float:
    """53-bit uniform in [0, 1): counter (shot, stream), key seed."""
    r = philox4x32_10(
        (counter & _U32, (counter >> 32) & _U32, stream_id & _U32, (stream_id >> 32) & _U32),
        (seed & _U32, (seed >> 32) & _U32),
    )
    a = r[0] >> 5
    b = r[1] >> 6
    return (float(a) * 67108864.0 + float(b)) * (1.0 / 9007199254740992.0)


BLOCK = 4096
LANES = 64


def amp_probs_c64(state_c64: np.ndarray) -> np.ndarray:
    """|a|^2 in FP64 from complex64 amplitudes: re*re + im*im, both products
    exact in double, one rounding for the sum."""
    a = np.asarray(state_c64, dtype=np.complex64)
    re = a.real.astype(np.float64)
    im = a.imag.astype(np.float64)
    return re * re + im * im


def _lane_sums(pb: np.ndarray) -> np.ndarray:
    """Lane l sums elements j*64 + l for j = 0..63, sequentially in j."""
    m = pb.reshape(BLOCK // LANES, LANES)
    s = np.zeros(LANES)
    for j in range(BLOCK // LANES):
        s = s + m[j]
    return s


def block_sums(p: np.ndarray) -> np.ndarray:
    """Per 4096-element block: lane sums, then the lane sums added in lane
    order (the fixed order the device kernel uses)."""
    nb = (len(p) + BLOCK - 1) // BLOCK
    pad = np.zeros(nb * BLOCK)
    pad[: len(p)] = p
    out = np.zeros(nb)
    for b in range(nb):
        ls = _lane_sums(pad[b * BLOCK:(b + 1) * BLOCK])
        t = 0.0
        for v in ls:
            t = t + float(v)
        out[b] = t
    return out


def sample_indices(p: np.ndarray, shots: int, seed: int, stream_id: int = 0) -> np.ndarray:
    """Inverse-CDF sampling over the fixed-order FP64 CDF: block by prefix
    search, then lane, then element (first position whose running sum
    exceeds the target)."""
    nb = (len(p) + BLOCK - 1) // BLOCK
    pad = np.zeros(nb * BLOCK)
    pad[: len(p)] = p
    bs = block_sums(p)
    prefix = [0.0]
    for v in bs:
        prefix.append(prefix[-1] + float(v))
    total = prefix[-1]
    if not total > 0.0:
        raise ValueError("zero-norm state")
    pref = np.asarray(prefix[1:])
    out = np.zeros(shots, dtype=np.uint64)
    for k in range(shots):
        t = philox_u01(seed, stream_id, k) * total
        if t >= total:
            t = float(np.nextafter(total, 0.0))
        b = int(np.searchsorted(pref, t, side="right"))
        b = min(b, nb - 1)
        r = t - prefix[b]
        blk = pad[b * BLOCK:(b + 1) * BLOCK]
        ls = _lane_sums(blk)
        c, L, r2, last = 0.0, -1, 0.0, -1
        for lane in range(LANES):
            v = float(ls[lane])
            if v > 0.0:
                last = lane
            cn = c + v
            if L < 0 and cn > r:
                L, r2 = lane, r - c
            c = cn
        if L < 0:
            L, r2 = (0 if last < 0 else last), math.inf
        c2, J, lastj = 0.0, -1, -1
        for j in range(BLOCK // LANES):
            v = float(blk[j * LANES + L])
            if v > 0.0:
                lastj = j
            cn = c2 + v
            if J < 0 and cn > r2:
                J = j
            c2 = cn
        if J < 0:
            J = 0 if lastj < 0 else lastj
        out[k] = b * BLOCK + J * LANES + L
    return out


def pack_wires(indices: np.ndarray, n: int, wires: Sequence[int]) -> np.ndarray:
    """Requested wires' bits of each index, first wire = MSB (sim.py:414-419)."""
    out = np.zeros(len(indices), dtype=np.uint64)
    for w in wires:
        out = (out << np.uint64(1)) | ((indices >> np.uint64(n - 1 - w)) & np.uint64(1))
    return out


# ---------------------------------------------------------------------------
# layered rotation ansatz: energies of a diagonal observable and gradients
# (the batched estimator / adjoint path of mettleq_amd.ansatz)
# ---------------------------------------------------------------------------


def ansatz_state(n: int, gates: Sequence[tuple], params: Sequence[float]) -> np.ndarray:
    """gates: ("RX"|"RY"|"RZ", qubit, param_index) or
    ("CPHASE"|"ZZPHASE", q0, q1, theta), applied in order from |0...0>."""
    s = zero_state(n)
    for g in gates:
        if g[0] in ("RX", "RY", "RZ"):
            s = apply_dense(s, n, gate_matrix(g[0], [params[g[2]]]), [g[1]])
        else:
            s = apply_dense(s, n, gate_matrix(g[0], [g[3]]), [g[1], g[2]])
    return s


def z_observable_diag(n: int, terms: Sequence[tuple]) -> np.ndarray:
    """h(x) = sum_t w_t (-1)^{sum of qubit bits of x in term t}; terms are
    (qubits, weight), qubit q = index bit n-1-q."""
    x = np.arange(1 << n, dtype=np.int64)
    h = np.zeros(1 << n)
    for qubits, w in terms:
        par = np.zeros(1 << n, dtype=np.int64)
        for q in qubits:
            par ^= (x >> (n - 1 - q)) & 1
        h += w * (1 - 2 * par)
    return h


def ansatz_energy(n: int, gates, params, terms) -> float:
    s = ansatz_state(n, gates, params)
    return float(np.sum(np.abs(s) ** 2 * z_observable_diag(n, terms)))


def ansatz_param_shift(n: int, gates, params, terms) -> np.ndarray:
    """Exact gradient of Pauli rotations: per gate occurrence,
    (E(theta + pi/2) - E(theta - pi/2)) / 2 with only that occurrence shifted."""
    params = np.asarray(params, dtype=np.float64)
    grad = np.zeros_like(params)
    for gi, g in enumerate(gates):
        if g[0] not in ("RX", "RY", "RZ"):
            continue
        vals = []
        for sh in (math.pi / 2, -math.pi / 2):
            s = zero_state(n)
            for gj, h in enumerate(gates):
                if h[0] in ("RX", "RY", "RZ"):
                    th = params[h[2]] + (sh if gj == gi else 0.0)
                    s = apply_dense(s, n, gate_matrix(h[0], [th]), [h[1]])
                else:
                    s = apply_dense(s, n, gate_matrix(h[0], [h[3]]), [h[1], h[2]])
            vals.append(float(np.sum(np.abs(s) ** 2 * z_observable_diag(n, terms))))
        grad[g[2]] += 0.5 * (vals[0] - vals[1])
    return grad


def ansatz_central_diff(n: int, gates, params, terms, h: float = 1e-4) -> np.ndarray:
    params = np.asarray(params, dtype=np.float64)
    grad = np.zeros_like(params)
    for p in range(len(params)):
        hi = params.copy()
        lo = params.copy()
        hi[p] += h
        lo[p] -= h
        grad[p] = (ansatz_energy(n, gates, hi, terms) - ansatz_energy(n, gates, lo, terms)) / (2 * h)
    return grad
