"""Independent pins of every gate matrix and state routine of the CPU oracle.

Each oracle gate is compared with the matrix exponential of its generator
(scipy.linalg.expm over Pauli matrices built here from the computational
basis), with projector/block constructions, or with truth tables. None of
these restates the oracle's own formula, so a dropped factor, a wrong sign or
a halved angle fails at least one check. Conventions follow the reference:
RP(t) = exp(-i t P / 2) (gates.py:54-72), PhaseShift(p) = diag(1, e^{ip})
(gates.py:75-77), ZZ/XX/YYPHASE(t) = exp(-i t P (x) P) (device.py:53-70),
controlled gates act on |control, target> with the control as the most
significant bit (gates.py:158-170, 227-234).
"""
import math

import numpy as np
import pytest
from scipy.linalg import expm

from oracle import sv_oracle as O

# Pauli matrices from their action on |0>, |1> (independent of the oracle).
_E = np.eye(2, dtype=complex)
_P0 = np.outer(_E[0], _E[0])          # |0><0|
_P1 = np.outer(_E[1], _E[1])          # |1><1|
_SX = np.outer(_E[1], _E[0]) + np.outer(_E[0], _E[1])    # |1><0| + |0><1|
_SZ = _P0 - _P1
_SY = -1j * _SZ @ _SX                 # Y = -i Z X  (so that X Y = i Z)
I2 = np.eye(2, dtype=complex)

ANGLES = [0.0, 0.3, -1.1, 2.0, math.pi, -math.pi / 2, 5.9]


def _kron(*ms):
    out = np.eye(1, dtype=complex)
    for m in ms:
        out = np.kron(out, m)
    return out


def _close(a, b, tol=1e-12):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) < tol


def test_pauli_algebra():
    assert _close(O.X(), _SX) and _close(O.Z(), _SZ) and _close(O.Y(), _SY)
    assert _close(O.X() @ O.Y(), 1j * O.Z())
    assert _close(O.Y() @ O.Z(), 1j * O.X())
    for p in (O.X(), O.Y(), O.Z()):
        assert _close(p @ p, I2) and _close(p, p.conj().T)
    assert _close(O.I(), I2)


def test_hadamard_and_clifford_phases():
    assert _close(O.H(), (_SX + _SZ) / math.sqrt(2))
    assert _close(O.H() @ _SX @ O.H(), _SZ)
    # S = e^{i pi/4} RZ(pi/2), T = e^{i pi/8} RZ(pi/4), daggers are adjoints
    assert _close(O.S(), np.exp(1j * math.pi / 4) * expm(-1j * math.pi / 4 * _SZ))
    assert _close(O.T(), np.exp(1j * math.pi / 8) * expm(-1j * math.pi / 8 * _SZ))
    assert _close(O.SDG(), expm(-1j * math.pi / 4 * _SZ).conj().T * np.exp(-1j * math.pi / 4))
    assert _close(O.SDG(), O.S().conj().T) and _close(O.SDG() @ O.S(), I2)
    assert _close(O.TDG(), O.T().conj().T) and _close(O.TDG() @ O.T(), I2)
    # SX is the principal square root of X: e^{i pi/4} exp(-i pi/4 X)
    assert _close(O.SX(), np.exp(1j * math.pi / 4) * expm(-1j * math.pi / 4 * _SX))


@pytest.mark.parametrize("t", ANGLES)
def test_rotations_are_exponentials_of_half_angles(t):
    assert _close(O.RX(t), expm(-0.5j * t * _SX))
    assert _close(O.RY(t), expm(-0.5j * t * _SY))
    assert _close(O.RZ(t), expm(-0.5j * t * _SZ))
    # PhaseShift(t) = U1(t): |0> untouched, |1> picks up e^{it}
    assert _close(O.PhaseShift(t), _P0 + np.exp(1j * t) * _P1)
    assert _close(O.U1(t), np.exp(0.5j * t) * expm(-0.5j * t * _SZ))


@pytest.mark.parametrize("t", ANGLES)
def test_controlled_rotations_block_structure(t):
    for name, p in (("CRX", _SX), ("CRY", _SY), ("CRZ", _SZ)):
        got = O.gate_matrix(name, [t])
        assert _close(got, expm(-0.5j * t * _kron(_P1, p)))
        assert _close(got, _kron(_P0, I2) + _kron(_P1, expm(-0.5j * t * p)))
    assert _close(O.CPHASE(t), expm(1j * t * _kron(_P1, _P1)))
    assert _close(O.gate_matrix("CP", [t]), O.CPHASE(t))


def test_two_qubit_constant_gates():
    assert _close(O.CNOT(), _kron(_P0, I2) + _kron(_P1, _SX))
    assert _close(O.CZ(), np.eye(4) - 2 * _kron(_P1, _P1))
    assert _close(O.CH(), _kron(_P0, I2) + _kron(_P1, (_SX + _SZ) / math.sqrt(2)))
    swap = (np.eye(4) + _kron(_SX, _SX) + _kron(_SY, _SY) + _kron(_SZ, _SZ)) / 2
    assert _close(O.SWAP(), swap)
    assert _close(O.ISWAP(), expm(0.25j * math.pi * (_kron(_SX, _SX) + _kron(_SY, _SY))))
    # CNOT truth table with the control as the first (most significant) wire
    for x in range(4):
        c, tq = x >> 1, x & 1
        col = O.CNOT()[:, x]
        assert np.argmax(np.abs(col)) == (c << 1) | (tq ^ c)


@pytest.mark.parametrize("t", ANGLES)
def test_pauli_pair_phases_are_exp_minus_i_theta_pp(t):
    """device.py:53-70: exp(-i theta P (x) P), the angle NOT halved (the SDK
    adapters halve rxx/ryy/rzz angles before building these)."""
    for name, p in (("ZZPHASE", _SZ), ("XXPHASE", _SX), ("YYPHASE", _SY)):
        want = expm(-1j * t * _kron(p, p))
        assert _close(O.pauli_pair_phase(name, t), want)
        assert _close(O.gate_matrix(name, [t]), want)


def test_multi_qubit_truth_tables_from_projectors():
    tof = _kron(np.eye(4) - _kron(_P1, _P1), I2) + _kron(_P1, _P1, _SX)
    assert _close(O.TOFFOLI(), tof)
    fred = _kron(_P0, np.eye(4)) + _kron(_P1, O.SWAP())
    assert _close(O.FREDKIN(), fred)


def test_gate_matrix_names_and_aliases():
    cases = {
        "I": I2, "X": _SX, "Y": _SY, "Z": _SZ, "H": (_SX + _SZ) / math.sqrt(2),
        "SDAG": O.S().conj().T, "TDAG": O.T().conj().T,
        "CX": _kron(_P0, I2) + _kron(_P1, _SX),
        "CCX": O.TOFFOLI(), "CSWAP": O.FREDKIN(),
    }
    for name, want in cases.items():
        assert _close(O.gate_matrix(name), want), name
    t, p, l = 0.4, -0.8, 1.9
    assert _close(O.gate_matrix("U3", [t, p, l]),
                  expm(-0.5j * p * _SZ) @ expm(-0.5j * t * _SY) @ expm(-0.5j * l * _SZ))
    assert _close(O.gate_matrix("U2", [p, l]),
                  expm(-0.5j * p * _SZ) @ expm(-0.25j * math.pi * _SY) @ expm(-0.5j * l * _SZ))
    with pytest.raises(ValueError):
        O.gate_matrix("NOT_A_GATE")


# --- state routines against explicit Kronecker products ----------------------


def _embed(n, gate, qubits):
    """Full 2^n x 2^n operator of `gate` on `qubits` (first = MSB of the
    gate's index), by summing over basis states; qubit q = index bit n-1-q."""
    k = len(qubits)
    dim = 1 << n
    full = np.zeros((dim, dim), dtype=complex)
    for col in range(dim):
        sub_in = 0
        for j, q in enumerate(qubits):
            sub_in |= ((col >> (n - 1 - q)) & 1) << (k - 1 - j)
        for sub_out in range(1 << k):
            amp = gate[sub_out, sub_in]
            if amp == 0:
                continue
            row = col
            for j, q in enumerate(qubits):
                bit = (sub_out >> (k - 1 - j)) & 1
                row = (row & ~(1 << (n - 1 - q))) | (bit << (n - 1 - q))
            full[row, col] += amp
    return full


@pytest.mark.parametrize("qubits", [[0], [3], [1, 3], [3, 0], [2, 0, 4]])
def test_apply_dense_matches_embedded_operator(qubits):
    n = 5
    rng = np.random.default_rng(len(qubits) * 7 + qubits[0])
    k = len(qubits)
    g = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    assert _close(O.apply_dense(psi, n, g, qubits), _embed(n, g, qubits) @ psi, 1e-10)
    if qubits == [0]:
        # qubit 0 is the leftmost Kronecker factor
        assert _close(O.apply_dense(psi, n, g, [0]), _kron(g, np.eye(16)) @ psi, 1e-10)


def test_apply_controlled_and_1q_layer():
    n = 4
    rng = np.random.default_rng(11)
    psi = rng.normal(size=16) + 1j * rng.normal(size=16)
    u = O.RX(0.7) @ O.RZ(-0.3)
    # U on qubit 3 when qubits 0 and 2 are both set
    ctl = _kron(_P1, I2, _P1, u) + (np.eye(16) - _kron(_P1, I2, _P1, I2))
    assert _close(O.apply_controlled(psi, n, u, [0, 2], 3), ctl @ psi, 1e-10)
    mats = [O.U3(*rng.uniform(0, 6, 3)) for _ in range(n)]
    assert _close(O.apply_1q_layer(psi, n, mats), _kron(*mats) @ psi, 1e-10)


def test_cphase_layer_diag_matches_dense_cphase_products():
    n = 5
    bonds = [(0, 1), (3, 1), (4, 2), (0, 4)]
    thetas = [0.3, math.pi, -2.1, 1.0]
    psi = np.random.default_rng(4).normal(size=32) + 0j
    want = psi
    for (a, b), t in zip(bonds, thetas):
        want = _embed(n, expm(1j * t * _kron(_P1, _P1)), [a, b]) @ want
    assert _close(psi * O.cphase_layer_diag(n, bonds, thetas), want, 1e-12)


def test_zz_layer_diag_matches_exponentials():
    n = 4
    bonds = [(0, 3), (2, 1), (1, 0)]
    thetas = [0.45, -1.3, 2.7]
    gen = sum(t * _embed(n, _kron(_SZ, _SZ), [a, b]) for (a, b), t in zip(bonds, thetas))
    assert _close(np.diag(expm(-1j * gen)), O.zz_layer_diag(n, bonds, thetas), 1e-12)


def test_marginal_probabilities_brute_force():
    n = 4
    rng = np.random.default_rng(8)
    psi = rng.normal(size=16) + 1j * rng.normal(size=16)
    p = np.abs(psi) ** 2
    for wires in ([2], [3, 0], [1, 2, 0], [0, 1, 2, 3]):
        want = np.zeros(1 << len(wires))
        for i in range(16):
            k = 0
            for w in wires:
                k = (k << 1) | ((i >> (n - 1 - w)) & 1)
            want[k] += p[i]
        assert _close(O.probabilities(psi, n, wires), want, 1e-12)


def test_pauli_expectation_against_kronecker_words():
    n = 3
    rng = np.random.default_rng(12)
    psi = rng.normal(size=8) + 1j * rng.normal(size=8)
    psi /= np.linalg.norm(psi)
    pm = {"X": _SX, "Y": _SY, "Z": _SZ, "I": I2}
    for word in ({0: "X", 2: "Y"}, {1: "Z"}, {0: "Y", 1: "Y", 2: "X"}):
        full = _kron(*[pm[word.get(q, "I")] for q in range(n)])
        assert abs(O.pauli_expectation(psi, n, word) - np.vdot(psi, full @ psi)) < 1e-12


def test_amp_probs_c64_exact_products():
    a = np.array([0.1 + 0.7j, -3e-4 + 2j, 1.5 - 0j], dtype=np.complex64)
    want = [float(np.float64(z.real)) ** 2 + float(np.float64(z.imag)) ** 2 for z in a]
    assert O.amp_probs_c64(a).tolist() == want


def test_qft_gate_sequence_matches_dense_hadamard_ladder():
    """sim.py:522-546 gate order, built from the expm-pinned CPHASE and H."""
    n = 4
    rng = np.random.default_rng(13)
    psi = rng.normal(size=16) + 1j * rng.normal(size=16)
    full = np.eye(16, dtype=complex)
    h = (_SX + _SZ) / math.sqrt(2)
    for j in range(n):
        full = _embed(n, h, [j]) @ full
        for k in range(j + 1, n):
            full = _embed(n, expm(1j * math.pi / 2 ** (k - j) * _kron(_P1, _P1)), [k, j]) @ full
    assert _close(O.qft(psi, n, list(range(n))), full @ psi, 1e-10)
    assert _close(O.iqft(full @ psi, n, list(range(n))), psi, 1e-10)


# --- the pins reject the plausible convention slips --------------------------


@pytest.mark.parametrize("mutant,pin", [
    (lambda t: O.pauli_pair_phase("XXPHASE", t / 2), lambda t: expm(-1j * t * _kron(_SX, _SX))),
    (lambda t: O.pauli_pair_phase("YYPHASE", -t), lambda t: expm(-1j * t * _kron(_SY, _SY))),
    (lambda t: O.pauli_pair_phase("ZZPHASE", t / 2), lambda t: expm(-1j * t * _kron(_SZ, _SZ))),
    (lambda t: O.RX(-t), lambda t: expm(-0.5j * t * _SX)),
    (lambda t: O.RX(2 * t), lambda t: expm(-0.5j * t * _SX)),
    (lambda t: O.PhaseShift(t / 2), lambda t: _P0 + np.exp(1j * t) * _P1),
    (lambda t: _kron(_P0, I2) + _kron(_P1, O.RY(-t)), lambda t: O.CRY(t)),
    (lambda t: O.ISWAP().conj(), lambda t: expm(0.25j * math.pi * (_kron(_SX, _SX) + _kron(_SY, _SY)))),
    (lambda t: O.S(), lambda t: O.SDG()),
])
def test_pins_reject_sign_and_half_angle_slips(mutant, pin):
    assert not _close(mutant(0.9), pin(0.9), 1e-6)
