"""Pins of the CPU oracle (oracle/sv_oracle.py, oracle/sv_cpu.c).

Each check is independent of the formula it pins: published known-answer
vectors, closed forms, truth tables and identities from the reference's own
tests, and brute force against dense products.
"""
import math

import numpy as np
import pytest

from oracle import sv_oracle as O

# --- Philox4x32-10 known-answer vectors (Random123 kat_vectors) -------------


@pytest.mark.parametrize("ctr,key,want", [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
])
def test_philox_known_answers(ctr, key, want):
    assert O.philox4x32_10(ctr, key) == want


def test_philox_c_restatement_matches_python():
    import ctypes
    from oracle.sv_cpu import lib
    for ctr, key in [((0, 0, 0, 0), (0, 0)), ((1, 2, 3, 4), (5, 6)), ((0xFFFFFFFF,) * 4, (7, 9))]:
        c = (ctypes.c_uint32 * 4)(*ctr)
        k = (ctypes.c_uint32 * 2)(*key)
        out = (ctypes.c_uint32 * 4)()
        lib().cpu_philox4x32_10(c, k, out)
        assert tuple(out) == O.philox4x32_10(ctr, key)
    for s in range(20):
        assert lib().cpu_philox_u01(123, 4, s) == O.philox_u01(123, 4, s)


def test_u01_range_and_uniformity():
    u = np.array([O.philox_u01(99, 0, k) for k in range(20000)])
    assert u.min() >= 0.0 and u.max() < 1.0
    hist, _ = np.histogram(u, bins=10, range=(0, 1))
    assert np.all(np.abs(hist - 2000) < 4 * math.sqrt(2000))


# --- gates and conventions (tests/mlxQCoreTest.py) ---------------------------


def test_gate_identities():
    eye = np.eye(2)
    assert np.allclose(O.H() @ O.H(), eye)
    assert np.allclose(O.X() @ O.X(), eye)
    assert np.allclose(O.S() @ O.S(), O.Z())
    assert np.allclose(O.T() @ O.T(), O.S())
    assert np.allclose(O.SX() @ O.SX(), O.X())
    # U3 = RZ(phi) RY(theta) RZ(lam) (gates.py:85-87) equals Qiskit's U up to
    # the global phase exp(-i(phi+lam)/2)
    t, p, l = 0.7, -1.1, 2.3
    qiskit_u = np.array([[math.cos(t / 2), -np.exp(1j * l) * math.sin(t / 2)],
                         [np.exp(1j * p) * math.sin(t / 2), np.exp(1j * (p + l)) * math.cos(t / 2)]])
    assert np.allclose(O.U3(t, p, l), np.exp(-0.5j * (p + l)) * qiskit_u)
    assert np.allclose(O.U2(p, l), O.U3(math.pi / 2, p, l))
    assert np.allclose(O.RZ(0.3), np.diag([np.exp(-0.15j), np.exp(0.15j)]))


def test_msb_first_convention():
    """qubit 0 is the most significant index bit (sim.py:10-14)."""
    n = 3
    s = O.apply_dense(O.zero_state(n), n, O.X(), [0])
    assert np.argmax(np.abs(s)) == 0b100
    s = O.apply_dense(O.zero_state(n), n, O.X(), [2])
    assert np.argmax(np.abs(s)) == 0b001


def test_toffoli_fredkin_truth_tables():
    n = 3
    for x in range(8):
        s = np.zeros(8, dtype=complex)
        s[x] = 1
        t = O.apply_dense(s, n, O.TOFFOLI(), [0, 1, 2])
        want = x ^ 1 if (x >> 1) == 0b11 else x
        assert np.argmax(np.abs(t)) == want
        f = O.apply_dense(s, n, O.FREDKIN(), [0, 1, 2])
        b0, b1, b2 = (x >> 2) & 1, (x >> 1) & 1, x & 1
        want = (b0 << 2) | ((b2 << 1) | b1 if b0 else (b1 << 1) | b2)
        assert np.argmax(np.abs(f)) == want


def test_bell_and_ghz():
    n = 3
    s = O.run_ops(n, [{"name": "H", "wires": [0]}, {"name": "CNOT", "wires": [0, 1]},
                      {"name": "CNOT", "wires": [1, 2]}])
    want = np.zeros(8)
    want[0] = want[7] = 1 / math.sqrt(2)
    assert np.allclose(s, want)


def test_bell_pauli_values():
    s = O.run_ops(2, [{"name": "H", "wires": [0]}, {"name": "CNOT", "wires": [0, 1]}])
    assert np.isclose(O.pauli_expectation(s, 2, {0: "X", 1: "X"}), 1)
    assert np.isclose(O.pauli_expectation(s, 2, {0: "Z", 1: "Z"}), 1)
    assert np.isclose(O.pauli_expectation(s, 2, {0: "Y", 1: "Y"}), -1)


# --- QFT (tests/mlxQQmlQftSubsetTest.py, mlxQCoreTest.py) --------------------


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7])
def test_gate_qft_equals_closed_form(n):
    rng = np.random.default_rng(n)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    assert np.allclose(O.qft(psi, n, list(range(n))), O.qft_matrix(n) @ psi, atol=1e-12)
    assert np.allclose(O.iqft(O.qft(psi, n, list(range(n))), n, list(range(n))), psi, atol=1e-12)


def test_qft_is_bit_reversed_inverse_fft():
    """sqrt(N) * ifft followed by bit reversal (sim.py:466-482 docstring)."""
    n = 6
    rng = np.random.default_rng(2)
    psi = rng.normal(size=64) + 1j * rng.normal(size=64)
    f = np.fft.ifft(psi) * math.sqrt(64)
    f = f[[O.bit_reverse(i, n) for i in range(64)]]
    assert np.allclose(O.qft_matrix(n) @ psi, f)


def test_qft_fft_form_equals_matrix_form():
    for n in (1, 4, 9):
        psi = np.random.default_rng(n).normal(size=1 << n) + 0j
        assert np.allclose(O.qft_fft(psi, n), O.qft_matrix(n) @ psi, atol=1e-12)


def test_qft_uniform_maps_to_zero():
    n = 5
    psi = np.full(32, 1 / math.sqrt(32))
    out = O.qft(psi, n, list(range(n)))
    assert np.isclose(abs(out[0]), 1.0)


# --- GF(2) affine maps (shaders/xor_affine.py) ---------------------------------


def test_compose_inverse_affine_inverts_the_block():
    """The reference's property test (tests/mlxQMetalKernelsTest.py:293-323)."""
    n = 6
    ops = ([{"name": "CNOT", "wires": [q, q + 1]} for q in range(n - 1)]
           + [{"name": "X", "wires": [3]}, {"name": "SWAP", "wires": [0, 5]},
              {"name": "CNOT", "wires": [4, 1]}])
    rows, c = O.compose_inverse_affine(ops, n)

    def forward(x):
        bits = [(x >> (n - 1 - q)) & 1 for q in range(n)]
        for op in ops:
            w = op["wires"]
            if op["name"] == "X":
                bits[w[0]] ^= 1
            elif op["name"] == "CNOT":
                bits[w[1]] ^= bits[w[0]]
            else:
                bits[w[0]], bits[w[1]] = bits[w[1]], bits[w[0]]
        return sum(b << (n - 1 - q) for q, b in enumerate(bits))

    src = O.affine_source_indices(n, rows, c)
    for x in range(1 << n):
        assert int(src[forward(x)]) == x


def test_cnot_chain_inverse_is_gray_code():
    n = 9
    rows, c = O.compose_inverse_affine([{"name": "CNOT", "wires": [i, i + 1]} for i in range(n - 1)], n)
    y = np.arange(1 << n, dtype=np.uint64)
    assert np.array_equal(O.affine_source_indices(n, rows, c), y ^ (y >> np.uint64(1)))


def test_affine_gather_matches_dense_gates():
    n = 5
    rng = np.random.default_rng(1)
    ops = []
    for _ in range(15):
        k = rng.integers(3)
        if k == 0:
            ops.append({"name": "X", "wires": [int(rng.integers(n))]})
        else:
            a, b = rng.choice(n, 2, replace=False)
            ops.append({"name": ["CNOT", "SWAP"][k - 1], "wires": [int(a), int(b)]})
    psi = rng.normal(size=32) + 1j * rng.normal(size=32)
    rows, c = O.compose_inverse_affine(ops, n)
    assert np.allclose(O.xor_affine_gather(psi, n, rows, c), O.run_ops(n, ops, psi))


# --- diagonal layers ------------------------------------------------------------


def test_zz_layer_diag_matches_dense_products():
    n = 5
    bonds = [(0, 1), (1, 2), (4, 0), (2, 4)]
    thetas = [0.3, -1.2, 2.0, 0.7]
    psi = np.random.default_rng(0).normal(size=32).astype(complex)
    want = psi
    for (a, b), t in zip(bonds, thetas):
        want = O.apply_dense(want, n, O.pauli_pair_phase("ZZPHASE", t), [a, b])
    assert np.allclose(psi * O.zz_layer_diag(n, bonds, thetas), want)


def test_zz_lut_closed_form():
    """phase_lut[m] = exp(-i theta (B - 2m)) for m mismatched bonds (zz.py:78-84)."""
    n, theta = 4, 0.37
    bonds = [(0, 1), (1, 2), (2, 3)]
    d = O.zz_layer_diag(n, bonds, [theta] * 3)
    for i in range(16):
        bits = [(i >> (n - 1 - q)) & 1 for q in range(n)]
        m = sum(bits[a] != bits[b] for a, b in bonds)
        assert np.isclose(d[i], np.exp(-1j * theta * (3 - 2 * m)))


def test_cz_layer_is_exact_sign():
    n = 6
    d = O.cphase_layer_diag(n, [(0, 1), (2, 3), (4, 5)], [math.pi] * 3)
    assert set(np.unique(d.real).tolist()) <= {-1.0, 1.0}
    assert np.all(d.imag == 0)


# --- probabilities (sim.py:322-343) ------------------------------------------


def test_marginal_order_is_msb_first():
    """First requested wire is the MSB of the output (tests/
    mlxQInternalConsistencyTest.py:148-162)."""
    n = 3
    s = np.zeros(8, dtype=complex)
    s[0b011] = 1  # qubit0=0, qubit1=1, qubit2=1
    assert np.argmax(O.probabilities(s, n, [1, 0])) == 0b10
    assert np.argmax(O.probabilities(s, n, [0, 2])) == 0b01
    assert np.argmax(O.probabilities(s, n, [2, 1, 0])) == 0b110


# --- sampling ------------------------------------------------------------------


def test_block_sums_fixed_order_equal_numpy_totals():
    rng = np.random.default_rng(3)
    p = rng.random(3 * 4096 + 17)
    bs = O.block_sums(p)
    assert len(bs) == 4
    assert np.allclose(bs, [p[:4096].sum(), p[4096:8192].sum(), p[8192:12288].sum(), p[12288:].sum()])


@pytest.mark.parametrize("n", [1, 3, 12, 14])
def test_sampling_python_and_c_restatements_bit_exact(n):
    from oracle.sv_cpu import CpuEngine
    rng = np.random.default_rng(n)
    psi = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)).astype(np.complex64)
    psi /= np.linalg.norm(psi)
    eng = CpuEngine(n, psi)
    got = eng.sample_indices(500, seed=42, stream_id=3)
    want = O.sample_indices(O.amp_probs_c64(eng.state), 500, 42, 3)
    assert np.array_equal(got, want)
    assert np.array_equal(eng.block_sums(), O.block_sums(O.amp_probs_c64(eng.state)))


def test_sampling_distribution_and_zero_probabilities():
    n = 4
    p = np.zeros(16)
    p[[1, 6, 13]] = [0.2, 0.3, 0.5]
    idx = O.sample_indices(p, 6000, 7)
    assert set(np.unique(idx).tolist()) == {1, 6, 13}
    freq = np.array([np.mean(idx == k) for k in (1, 6, 13)])
    assert np.all(np.abs(freq - [0.2, 0.3, 0.5]) < 0.03)


def test_pack_wires_msb_first():
    idx = np.array([0b101, 0b011], dtype=np.uint64)
    assert O.pack_wires(idx, 3, [0, 2]).tolist() == [0b11, 0b01]
    assert O.pack_wires(idx, 3, [1]).tolist() == [0, 1]


# --- C restatement of the CPU path vs the dense oracle --------------------------


@pytest.mark.parametrize("n", [3, 6, 10])
def test_cpu_engine_matches_dense_oracle(n):
    from oracle.sv_cpu import CpuEngine
    rng = np.random.default_rng(50 + n)
    names = ["H", "X", "Y", "Z", "S", "T", "SDG", "RX", "RY", "RZ", "U1", "U3", "CNOT", "CZ",
             "CPHASE", "CRX", "CRY", "CRZ", "CH", "SWAP", "ZZPHASE", "XXPHASE", "YYPHASE", "ISWAP",
             "TOFFOLI", "FREDKIN"]
    ops = []
    for _ in range(60):
        nm = names[int(rng.integers(len(names)))]
        arity = 3 if nm in ("TOFFOLI", "FREDKIN") else (2 if nm in (
            "CNOT", "CZ", "CPHASE", "CRX", "CRY", "CRZ", "CH", "SWAP", "ZZPHASE", "XXPHASE",
            "YYPHASE", "ISWAP") else 1)
        if arity > n:
            continue
        wires = [int(x) for x in rng.choice(n, arity, replace=False)]
        npar = {"RX": 1, "RY": 1, "RZ": 1, "U1": 1, "U3": 3, "CPHASE": 1, "CRX": 1, "CRY": 1,
                "CRZ": 1, "ZZPHASE": 1, "XXPHASE": 1, "YYPHASE": 1}.get(nm, 0)
        ops.append({"name": nm, "wires": wires, "parameters": rng.uniform(-3, 3, npar).tolist()})
    eng = CpuEngine(n)
    got = eng.run(ops)
    want = O.run_ops(n, ops)
    assert np.max(np.abs(got - want)) < 1e-5


def test_cpu_engine_qft_and_uniform_zz_layer():
    from oracle.sv_cpu import CpuEngine
    n = 8
    rng = np.random.default_rng(9)
    psi = (rng.normal(size=256) + 1j * rng.normal(size=256))
    psi /= np.linalg.norm(psi)
    eng = CpuEngine(n, psi)
    eng.qft(list(range(n)))
    assert np.max(np.abs(eng.state - O.qft_matrix(n) @ psi)) < 2e-6
    eng.qft(list(range(n)), inverse=True)
    assert np.max(np.abs(eng.state - psi)) < 2e-6
    ops = [{"name": "ZZPHASE", "wires": [i, i + 1], "parameters": [0.4]} for i in range(n - 1)]
    eng2 = CpuEngine(n, psi)
    eng2.run(ops)
    assert eng2.passes == 1  # uniform run fused into one diagonal (sim.py:271-300)
    assert np.max(np.abs(eng2.state - O.run_ops(n, ops, psi))) < 2e-6
