"""HIP kernels vs the CPU oracle (oracle/sv_oracle.py) through the C ABI.

Tolerances: complex64 amplitudes within 5e-6 max-abs of the complex128
oracle (the reference's kernel A/B budget, tests/mlxQMetalKernelsTest.py);
complex128 within 1e-12; XOR permutations and sampled indices bit-exact.
"""
import math

import numpy as np
import pytest

from oracle import sv_oracle as O

pytestmark = pytest.mark.gpu

TOL64 = 5e-6
TOL128 = 1e-12


def _state(n, dtype="complex64"):
    from mettleq_amd.native_state import NativeState
    return NativeState(n, dtype)


def _rand_state(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


def _rand_u(rng):
    t, p, l = rng.uniform(0, 2 * math.pi, 3)
    return O.U3(t, p, l)


class _Tuned:
    """Sets tuning keys for the duration of a with-block and restores them."""

    def __init__(self, **kv):
        from mettleq_amd import _native as nat
        self.kv = {getattr(nat, "MSV_TUNE_" + k): v for k, v in kv.items()}
        self.old = {}

    def __enter__(self):
        from mettleq_amd import _native as nat
        for k, v in self.kv.items():
            self.old[k] = nat.get_tuning(k)
            nat.set_tuning(k, v)
        return self

    def __exit__(self, *exc):
        from mettleq_amd import _native as nat
        for k, v in self.old.items():
            nat.set_tuning(k, v)


@pytest.mark.parametrize("n", [1, 2, 4, 5, 6, 9, 13, 14, 16, 17, 20, 22])
@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_1q_layer_matches_oracle(gpu, n, dtype):
    rng = np.random.default_rng(100 + n)
    psi = _rand_state(n, n)
    mats = [_rand_u(rng) for _ in range(n)]
    want = O.apply_1q_layer(psi, n, mats)
    with _state(n, dtype) as s:
        s.from_numpy(psi)
        s.apply_1q_layer(mats)
        got = s.to_numpy()
    tol = TOL64 if dtype == "complex64" else TOL128
    assert np.max(np.abs(got - want)) < tol


@pytest.mark.parametrize("lane_swap", [0, 1])
@pytest.mark.parametrize("n", [10, 13, 16, 21])
@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_1q_layer_lane_swaps(gpu, lane_swap, n, dtype):
    """Rotation sweeps whose segments reach lane bits 4 and 5 through
    in-wavefront swaps (MSV_TUNE_LANE_SWAP=1) against the oracle, and
    against the LDS-only segments: a plain 1q layer, a subset of its bits,
    and a layer read through a deferred RZZ ring + CNOT chain (fused first
    sweep). The swaps change where segments meet, never the sweep count."""
    from mettleq_amd import _native as nat
    rng = np.random.default_rng(900 + n)
    psi = _rand_state(n, n + 7)
    mats = [_rand_u(rng) for _ in range(n)]
    subset = sorted(int(q) for q in rng.choice(n, max(1, (2 * n) // 3), replace=False))
    thetas = rng.uniform(0, math.pi, n)
    bonds = [(i, (i + 1) % n) for i in range(n)]
    rows, c = _chain(n)
    want = O.apply_1q_layer(psi, n, mats)
    sub = [mats[q] if q in subset else np.eye(2) for q in range(n)]
    want_sub = O.apply_1q_layer(psi, n, sub)
    fused = O.xor_affine_gather(O.zz_layer_diag(n, bonds, thetas) * psi, n, rows, c)
    fused = O.apply_1q_layer(fused, n, mats)
    tol = TOL64 if dtype == "complex64" else TOL128
    with _Tuned(LANE_SWAP=lane_swap, DEFER_TAIL=0), _state(n, dtype) as s:
        s.from_numpy(psi)
        s0 = s.sweep_count()
        s.apply_1q_layer(mats)
        sweeps = s.sweep_count() - s0
        assert np.max(np.abs(s.to_numpy() - want)) < tol
        s.from_numpy(psi)
        s.apply_1q_layer(mats, subset)
        assert np.max(np.abs(s.to_numpy() - want_sub)) < tol
        s.from_numpy(psi)
        s.apply_phase_layer([nat.MSV_PHASE_ZZ] * n, bonds, list(thetas))
        s.apply_xor_affine(rows, c)
        s.apply_1q_layer(mats)
        assert np.max(np.abs(s.to_numpy() - fused)) < tol
    with _Tuned(LANE_SWAP=1 - lane_swap, DEFER_TAIL=0), _state(n, dtype) as s:
        s.from_numpy(psi)
        s0 = s.sweep_count()
        s.apply_1q_layer(mats)
        assert s.sweep_count() - s0 == sweeps


@pytest.mark.parametrize("reg_bits", [4, 5])
@pytest.mark.parametrize("n", [13, 15, 19])
def test_tile_register_bits_both_shapes(gpu, reg_bits, n):
    """complex64 tile sweeps with 16 (512-thread tiles) or 32 (256-thread
    tiles) amplitudes per thread: 1q layer, QFT and a generic gate list."""
    from mettleq_amd import _native as nat
    rng = np.random.default_rng(500 + n)
    psi = _rand_state(n, n + 1)
    mats = [_rand_u(rng) for _ in range(n)]
    ops = []
    for _ in range(60):
        q = [int(x) for x in rng.choice(n, 2, replace=False)]
        ops.append({"name": "CRY", "wires": q, "parameters": [float(rng.uniform(-3, 3))]})
        ops.append({"name": "U3", "wires": [q[0]], "parameters": [float(x) for x in rng.uniform(0, 6, 3)]})
    from mettleq_amd.fusion import gates_for_op
    gates = [g for op in ops for g in gates_for_op(op, n)]
    old = nat.get_tuning(nat.MSV_TUNE_TILE_REG_BITS)
    nat.set_tuning(nat.MSV_TUNE_TILE_REG_BITS, reg_bits)
    try:
        with _state(n) as s:
            s.from_numpy(psi)
            s.apply_1q_layer(mats)
            assert np.max(np.abs(s.to_numpy() - O.apply_1q_layer(psi, n, mats))) < TOL64
            s.from_numpy(psi)
            s.apply_qft(0, n)
            assert np.max(np.abs(s.to_numpy() - O.qft_fft(psi, n))) < TOL64
            s.from_numpy(psi)
            s.apply_gates(gates)
            assert np.max(np.abs(s.to_numpy() - O.run_ops(n, ops, psi))) < TOL64
    finally:
        nat.set_tuning(nat.MSV_TUNE_TILE_REG_BITS, old)


@pytest.mark.parametrize("kind", ["special", "nonunitary"])
def test_1q_layer_special_and_nonunitary_matrices(gpu, kind):
    """Diagonal, anti-diagonal and real rotations (edge cases of the
    rotation split), and non-unitary matrices (which must not be split)."""
    n = 15
    rng = np.random.default_rng(77)
    psi = _rand_state(n, 9)
    if kind == "special":
        pool = [O.X(), O.Y(), O.Z(), O.S(), O.T(), O.H(), O.RY(0.3), O.RZ(1.1), O.I(),
                O.RX(math.pi), O.PhaseShift(2.0)]
    else:
        pool = [np.array([[1, 0], [0, 0]], complex), 0.5 * O.H(),
                np.array([[1, 2j], [0.5, -1]], complex), _rand_u(rng)]
    mats = [pool[i % len(pool)] for i in range(n)]
    want = O.apply_1q_layer(psi, n, mats)
    with _state(n) as s:
        s.from_numpy(psi)
        s.apply_1q_layer(mats)
        got = s.to_numpy()
    assert np.max(np.abs(got - want)) < TOL64 * max(1.0, np.max(np.abs(want)) * 4)


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_1q_layer_page_balanced_split(gpu, dtype):
    """High sweeps with non-contiguous gate-bit sets (split policy 1)."""
    from mettleq_amd import _native as nat
    n = 24
    rng = np.random.default_rng(41)
    psi = _rand_state(n, 5)
    mats = [_rand_u(rng) for _ in range(n)]
    want = O.apply_1q_layer(psi, n, mats)
    nat.set_tuning(nat.MSV_TUNE_SWEEP_SPLIT, 1)
    try:
        with _state(n, dtype) as s:
            s.from_numpy(psi)
            s.apply_1q_layer(mats)
            got = s.to_numpy()
    finally:
        nat.set_tuning(nat.MSV_TUNE_SWEEP_SPLIT, 0)
    tol = TOL64 if dtype == "complex64" else TOL128
    assert np.max(np.abs(got - want)) < tol


@pytest.mark.parametrize("n", [6, 13, 17, 21])
def test_1q_layer_active_subset(gpu, n):
    rng = np.random.default_rng(7)
    psi = _rand_state(n, 3)
    active = sorted(rng.choice(n, size=max(1, n // 3), replace=False).tolist())
    mats = [O.I() for _ in range(n)]
    for q in active:
        mats[q] = _rand_u(rng)
    want = O.apply_1q_layer(psi, n, mats)
    with _state(n) as s:
        s.from_numpy(psi)
        # garbage matrices on inactive qubits must be ignored
        sent = [m if q in active else np.full((2, 2), 7.0) for q, m in enumerate(mats)]
        s.apply_1q_layer(sent, active=active)
        got = s.to_numpy()
    assert np.max(np.abs(got - want)) < TOL64


@pytest.mark.parametrize("n", [2, 3, 5, 8, 13, 14, 16, 18, 20])
@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_full_qft_matches_closed_form(gpu, n, dtype):
    psi = _rand_state(n, 11 + n)
    want = O.qft_matrix(n) @ psi if n <= 10 else O.qft_fft(psi, n)
    with _state(n, dtype) as s:
        s.from_numpy(psi)
        s.apply_qft(0, n)
        got = s.to_numpy()
        s.apply_qft(0, n, inverse=True)
        back = s.to_numpy()
    tol = 2e-5 if dtype == "complex64" else TOL128
    assert np.max(np.abs(got - want)) < tol
    assert np.max(np.abs(back - psi)) < tol


def test_qft_uniform_to_zero(gpu):
    """QFT(uniform) = |0> (tests/mlxQCoreTest.py:469-477, convention: no swaps
    maps the uniform state to |0>)."""
    n = 15
    psi = np.full(1 << n, 1 / math.sqrt(1 << n), dtype=np.complex128)
    with _state(n) as s:
        s.from_numpy(psi)
        s.apply_qft(0, n)
        got = s.to_numpy()
    want = np.zeros(1 << n)
    want[0] = 1
    assert np.max(np.abs(got - want)) < 1e-5


@pytest.mark.parametrize("n,first,m", [(6, 1, 5), (10, 2, 9), (16, 0, 7), (16, 5, 16), (18, 3, 17)])
def test_subregister_qft_matches_gate_decomposition(gpu, n, first, m):
    psi = _rand_state(n, n + first)
    wires = list(range(first, m))
    want = O.qft(psi, n, wires)
    with _state(n) as s:
        s.from_numpy(psi)
        s.apply_qft(first, m)
        got = s.to_numpy()
        s.apply_qft(first, m, inverse=True)
        back = s.to_numpy()
    assert np.max(np.abs(got - want)) < 2e-5
    assert np.max(np.abs(back - psi)) < 2e-5


@pytest.mark.parametrize("n,j,m,inv", [(5, 0, 5, False), (9, 2, 7, False), (9, 2, 7, True),
                                       (15, 4, 15, True), (15, 14, 15, False)])
def test_single_qft_stage(gpu, n, j, m, inv):
    """qft_stage_sub semantics (shaders/qft.py:144-181): H(j) then the ladder
    CPHASE(k -> j, pi/2^(k-j)); inverse = adjoint."""
    psi = _rand_state(n, 5)
    want = psi
    if not inv:
        want = O.apply_dense(want, n, O.H(), [j])
        for k in range(j + 1, m):
            want = O.apply_dense(want, n, O.CPHASE(math.pi / 2 ** (k - j)), [k, j])
    else:
        for k in reversed(range(j + 1, m)):
            want = O.apply_dense(want, n, O.CPHASE(-math.pi / 2 ** (k - j)), [k, j])
        want = O.apply_dense(want, n, O.H(), [j])
    with _state(n) as s:
        s.from_numpy(psi)
        s.apply_qft_stage(j, m, inverse=inv)
        got = s.to_numpy()
    assert np.max(np.abs(got - want)) < TOL64


@pytest.mark.parametrize("n", [2, 5, 12, 17, 20])
def test_zz_ring_distinct_angles(gpu, n):
    rng = np.random.default_rng(n)
    bonds = [(i, i + 1) for i in range(n - 1)] + ([(n - 1, 0)] if n > 2 else [])
    thetas = rng.uniform(0, math.pi, len(bonds)).tolist()
    psi = _rand_state(n, 9)
    want = psi * O.zz_layer_diag(n, bonds, thetas)
    with _state(n) as s:
        s.from_numpy(psi)
        s.apply_phase_layer([0] * len(bonds), bonds, thetas)
        got = s.to_numpy()
    assert np.max(np.abs(got - want)) < TOL64


@pytest.mark.parametrize("n", [3, 8, 14, 19])
def test_cphase_random_bonds_and_cz(gpu, n):
    rng = np.random.default_rng(n + 1)
    bonds = []
    while len(bonds) < 3 * n:
        a, b = rng.choice(n, 2, replace=False)
        bonds.append((int(a), int(b)))
    thetas = [math.pi if k % 3 == 0 else float(rng.uniform(-3, 3)) for k in range(len(bonds))]
    psi = _rand_state(n, 2)
    want = psi * O.cphase_layer_diag(n, bonds, thetas)
    with _state(n) as s:
        s.from_numpy(psi)
        s.apply_phase_layer([1] * len(bonds), bonds, thetas)
        got = s.to_numpy()
    assert np.max(np.abs(got - want)) < TOL64


def test_cz_brick_exact_signs(gpu):
    """CZ layers use exact +-1 entries (diag.py:89-98): on a real state the
    result equals the oracle's sign flips exactly."""
    n = 12
    rng = np.random.default_rng(3)
    psi = rng.normal(size=1 << n).astype(np.complex64)
    bonds = [(q, q + 1) for q in range(0, n - 1, 2)]
    want = (psi.astype(np.complex128) * O.cphase_layer_diag(n, bonds, [math.pi] * len(bonds))).astype(np.complex64)
    with _state(n) as s:
        s.from_numpy(psi)
        s.apply_phase_layer([1] * len(bonds), bonds, [math.pi] * len(bonds))
        got = s.to_numpy()
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n", [2, 6, 11, 16, 21])
def test_xor_cnot_chain_bit_exact(gpu, n):
    ops = [{"name": "CNOT", "wires": [q, q + 1]} for q in range(n - 1)]
    rows, c = O.compose_inverse_affine(ops, n)
    rng = np.random.default_rng(n)
    psi = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)).astype(np.complex64)
    want = psi[O.affine_source_indices(n, rows, c).astype(np.int64)]
    # Gray-code closed form of the chain's inverse: src = y ^ (y >> 1)
    y = np.arange(1 << n)
    assert np.array_equal(want, psi[y ^ (y >> 1)])
    with _state(n) as s:
        s.from_numpy(psi)
        s.apply_xor_affine(rows, c)
        got = s.to_numpy()
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n", [3, 9, 17])
def test_xor_random_affine_bit_exact(gpu, n):
    rng = np.random.default_rng(42 + n)
    ops = []
    for _ in range(4 * n):
        kind = rng.choice(["X", "CNOT", "SWAP"])
        if kind == "X":
            ops.append({"name": "X", "wires": [int(rng.integers(n))]})
        else:
            a, b = rng.choice(n, 2, replace=False)
            ops.append({"name": str(kind), "wires": [int(a), int(b)]})
    rows, c = O.compose_inverse_affine(ops, n)
    psi = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)).astype(np.complex64)
    want_dense = O.run_ops(n, ops, psi.astype(np.complex128))
    with _state(n) as s:
        s.from_numpy(psi)
        s.apply_xor_affine(rows, c)
        got = s.to_numpy()
    assert np.array_equal(got, psi[O.affine_source_indices(n, rows, c).astype(np.int64)])
    assert np.max(np.abs(got - want_dense)) < 1e-6


def _random_gate_list(n, count, rng):
    from mettleq_amd.native_state import gate_u, gate_diag1, gate_diag2, gate_swap
    gates, ops = [], []
    for _ in range(count):
        r = rng.random()
        if r < 0.4:
            q = int(rng.integers(n))
            t, p, l = rng.uniform(0, 2 * math.pi, 3)
            gates.append(gate_u(q, O.U3(t, p, l)))
            ops.append({"name": "U3", "wires": [q], "parameters": [t, p, l]})
        elif r < 0.6 and n >= 2:
            a, b = (int(x) for x in rng.choice(n, 2, replace=False))
            gates.append(gate_u(b, O.X(), [a]))
            ops.append({"name": "CNOT", "wires": [a, b]})
        elif r < 0.7 and n >= 2:
            a, b = (int(x) for x in rng.choice(n, 2, replace=False))
            th = float(rng.uniform(0, math.pi))
            d = np.diag(O.pauli_pair_phase("ZZPHASE", th))
            gates.append(gate_diag2(a, b, d))
            ops.append({"name": "ZZPHASE", "wires": [a, b], "parameters": [th]})
        elif r < 0.8:
            q = int(rng.integers(n))
            th = float(rng.uniform(-3, 3))
            d = np.diag(O.RZ(th))
            gates.append(gate_diag1(q, d[0], d[1]))
            ops.append({"name": "RZ", "wires": [q], "parameters": [th]})
        elif r < 0.9 and n >= 3:
            a, b, c = (int(x) for x in rng.choice(n, 3, replace=False))
            gates.append(gate_u(c, O.X(), [a, b]))
            ops.append({"name": "TOFFOLI", "wires": [a, b, c]})
        elif n >= 2:
            a, b = (int(x) for x in rng.choice(n, 2, replace=False))
            gates.append(gate_swap(a, b))
            ops.append({"name": "SWAP", "wires": [a, b]})
    return gates, ops


@pytest.mark.parametrize("n,count", [(3, 30), (7, 80), (12, 120), (15, 150), (19, 100)])
def test_generic_gate_list_matches_dense_oracle(gpu, n, count):
    rng = np.random.default_rng(1000 + n)
    gates, ops = _random_gate_list(n, count, rng)
    want = O.run_ops(n, ops)
    with _state(n) as s:
        s.apply_gates(gates)
        got = s.to_numpy()
    assert np.max(np.abs(got - want)) < 2e-5


def test_generic_gates_complex128(gpu):
    n = 11
    rng = np.random.default_rng(77)
    gates, ops = _random_gate_list(n, 120, rng)
    want = O.run_ops(n, ops)
    with _state(n, "complex128") as s:
        s.apply_gates(gates)
        got = s.to_numpy()
    assert np.max(np.abs(got - want)) < TOL128


@pytest.mark.parametrize("n,wires", [(5, None), (8, [0, 3, 7]), (12, [11, 0, 5, 6]),
                                     (17, list(range(10))), (15, [14]), (14, [3, 2, 1, 0])])
def test_probabilities_and_marginals(gpu, n, wires):
    psi = _rand_state(n, 21)
    with _state(n) as s:
        s.from_numpy(psi)
        got = s.probabilities(wires)
        amps = s.to_numpy()
    want = O.probabilities(amps.astype(np.complex128), n, wires)
    assert np.max(np.abs(got - want)) < 1e-12


@pytest.mark.parametrize("n", [1, 4, 12, 13, 18])
def test_sampling_bit_exact_vs_cpu_restatement(gpu, n):
    psi = _rand_state(n, 31 + n)
    with _state(n) as s:
        s.from_numpy(psi)
        amps = s.to_numpy()
        got = s.sample(700, None, seed=20261204, stream_id=5)
    p = O.amp_probs_c64(amps)
    want = O.sample_indices(p, 700, 20261204, 5)
    assert np.array_equal(got, want)


def test_sampling_never_hits_zero_probability(gpu):
    """Bell state: only 00 and 11 (tests/mlxQCoreTest.py:1244-1250)."""
    n = 2
    psi = np.array([1, 0, 0, 1], dtype=np.complex128) / math.sqrt(2)
    with _state(n) as s:
        s.from_numpy(psi)
        got = s.sample(4000, [0, 1], seed=1)
    assert set(np.unique(got).tolist()) <= {0, 3}
    frac = np.mean(got == 3)
    assert abs(frac - 0.5) < 0.05


def test_sampling_wire_subset_packing(gpu):
    n = 6
    psi = _rand_state(n, 8)
    wires = [4, 0, 2]
    with _state(n) as s:
        s.from_numpy(psi)
        amps = s.to_numpy()
        full = s.sample(300, None, seed=9)
        sub = s.sample(300, wires, seed=9)
    assert np.array_equal(sub, O.pack_wires(full, n, wires))
    want = O.sample_indices(O.amp_probs_c64(amps), 300, 9, 0)
    assert np.array_equal(full, want)


@pytest.mark.parametrize("n", [2, 7, 15])
def test_pauli_expectations(gpu, n):
    rng = np.random.default_rng(n)
    psi = _rand_state(n, 44)
    words = []
    for _ in range(20):
        w = {}
        for q in rng.choice(n, size=min(n, 3), replace=False):
            w[int(q)] = str(rng.choice(["X", "Y", "Z"]))
        words.append(w)
    words.append({})
    with _state(n) as s:
        s.from_numpy(psi)
        amps = s.to_numpy().astype(np.complex128)
        got = s.expval_pauli(words)
    want = np.array([O.pauli_expectation(amps, n, w) for w in words])
    assert np.max(np.abs(got - want)) < 1e-10


def test_bell_pauli_algebra(gpu):
    """<XX> = <ZZ> = 1, <YY> = -1 on the Bell state (tests/mlxQCoreTest.py:817-841)."""
    psi = np.array([1, 0, 0, 1], dtype=np.complex128) / math.sqrt(2)
    with _state(2) as s:
        s.from_numpy(psi)
        v = s.expval_pauli([{0: "X", 1: "X"}, {0: "Z", 1: "Z"}, {0: "Y", 1: "Y"}])
    assert np.allclose(v, [1, 1, -1], atol=1e-7)


def test_inner_norm_copy_scale(gpu):
    n = 16
    psi = _rand_state(n, 1)
    with _state(n) as a, _state(n, "complex128") as b:
        a.from_numpy(psi)
        b.copy_from(a)
        assert abs(a.norm2() - 1.0) < 1e-6
        ip = a.inner(b)
        assert abs(ip - 1.0) < 1e-6
        a.scale(1j)
        assert abs(a.inner(b) - (-1j)) < 1e-6
        assert np.max(np.abs(b.to_numpy() - psi)) < 1e-7


def test_reset_and_basis(gpu):
    with _state(10) as s:
        s.set_basis(777)
        v = s.to_numpy()
        assert v[777] == 1 and np.count_nonzero(v) == 1
        s.reset()
        v = s.to_numpy()
        assert v[0] == 1 and np.count_nonzero(v) == 1


def test_invalid_arguments_raise_value_error(gpu):
    with _state(4) as s:
        with pytest.raises(ValueError):
            s.apply_phase_layer([0], [(1, 1)], [0.1])
        with pytest.raises(ValueError):
            s.apply_qft(3, 2)
        with pytest.raises(ValueError):
            s.probabilities([0, 0])
        with pytest.raises(ValueError):
            s.apply_xor_affine([1, 1, 1, 1], 0)  # singular over GF(2)


def test_tile_prefetch_bit_identical_at_28q(gpu):
    """The prefetching tile sweep (next tile's loads overlap the current
    tile; used from 2^15 tiles up) only reorders HBM traffic: a 28-qubit
    U3 layer + QFT + U3 layer + generic gate list must give bit-identical
    amplitudes with prefetch on and off, with the norm kept."""
    from mettleq_amd import _native as nat
    n = 28
    rng = np.random.default_rng(2828)
    m1 = [_rand_u(rng) for _ in range(n)]
    m2 = [_rand_u(rng) for _ in range(n)]
    from mettleq_amd.native_state import gate_u
    gates = [gate_u(int(rng.integers(n)), _rand_u(rng)) for _ in range(20)]
    old_pf = nat.get_tuning(nat.MSV_TUNE_TILE_PREFETCH)
    old_rb = nat.get_tuning(nat.MSV_TUNE_TILE_REG_BITS)
    outs = []
    try:
        nat.set_tuning(nat.MSV_TUNE_TILE_REG_BITS, 4)
        for pf in (0, 1):
            nat.set_tuning(nat.MSV_TUNE_TILE_PREFETCH, pf)
            with _state(n) as s:
                s.apply_1q_layer(m1)
                s.apply_qft(0, n)
                s.apply_1q_layer(m2)
                s.apply_gates(gates)
                outs.append(s.to_numpy())
                assert abs(s.norm2() - 1.0) < 1e-5
    finally:
        nat.set_tuning(nat.MSV_TUNE_TILE_PREFETCH, old_pf)
        nat.set_tuning(nat.MSV_TUNE_TILE_REG_BITS, old_rb)
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))


@pytest.mark.parametrize("split_min", [0, 1, 4, 13])
@pytest.mark.parametrize("n", [9, 15, 21])
def test_1q_layer_rotation_split_thresholds(gpu, split_min, n):
    """1q layers with the Z-Y-Z split (two table diagonals around fast
    Givens rotations, both forms |tan| <= 1 and |cot| < 1) from `split_min`
    gates per sweep up, against the dense oracle; includes angles near 0,
    pi/2 and pi where the form switches."""
    from mettleq_amd import _native as nat
    rng = np.random.default_rng(900 + n + split_min)
    psi = _rand_state(n, 3 * n)
    special = [0.0, 1e-7, math.pi / 2 - 1e-7, math.pi / 2, math.pi / 2 + 1e-7, math.pi - 1e-7, math.pi]
    mats = []
    for q in range(n):
        t = special[q % len(special)] if q % 2 == 0 else rng.uniform(0, 2 * math.pi)
        p, l = rng.uniform(0, 2 * math.pi, 2)
        mats.append(O.U3(t, p, l))
    want = O.apply_1q_layer(psi, n, mats)
    old = nat.get_tuning(nat.MSV_TUNE_ROT_SPLIT_MIN)
    nat.set_tuning(nat.MSV_TUNE_ROT_SPLIT_MIN, split_min)
    try:
        with _state(n) as s:
            s.from_numpy(psi)
            s.apply_1q_layer(mats)
            got = s.to_numpy()
    finally:
        nat.set_tuning(nat.MSV_TUNE_ROT_SPLIT_MIN, old)
    assert np.max(np.abs(got - want)) < TOL64


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
@pytest.mark.parametrize("after", ["xor", "probs", "download", "phase", "layer", "qft", "inner",
                                   "sample", "reset", "scale", "gates"])
def test_deferred_phase_layer_then_next_operation(gpu, dtype, after):
    """A phase layer is held until the state's next operation: the XOR
    permutation applies it in its own gather (D at the source index), any
    other call launches it first. Each sequence must match the oracle."""
    from mettleq_amd import _native as nat
    from mettleq_amd.native_state import gate_u
    n = 14
    rng = np.random.default_rng(4242)
    psi = _rand_state(n, 99)
    bonds = [(i, (i + 1) % n) for i in range(n)] + [(0, 7), (3, 11)]
    th = rng.uniform(0, math.pi, len(bonds))
    want = O.zz_layer_diag(n, bonds, th) * psi
    rows, c = O.compose_inverse_affine([{"name": "CNOT", "wires": [i, i + 1]} for i in range(n - 1)]
                                       + [{"name": "X", "wires": [2]}], n)
    mats = [_rand_u(rng) for _ in range(n)]
    tol = TOL64 if dtype == "complex64" else TOL128
    assert nat.get_tuning(nat.MSV_TUNE_DEFER_DIAG) == 1
    with _state(n, dtype) as s:
        s.from_numpy(psi)
        s.apply_phase_layer([nat.MSV_PHASE_ZZ] * len(bonds), bonds, list(th))
        if after == "xor":
            s.apply_xor_affine(rows, c)
            want = O.xor_affine_gather(want, n, rows, c)
        elif after == "probs":
            assert np.max(np.abs(s.probabilities() - np.abs(want) ** 2)) < tol
        elif after == "download":
            pass
        elif after == "phase":
            s.apply_phase_layer([nat.MSV_PHASE_CPHASE] * 3, [(0, 1), (4, 9), (12, 13)], [0.3, 1.1, math.pi])
            want = O.cphase_layer_diag(n, [(0, 1), (4, 9), (12, 13)], [0.3, 1.1, math.pi]) * want
            s.apply_xor_affine(rows, c)
            want = O.xor_affine_gather(want, n, rows, c)
        elif after == "layer":
            s.apply_1q_layer(mats)
            want = O.apply_1q_layer(want, n, mats)
        elif after == "qft":
            s.apply_qft(0, n)
            want = O.qft_fft(want, n)
        elif after == "inner":
            with _state(n, dtype) as t:
                t.from_numpy(psi)
                ip = t.inner(s)
            assert abs(ip - np.vdot(psi, want)) < 10 * tol
        elif after == "sample":
            got = s.sample(300, None, seed=11)
            ref = O.sample_indices(O.amp_probs_c64(s.to_numpy()), 300, 11, 0)
            assert np.array_equal(got, ref)
        elif after == "reset":
            s.reset()
            want = O.zero_state(n)
        elif after == "scale":
            s.scale(1j)
            want = 1j * want
        elif after == "gates":
            g = [gate_u(3, mats[0]), gate_u(9, mats[1], [2])]
            s.apply_gates(g)
            want = O.apply_controlled(O.apply_dense(want, n, mats[0], [3]), n, mats[1], [2], 9)
        got = s.to_numpy()
    assert np.max(np.abs(got - want)) < tol


def test_deferred_phase_matches_immediate_at_26q(gpu):
    """RZZ ring then CNOT chain at 26 qubits: the gather that absorbs the
    deferred phase layer equals the two separate sweeps (within complex64
    rounding of the phase products), in one sweep instead of two."""
    from mettleq_amd import _native as nat
    from mettleq_amd.workloads import ring_bonds
    n = 26
    rng = np.random.default_rng(26)
    mats = [_rand_u(rng) for _ in range(n)]
    bonds = ring_bonds(n)
    th = [float(x) for x in rng.uniform(0, math.pi, len(bonds))]
    rows, c = O.compose_inverse_affine([{"name": "CNOT", "wires": [i, i + 1]} for i in range(n - 1)], n)
    outs, sweeps = [], []
    try:
        for defer in (0, 1):
            nat.set_tuning(nat.MSV_TUNE_DEFER_DIAG, defer)
            with _Tuned(DEFER_TAIL=0), _state(n) as s:
                s.apply_1q_layer(mats)
                s0 = s.sweep_count()
                s.apply_phase_layer([nat.MSV_PHASE_ZZ] * len(bonds), bonds, th)
                s.apply_xor_affine(rows, c)
                outs.append(s.to_numpy())  # runs the deferred work
                sweeps.append(s.sweep_count() - s0)
    finally:
        nat.set_tuning(nat.MSV_TUNE_DEFER_DIAG, 1)
    assert sweeps == [2, 1]
    assert np.max(np.abs(outs[0] - outs[1])) < 1e-6


def _chain(n):
    return O.compose_inverse_affine([{"name": "CNOT", "wires": [i, i + 1]} for i in range(n - 1)], n)


def _random_affine(n, rng):
    ops = []
    for _ in range(3 * n):
        a, b = rng.choice(n, 2, replace=False)
        ops.append({"name": "CNOT", "wires": [int(a), int(b)]})
    ops.append({"name": "X", "wires": [int(rng.integers(n))]})
    return O.compose_inverse_affine(ops, n)


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
@pytest.mark.parametrize("n", [9, 13, 14, 20])
@pytest.mark.parametrize("with_phase", [False, True])
@pytest.mark.parametrize("perm", ["chain", "random"])
def test_deferred_permutation_read_by_next_1q_layer(gpu, dtype, n, with_phase, perm):
    """An XOR permutation (with or without a deferred phase layer before it)
    is held until the next operation; a 1q layer reads its first sweep
    through it when the map keeps the lowest column bits among themselves
    (the CNOT chain does), otherwise the gather runs first (random maps)."""
    from mettleq_amd import _native as nat
    rng = np.random.default_rng(7000 + n)
    psi = _rand_state(n, 5 * n)
    rows, c = _chain(n) if perm == "chain" else _random_affine(n, rng)
    mats = [_rand_u(rng) for _ in range(n)]
    bonds = [(i, (i + 1) % n) for i in range(n)]
    th = rng.uniform(0, math.pi, n)
    want = psi
    with _state(n, dtype) as s:
        s.from_numpy(psi)
        if with_phase:
            s.apply_phase_layer([nat.MSV_PHASE_ZZ] * n, bonds, list(th))
            want = O.zz_layer_diag(n, bonds, th) * want
        s.apply_xor_affine(rows, c)
        want = O.xor_affine_gather(want, n, rows, c)
        s.apply_1q_layer(mats)
        want = O.apply_1q_layer(want, n, mats)
        got = s.to_numpy()
    tol = TOL64 if dtype == "complex64" else TOL128
    assert np.max(np.abs(got - want)) < tol


@pytest.mark.parametrize("after", ["probs", "qft", "gates", "xor", "phase", "inner", "sample",
                                   "reset", "layer_subset", "scale"])
def test_deferred_permutation_then_other_operations(gpu, after):
    from mettleq_amd import _native as nat
    from mettleq_amd.native_state import gate_u
    n = 13
    rng = np.random.default_rng(31337)
    psi = _rand_state(n, 3)
    rows, c = _chain(n)
    rows2, c2 = _random_affine(n, rng)
    mats = [_rand_u(rng) for _ in range(n)]
    want = O.xor_affine_gather(psi, n, rows, c)
    with _state(n) as s:
        s.from_numpy(psi)
        s.apply_xor_affine(rows, c)
        if after == "probs":
            assert np.max(np.abs(s.probabilities() - np.abs(want) ** 2)) < TOL64
        elif after == "qft":
            s.apply_qft(0, n)
            want = O.qft_fft(want, n)
        elif after == "gates":
            s.apply_gates([gate_u(4, mats[0]), gate_u(0, mats[1], [12])])
            want = O.apply_controlled(O.apply_dense(want, n, mats[0], [4]), n, mats[1], [12], 0)
        elif after == "xor":
            s.apply_xor_affine(rows2, c2)
            want = O.xor_affine_gather(want, n, rows2, c2)
        elif after == "phase":
            s.apply_phase_layer([nat.MSV_PHASE_CPHASE], [(2, 9)], [0.7])
            want = O.cphase_layer_diag(n, [(2, 9)], [0.7]) * want
        elif after == "inner":
            with _state(n) as t:
                t.from_numpy(psi)
                assert abs(t.inner(s) - np.vdot(psi, want)) < 1e-5
        elif after == "sample":
            got = s.sample(200, None, seed=3)
            assert np.array_equal(got, O.sample_indices(O.amp_probs_c64(s.to_numpy()), 200, 3, 0))
        elif after == "reset":
            s.reset()
            want = O.zero_state(n)
        elif after == "layer_subset":
            act = [0, 5, 12]
            s.apply_1q_layer(mats, active=act)
            full = [mats[q] if q in act else np.eye(2) for q in range(n)]
            want = O.apply_1q_layer(want, n, full)
        elif after == "scale":
            s.scale(-1j)
            want = -1j * want
        got = s.to_numpy()
    assert np.max(np.abs(got - want)) < TOL64


def test_deferred_permutation_fused_sweep_counts_at_26q(gpu):
    """U3 layer, RZZ ring, CNOT chain, U3 layer at 26 qubits: with deferral
    the ring and the chain ride in the second layer's first sweep (two
    sweeps fewer); amplitudes agree with the undeferred run."""
    from mettleq_amd import _native as nat
    from mettleq_amd.workloads import ring_bonds
    n = 26
    rng = np.random.default_rng(2626)
    m1 = [_rand_u(rng) for _ in range(n)]
    m2 = [_rand_u(rng) for _ in range(n)]
    bonds = ring_bonds(n)
    th = [float(x) for x in rng.uniform(0, math.pi, len(bonds))]
    rows, c = _chain(n)
    outs, sweeps = [], []
    with _Tuned(DEFER_TAIL=0), _state(n) as s:
        s0 = s.sweep_count()
        s.apply_1q_layer(m1)
        u3 = s.sweep_count() - s0
    try:
        for d in (0, 1):
            nat.set_tuning(nat.MSV_TUNE_DEFER_DIAG, d)
            nat.set_tuning(nat.MSV_TUNE_DEFER_PERM, d)
            with _Tuned(DEFER_TAIL=0), _state(n) as s:
                s0 = s.sweep_count()
                s.apply_1q_layer(m1)
                s.apply_phase_layer([nat.MSV_PHASE_ZZ] * len(bonds), bonds, th)
                s.apply_xor_affine(rows, c)
                s.apply_1q_layer(m2)
                outs.append(s.to_numpy())
                sweeps.append(s.sweep_count() - s0)
    finally:
        nat.set_tuning(nat.MSV_TUNE_DEFER_DIAG, 1)
        nat.set_tuning(nat.MSV_TUNE_DEFER_PERM, 1)
    assert sweeps == [u3 + 1 + 1 + u3, 2 * u3]
    assert np.max(np.abs(outs[0] - outs[1])) < 2e-6


def _iqft_fft(psi, n):
    # inverse of O.qft_fft: undo the bit reversal, then the forward FFT / sqrt(N)
    size = 1 << n
    idx = np.array([O.bit_reverse(i, n) for i in range(size)])
    x = np.empty_like(np.asarray(psi, dtype=np.complex128))
    x[idx] = psi
    return np.fft.fft(x) / math.sqrt(size)


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
@pytest.mark.parametrize("n", [9, 14, 20])
@pytest.mark.parametrize("work", ["phase+chain", "chain", "phase", "phase+random"])
@pytest.mark.parametrize("qft", ["full", "inverse", "sub"])
def test_deferred_work_read_by_next_qft(gpu, dtype, n, work, qft):
    """Deferred phase layer and/or XOR permutation followed by a QFT: the
    first QFT sweep reads through them when the map keeps the lowest column
    bits among themselves (the CNOT chain), else they run first (random
    maps). A lone phase layer is read when the second buffer exists (here:
    after an earlier gather). Each sequence must match the oracle."""
    from mettleq_amd import _native as nat
    rng = np.random.default_rng(9100 + n)
    psi = _rand_state(n, 17 * n)
    bonds = [(i, (i + 1) % n) for i in range(n)] + [(0, n // 2)]
    th = rng.uniform(0, math.pi, len(bonds))
    rows, c = _random_affine(n, rng) if work == "phase+random" else _chain(n)
    pre_rows, pre_c = _random_affine(n, rng)
    want = psi
    with _state(n, dtype) as s:
        s.from_numpy(psi)
        if work == "phase":
            s.apply_xor_affine(pre_rows, pre_c)  # its gather runs at the phase layer
            want = O.xor_affine_gather(want, n, pre_rows, pre_c)
        if "phase" in work:
            s.apply_phase_layer([nat.MSV_PHASE_ZZ] * len(bonds), bonds, list(th))
            want = O.zz_layer_diag(n, bonds, th) * want
        if work != "phase":
            s.apply_xor_affine(rows, c)
            want = O.xor_affine_gather(want, n, rows, c)
        if qft == "full":
            s.apply_qft(0, n)
            want = O.qft_fft(want, n)
        elif qft == "inverse":
            s.apply_qft(0, n, inverse=True)
            want = _iqft_fft(want, n)
        else:
            s.apply_qft(1, n - 1)
            want = O.qft(want, n, list(range(1, n - 1)))
        got = s.to_numpy()
    tol = 2e-5 if dtype == "complex64" else TOL128
    assert np.max(np.abs(got - want)) < tol


def test_deferred_work_read_by_qft_sweep_counts_at_26q(gpu):
    """U3 layer, RZZ ring, CNOT chain, full QFT at 26 qubits: with deferral
    the ring and the chain ride in the QFT's first sweep (two sweeps fewer);
    amplitudes agree with the undeferred run."""
    from mettleq_amd import _native as nat
    from mettleq_amd.workloads import ring_bonds
    n = 26
    rng = np.random.default_rng(2727)
    m1 = [_rand_u(rng) for _ in range(n)]
    bonds = ring_bonds(n)
    th = [float(x) for x in rng.uniform(0, math.pi, len(bonds))]
    rows, c = _chain(n)
    with _Tuned(DEFER_TAIL=0), _state(n) as s:
        s0 = s.sweep_count()
        s.apply_1q_layer(m1)
        u3 = s.sweep_count() - s0
        s.apply_qft(0, n)
        q = s.sweep_count() - s0 - u3
    outs, sweeps = [], []
    try:
        for d in (0, 1):
            nat.set_tuning(nat.MSV_TUNE_DEFER_DIAG, d)
            nat.set_tuning(nat.MSV_TUNE_DEFER_PERM, d)
            with _Tuned(DEFER_TAIL=0), _state(n) as s:
                s0 = s.sweep_count()
                s.apply_1q_layer(m1)
                s.apply_phase_layer([nat.MSV_PHASE_ZZ] * len(bonds), bonds, th)
                s.apply_xor_affine(rows, c)
                s.apply_qft(0, n)
                outs.append(s.to_numpy())
                sweeps.append(s.sweep_count() - s0)
    finally:
        nat.set_tuning(nat.MSV_TUNE_DEFER_DIAG, 1)
        nat.set_tuning(nat.MSV_TUNE_DEFER_PERM, 1)
    assert sweeps == [u3 + 1 + 1 + q, u3 + q]
    assert np.max(np.abs(outs[0] - outs[1])) < 2e-6


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_lone_deferred_diagonal_read_in_place_by_next_layer(gpu, dtype):
    """Brick circuit (U3 layer, CZ brick, U3 layer, CZ brick, QFT) at 20
    qubits on a fresh state: each CZ brick rides in the next layer's first
    sweep, in place (no second buffer), one sweep per brick fewer; the
    amplitudes match the undeferred run and the oracle."""
    from mettleq_amd import _native as nat
    n = 20
    rng = np.random.default_rng(2020)
    m1 = [_rand_u(rng) for _ in range(n)]
    m2 = [_rand_u(rng) for _ in range(n)]
    b0 = [(q, q + 1) for q in range(0, n - 1, 2)]
    b1 = [(q, q + 1) for q in range(1, n - 1, 2)]
    psi = _rand_state(n, 77)
    want = O.apply_1q_layer(psi, n, m1)
    want = O.cphase_layer_diag(n, b0, [math.pi] * len(b0)) * want
    want = O.apply_1q_layer(want, n, m2)
    want = O.cphase_layer_diag(n, b1, [math.pi] * len(b1)) * want
    want = O.qft_fft(want, n)
    outs, sweeps = [], []
    try:
        for d in (0, 1):
            nat.set_tuning(nat.MSV_TUNE_DEFER_DIAG, d)
            with _Tuned(DEFER_TAIL=0), _state(n, dtype) as s:
                s.from_numpy(psi)
                s0 = s.sweep_count()
                s.apply_1q_layer(m1)
                u3 = s.sweep_count() - s0
                s.apply_phase_layer([nat.MSV_PHASE_CPHASE] * len(b0), b0, [math.pi] * len(b0))
                s.apply_1q_layer(m2)
                s.apply_phase_layer([nat.MSV_PHASE_CPHASE] * len(b1), b1, [math.pi] * len(b1))
                s.apply_qft(0, n)
                outs.append(s.to_numpy())
                sweeps.append(s.sweep_count() - s0)
    finally:
        nat.set_tuning(nat.MSV_TUNE_DEFER_DIAG, 1)
    assert sweeps[0] == sweeps[1] + 2
    tol = 2e-5 if dtype == "complex64" else TOL128
    assert np.max(np.abs(outs[1] - want)) < tol
    assert np.max(np.abs(outs[0] - outs[1])) < tol


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
@pytest.mark.parametrize("n", [1, 3, 10, 13, 20, 24])
@pytest.mark.parametrize("subset", [False, True])
def test_deferred_basis_state_then_1q_layer_writes_the_product_state(gpu, dtype, n, subset):
    """set_basis(k) is kept pending; the next 1q layer writes the product
    state prod_q U_q|k_q> in one pass (no sweep)."""
    rng = np.random.default_rng(4100 + n + 7 * subset)
    k = int(rng.integers(1 << n))
    mats = [_rand_u(rng) for _ in range(n)]
    act = sorted(int(q) for q in rng.choice(n, max(1, n // 2), replace=False)) if subset else None
    with _state(n, dtype) as s:
        s.set_basis(k)
        s0 = s.sweep_count()
        s.apply_1q_layer(mats, active=act)
        sweeps = s.sweep_count() - s0
        if n <= 20:
            got = s.to_numpy()
        else:
            win = ((1 << n) - (1 << 16), 1 << 16)
            got = s.to_numpy(*win)
    assert sweeps == 0
    # the product state, from per-qubit columns U_q[:, k_q]
    cols = []
    for q in range(n):
        kq = (k >> (n - 1 - q)) & 1
        u = mats[q] if (act is None or q in act) else np.eye(2)
        cols.append(np.asarray(u)[:, kq])
    if n <= 20:
        want = cols[0]
        for c in cols[1:]:
            want = np.kron(want, c)
        if n <= 13:
            e = np.zeros(1 << n, dtype=np.complex128)
            e[k] = 1
            full = [mats[q] if (act is None or q in act) else np.eye(2) for q in range(n)]
            assert np.max(np.abs(O.apply_1q_layer(e, n, full) - want)) < 1e-12
    else:
        idx = np.arange(*[win[0], win[0] + win[1]])
        want = np.ones(len(idx), dtype=np.complex128)
        for q in range(n):
            want *= cols[q][(idx >> (n - 1 - q)) & 1]
    tol = 2e-6 if dtype == "complex64" else TOL128
    assert np.max(np.abs(got - want)) < tol


@pytest.mark.parametrize("after", ["probs", "phase", "xor", "qft", "gates", "inner", "scale", "sample",
                                   "download", "copy", "upload_part"])
def test_deferred_basis_state_then_other_operations(gpu, after):
    """Every operation other than a 1q layer writes the pending basis state
    first: results equal those of |k>."""
    from mettleq_amd import _native as nat
    from mettleq_amd.native_state import gate_u
    n = 12
    rng = np.random.default_rng(4242)
    k = int(rng.integers(1 << n))
    e = np.zeros(1 << n, dtype=np.complex128)
    e[k] = 1
    want = e
    with _state(n) as s:
        s.reset()
        s.set_basis(k)
        if after == "probs":
            assert np.array_equal(s.probabilities(), np.abs(e) ** 2)
        elif after == "phase":
            bonds = [(0, 5), (3, 4)]
            s.apply_phase_layer([nat.MSV_PHASE_ZZ] * 2, bonds, [0.3, 0.9])
            want = O.zz_layer_diag(n, bonds, [0.3, 0.9]) * e
        elif after == "xor":
            rows, c = _chain(n)
            s.apply_xor_affine(rows, c)
            want = O.xor_affine_gather(e, n, rows, c)
        elif after == "qft":
            s.apply_qft(0, n)
            want = O.qft_fft(e, n)
        elif after == "gates":
            m = _rand_u(rng)
            s.apply_gates([gate_u(4, m, [2])])
            want = O.apply_controlled(e, n, m, [2], 4)
        elif after == "inner":
            with _state(n) as t:
                t.from_numpy(e)
                assert abs(t.inner(s) - 1.0) < 1e-6
        elif after == "scale":
            s.scale(1j)
            want = 1j * e
        elif after == "sample":
            got = s.sample(50, None, seed=1)
            assert np.all(got == k)
        elif after == "copy":
            with _state(n) as t:
                t.copy_from(s)
                assert np.array_equal(t.to_numpy(), e.astype(np.complex64))
        elif after == "upload_part":
            s.from_numpy(np.full(4, 0.5, dtype=np.complex64), offset=0)
            want = e.copy()
            want[:4] = 0.5
        got = s.to_numpy()
    assert np.max(np.abs(got - want)) < 2e-6


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
@pytest.mark.parametrize("n", [9, 14, 21])
def test_deferred_basis_state_then_gate_list_is_generated_by_the_first_tile_sweep(gpu, dtype, n):
    """set_basis(k), then a served-style gate list (U3, CZ, ZZPHASE, CNOT):
    the first tile sweep generates |k> instead of reading a written state
    (no basis-state kernel), and the result equals the oracle's run from |k>."""
    from mettleq_amd.circuit import encode_ops
    from mettleq_amd.serving import random_circuit
    rng = np.random.default_rng(9100 + n)
    k = int(rng.integers(1 << n))
    ops = random_circuit(n, 60, rng)
    e = np.zeros(1 << n, dtype=np.complex128)
    e[k] = 1
    want = O.run_ops(n, ops, e)
    with _state(n, dtype) as s:
        s.set_basis(k)
        s.dispatch_log(reset=True)
        s.apply_circuit(*encode_ops(ops))
        kinds = s.dispatch_log()
        got = s.to_numpy()
    assert kinds and set(kinds) == {"k_tile"}, kinds
    tol = 5e-6 if dtype == "complex64" else TOL128
    assert np.max(np.abs(got - want)) < tol


@pytest.mark.parametrize("n", [6, 11, 15])
@pytest.mark.parametrize("reorder", [0, 1])
def test_gate_lists_with_controls_swaps_and_diagonals(gpu, n, reorder):
    """Random gate lists with multi-controlled 2x2 gates (X among them),
    controlled SWAPs and one- and two-qubit diagonals (+-1 ones among
    them), in order and regrouped (MSV_TUNE_REORDER_GATES): the state
    matches the oracle's gate-by-gate application."""
    from mettleq_amd import _native as nat
    from mettleq_amd.native_state import gate_diag1, gate_diag2, gate_swap, gate_u
    rng = np.random.default_rng(5150 + n)
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    gates = []
    want = _rand_state(n, 3 * n)
    psi = want.copy()
    for _ in range(120):
        kind = rng.integers(6)
        qs = [int(q) for q in rng.choice(n, min(n, 4), replace=False)]
        if kind == 0:  # U, 0-2 controls
            nc = int(rng.integers(3))
            u = _rand_u(rng)
            gates.append(gate_u(qs[0], u, qs[1:1 + nc]))
            want = O.apply_controlled(want, n, u, qs[1:1 + nc], qs[0]) if nc else O.apply_dense(want, n, u, [qs[0]])
        elif kind == 1:  # X with 1-2 controls (register swaps)
            nc = int(rng.integers(1, 3))
            gates.append(gate_u(qs[0], X, qs[1:1 + nc]))
            want = O.apply_controlled(want, n, X, qs[1:1 + nc], qs[0])
        elif kind == 2:  # two-qubit diagonal, sometimes +-1 (CZ-like)
            d = [1, 1, 1, -1] if rng.random() < 0.5 else list(np.exp(1j * rng.uniform(0, 6, 4)))
            gates.append(gate_diag2(qs[0], qs[1], d))
            want = O.apply_dense(want, n, np.diag(d), [qs[0], qs[1]])
        elif kind == 3:  # one-qubit diagonal, sometimes Z
            d = (1, -1) if rng.random() < 0.5 else tuple(np.exp(1j * rng.uniform(0, 6, 2)))
            gates.append(gate_diag1(qs[0], *d))
            want = O.apply_dense(want, n, np.diag(d), [qs[0]])
        elif kind == 4:  # SWAP, 0-1 controls
            nc = int(rng.integers(2))
            gates.append(gate_swap(qs[0], qs[1], qs[2:2 + nc]))
            if nc:
                m = np.eye(8, dtype=complex)
                m[-4:, -4:] = O.SWAP()
                want = O.apply_dense(want, n, m, qs[2:3] + qs[:2])
            else:
                want = O.apply_dense(want, n, O.SWAP(), qs[:2])
        else:  # repeated gate on one wire (merged when adjacent)
            u = _rand_u(rng)
            gates.append(gate_u(qs[0], u))
            gates.append(gate_u(qs[0], u.conj().T))
    old = nat.get_tuning(nat.MSV_TUNE_REORDER_GATES)
    try:
        nat.set_tuning(nat.MSV_TUNE_REORDER_GATES, reorder)
        with _state(n) as s:
            s.from_numpy(psi)
            s.apply_gates(gates)
            got = s.to_numpy()
    finally:
        nat.set_tuning(nat.MSV_TUNE_REORDER_GATES, old)
    assert np.max(np.abs(got - want)) < 2e-5


def _layer_cycle(s, want, n, rng, layers, ring, perm):
    """U3 layer, then (RZZ ring) (permutation) before each further U3 layer;
    returns the oracle's state."""
    from mettleq_amd import _native as nat
    bonds = [(i, (i + 1) % n) for i in range(n)]
    for layer in range(layers):
        if layer:
            if ring:
                th = rng.uniform(0, math.pi, n)
                s.apply_phase_layer([nat.MSV_PHASE_ZZ] * n, bonds, list(th))
                want = O.zz_layer_diag(n, bonds, th) * want
            if perm is not None:
                rows, c = perm
                s.apply_xor_affine(rows, c)
                want = O.xor_affine_gather(want, n, rows, c)
        mats = [_rand_u(rng) for _ in range(n)]
        s.apply_1q_layer(mats)
        want = O.apply_1q_layer(want, n, mats)
    return want


def _sparse_affine(n, rng, count=4):
    # a few CNOTs and an X among the wires of index bits >= 4: the map keeps
    # the column bits and moves few bits, so a crossing tile exists that is
    # not the CNOT chain's
    ops = []
    for _ in range(count):
        a, b = rng.choice(n - 4, 2, replace=False)
        ops.append({"name": "CNOT", "wires": [int(a), int(b)]})
    ops.append({"name": "X", "wires": [int(rng.integers(n - 4))]})
    return O.compose_inverse_affine(ops, n)


@pytest.mark.parametrize("n", [16, 17, 20])
@pytest.mark.parametrize("perm", ["chain", "random", "sparse", "none"])
@pytest.mark.parametrize("ring", [False, True])
def test_pending_tail_crosses_into_the_next_layer(gpu, n, perm, ring):
    """With 9-bit tiles a complex64 layer takes three or four sweeps; its last
    sweep's rotations stay pending and the next layer's first sweep runs
    them, the deferred ring and permutation, and its own first rotations,
    changing frame inside the tile (k_rot_relabel). The chain and a sparse
    map on the high bits keep the column bits, so every layer after the
    first crosses (the sparse map's tile has extra basis vectors other than
    the chain's); a random map does not (its gather runs after the pending
    rotations). Four layers against the oracle, with pending tails on and
    off; fewer sweeps when on."""
    rng0 = np.random.default_rng(8100 + n)
    psi = _rand_state(n, 81 * n)
    p = {"chain": _chain(n), "random": _random_affine(n, rng0), "sparse": _sparse_affine(n, rng0),
         "none": None}[perm]
    outs, sweeps = [], []
    # without a permutation: on request (2), and by default (1) for states
    # up to 64 GiB
    for tail in ((0, 1) if perm != "none" else (0, 2, 1)):
        with _Tuned(TILE_LOCAL_BITS=9, DEFER_TAIL=tail):
            rng = np.random.default_rng(8200 + n)
            with _state(n) as s:
                s.from_numpy(psi)
                s0 = s.sweep_count()
                want = _layer_cycle(s, psi, n, rng, 4, ring, p)
                got = s.to_numpy()
                sweeps.append(s.sweep_count() - s0)
        assert np.max(np.abs(got - want)) < TOL64
        outs.append(got)
    for o in outs[1:]:
        assert np.max(np.abs(outs[0] - o)) < 2e-6
    assert sweeps[1] <= sweeps[0]
    if perm in ("chain", "sparse", "none"):
        assert sweeps[1] < sweeps[0]
    if perm == "none":
        assert sweeps[2] == sweeps[1]


@pytest.mark.parametrize("perm", ["chain", "none"])
def test_pending_tail_crosses_complex128(gpu, perm):
    """complex128 (12-bit tiles, three sweeps per layer at 22 qubits)."""
    n = 22
    rng = np.random.default_rng(8300)
    psi = _rand_state(n, 83)
    p = _chain(n) if perm == "chain" else None
    with _Tuned(DEFER_TAIL=1 if perm == "chain" else 2), _state(n, "complex128") as s:
        s.from_numpy(psi)
        s0 = s.sweep_count()
        want = _layer_cycle(s, psi, n, rng, 3, True, p)
        got = s.to_numpy()
        sw = s.sweep_count() - s0
    assert np.max(np.abs(got - want)) < TOL128
    assert sw == 2 + 2 + 2 + 1  # the last tail runs when the state is read


@pytest.mark.parametrize("after", ["probs", "qft", "qft_after_ring_chain", "gates", "xor", "phase",
                                   "diagonal", "inner", "sample", "reset", "layer_subset", "scale",
                                   "copy", "upload_part", "second_layer"])
def test_pending_tail_then_other_operations(gpu, after):
    """A 1q layer leaves rotations pending; every other operation runs them
    first (a QFT reads the deferred ring and chain after them), a reset drops
    them, and a second layer runs them in its first sweep."""
    from mettleq_amd import _native as nat
    from mettleq_amd.native_state import gate_u
    n = 16
    rng = np.random.default_rng(8400)
    psi = _rand_state(n, 84)
    mats = [_rand_u(rng) for _ in range(n)]
    rows, c = _chain(n)
    rows2, c2 = _random_affine(n, rng)
    want = O.apply_1q_layer(psi, n, mats)
    # 9-bit tiles (three sweeps per layer) except before a QFT, whose sweep
    # planner takes the default tiles
    bits = 13 if after.startswith("qft") else 9
    with _Tuned(TILE_LOCAL_BITS=bits, DEFER_TAIL=1), _state(n) as s:
        s.from_numpy(psi)
        s.apply_1q_layer(mats)
        if after == "probs":
            assert np.max(np.abs(s.probabilities() - np.abs(want) ** 2)) < TOL64
        elif after == "qft":
            s.apply_qft(0, n)
            want = O.qft_fft(want, n)
        elif after == "qft_after_ring_chain":
            bonds = [(i, (i + 1) % n) for i in range(n)]
            th = rng.uniform(0, math.pi, n)
            s.apply_phase_layer([nat.MSV_PHASE_ZZ] * n, bonds, list(th))
            s.apply_xor_affine(rows, c)
            s.apply_qft(0, n)
            want = O.qft_fft(O.xor_affine_gather(O.zz_layer_diag(n, bonds, th) * want, n, rows, c), n)
        elif after == "gates":
            s.apply_gates([gate_u(4, mats[0]), gate_u(0, mats[1], [12])])
            want = O.apply_controlled(O.apply_dense(want, n, mats[0], [4]), n, mats[1], [12], 0)
        elif after == "xor":
            s.apply_xor_affine(rows2, c2)
            want = O.xor_affine_gather(want, n, rows2, c2)
        elif after == "phase":
            s.apply_phase_layer([nat.MSV_PHASE_CPHASE], [(2, 9)], [0.7])
            want = O.cphase_layer_diag(n, [(2, 9)], [0.7]) * want
        elif after == "diagonal":
            d0, d1 = np.exp(0.3j), np.exp(-1.1j)
            s.apply_diagonal(one=[(3, d0, d1)])
            want = O.apply_dense(want, n, np.diag([d0, d1]), [3])
        elif after == "inner":
            with _state(n) as t:
                t.from_numpy(psi)
                assert abs(t.inner(s) - np.vdot(psi, want)) < 1e-5
        elif after == "sample":
            got = s.sample(200, None, seed=3)
            assert np.array_equal(got, O.sample_indices(O.amp_probs_c64(s.to_numpy()), 200, 3, 0))
        elif after == "reset":
            s.reset()
            want = O.zero_state(n)
        elif after == "layer_subset":
            act = [0, 5, 12]
            s.apply_1q_layer(mats, active=act)
            full = [mats[q] if q in act else np.eye(2) for q in range(n)]
            want = O.apply_1q_layer(want, n, full)
        elif after == "scale":
            s.scale(-1j)
            want = -1j * want
        elif after == "copy":
            with _state(n) as t:
                t.copy_from(s)
                assert np.max(np.abs(t.to_numpy() - want)) < TOL64
        elif after == "upload_part":
            s.from_numpy(np.full(4, 0.5, dtype=np.complex64), offset=0)
            want = want.copy()
            want[:4] = 0.5
        elif after == "second_layer":
            m2 = [_rand_u(rng) for _ in range(n)]
            with _Tuned(DEFER_TAIL=2):  # cross with no permutation pending
                s.apply_1q_layer(m2)
            want = O.apply_1q_layer(want, n, m2)
        got = s.to_numpy()
    assert np.max(np.abs(got - want)) < TOL64


def test_pending_tail_layered_sweep_counts_at_30q(gpu):
    """U3 / RZZ ring / CNOT chain at 30 qubits: the first layer takes two
    sweeps and leaves eight bits pending; every further layer takes two (the
    crossing sweep and one plain sweep) and leaves nine; the last tail runs
    when the state is read. Checked against the same layers without pending
    tails."""
    from mettleq_amd import _native as nat
    from mettleq_amd.workloads import ring_bonds
    n = 30
    rng = np.random.default_rng(3000)
    bonds = ring_bonds(n)
    rows, c = _chain(n)
    mats = [[_rand_u(rng) for _ in range(n)] for _ in range(3)]
    ths = [[float(x) for x in rng.uniform(0, math.pi, len(bonds))] for _ in range(2)]
    res = []
    for tail in (1, 0):
        with _Tuned(DEFER_TAIL=tail), _state(n) as s:
            s.from_numpy(np.zeros(1, dtype=np.complex64) + 1, offset=0)
            counts = []
            for i, m in enumerate(mats):
                if i:
                    s.apply_phase_layer([nat.MSV_PHASE_ZZ] * len(bonds), bonds, ths[i - 1])
                    s.apply_xor_affine(rows, c)
                s0 = s.sweep_count()
                s.apply_1q_layer(m)
                counts.append(s.sweep_count() - s0)
            s0 = s.sweep_count()
            s.sync()
            counts.append(s.sweep_count() - s0)
            win = [s.to_numpy(o, 1 << 14) for o in (0, 1 << 29, (1 << 30) - (1 << 14))]
            res.append((counts, win))
    assert res[0][0] == [2, 2, 2, 1]
    assert res[1][0] == [3, 3, 3, 0]
    for a, b in zip(res[0][1], res[1][1]):
        assert np.max(np.abs(a - b)) < 2e-6


def test_pending_mask_follows_deferred_work(gpu):
    """msv_state_pending reports what the state holds back: the reset's basis
    state, a 1q layer's last rotations, a phase layer, a permutation; a sync
    runs all of it."""
    from mettleq_amd import _native as nat
    n = 20
    rng = np.random.default_rng(8500)
    bonds = [(i, (i + 1) % n) for i in range(n)]
    rows, c = _chain(n)
    with _state(n) as s:
        s.reset()
        assert s.pending() == nat.MSV_PENDING_BASIS
        s.apply_1q_layer([_rand_u(rng) for _ in range(n)])  # the product state: nothing left
        assert s.pending() == 0
        s.apply_1q_layer([_rand_u(rng) for _ in range(n)])  # two sweeps: the second stays pending
        assert s.pending() == nat.MSV_PENDING_TAIL
        s.apply_phase_layer([nat.MSV_PHASE_ZZ] * n, bonds, list(rng.uniform(0, 1, n)))
        assert s.pending() == nat.MSV_PENDING_TAIL | nat.MSV_PENDING_DIAG
        s.apply_xor_affine(rows, c)
        assert s.pending() == nat.MSV_PENDING_TAIL | nat.MSV_PENDING_DIAG | nat.MSV_PENDING_PERM
        s.sync()
        assert s.pending() == 0
