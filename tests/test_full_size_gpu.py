"""Full-size properties on the GPU: 30 qubits, the layered circuit's shape
(8 GiB complex64 states), where the oracle cannot run. Size-independent
checks only:

* a layer cycle (U3 layer, RZZ ring, CNOT chain, QFT, ...) followed by its
  exact inverse returns the start state (1 - F against an untouched copy),
  through the deferred and fused paths;
* an XOR permutation followed by its inverse restores the state bit for
  bit (amplitude windows and an FP64 marginal compared with ==);
* the QFT of a basis state, with and without an RZZ ring and a CNOT chain
  read in its first sweep, matches the closed form
  exp(2 pi i bitrev(y) k' / N) / sqrt(N) times the ring's phase at k, where
  k' is the chain's image of k (qft_matrix in the oracle, here evaluated at
  sampled output indices);
* complex64 against the complex128 mode after one cycle: 1 - F <= 1e-6.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 30
WINDOWS = [(0, 1 << 16), ((1 << 29) + 12345, 1 << 16), ((1 << 30) - (1 << 16), 1 << 16)]


def _mats(rng, n):
    from mettleq_amd import gates as G
    return [G.single_qubit_matrix("U3", rng.uniform(0, 2 * math.pi, 3)) for _ in range(n)]


def _chain():
    return [{"name": "CNOT", "wires": [i, i + 1]} for i in range(N - 1)]


def _prepare(st, seed):
    """A non-product start state: U3 layer, CZ brick, U3 layer."""
    from mettleq_amd import _native as nat
    rng = np.random.default_rng(seed)
    bonds = [(q, q + 1) for q in range(0, N - 1, 2)]
    st.reset()
    st.apply_1q_layer(_mats(rng, N))
    st.apply_phase_layer([nat.MSV_PHASE_CPHASE] * len(bonds), bonds, [math.pi] * len(bonds))
    st.apply_1q_layer(_mats(rng, N))


def _fidelity(a, b):
    ip = a.inner(b)
    return abs(ip) ** 2 / (a.norm2() * b.norm2())


def test_layer_cycle_then_inverse_returns_the_state_30q(gpu):
    from mettleq_amd import _native as nat
    from mettleq_amd.layers import compose_inverse_affine
    from mettleq_amd.native_state import NativeState
    from mettleq_amd.workloads import ring_bonds
    rng = np.random.default_rng(3030)
    bonds = ring_bonds(N)
    cycle = []
    for _ in range(2):
        cycle.append(("u3", _mats(rng, N)))
        cycle.append(("rzz", [float(x) for x in rng.uniform(0, math.pi, len(bonds))]))
        cycle.append(("cnot", None))
    cycle.append(("qft", None))
    cycle.append(("u3", _mats(rng, N)))
    fwd = compose_inverse_affine(_chain(), N)
    bwd = compose_inverse_affine(list(reversed(_chain())), N)

    def apply(st, kind, arg, inverse):
        if kind == "u3":
            st.apply_1q_layer([m.conj().T for m in arg] if inverse else arg)
        elif kind == "rzz":
            st.apply_phase_layer([nat.MSV_PHASE_ZZ] * len(bonds), bonds, [-t if inverse else t for t in arg])
        elif kind == "cnot":
            st.apply_xor_affine(*(bwd if inverse else fwd))
        else:
            st.apply_qft(0, N, inverse=inverse)

    with NativeState(N) as ref, NativeState(N) as st:
        _prepare(ref, 7)
        _prepare(st, 7)
        s0 = st.sweep_count()
        for kind, arg in cycle:
            apply(st, kind, arg, False)
        forward_sweeps = st.sweep_count() - s0
        # mid-way the state differs from the start
        assert _fidelity(ref, st) < 0.5
        for kind, arg in reversed(cycle):
            apply(st, kind, arg, True)
        f = _fidelity(ref, st)
        assert abs(st.norm2() - 1.0) < 1e-5
    # RZZ rings and CNOT chains ride in the next sweep; each U3 layer leaves
    # its last sweep's rotations pending and the next U3 layer runs them in
    # its first sweep (two sweeps per layer), the QFT runs them first (one
    # sweep + three)
    assert forward_sweeps == 2 + 2 + (1 + 3) + 2
    assert 1.0 - f < 1e-6


def test_permutation_round_trip_is_bit_exact_30q(gpu):
    from mettleq_amd.layers import compose_inverse_affine
    from mettleq_amd.native_state import NativeState
    rng = np.random.default_rng(4040)
    ops = []
    for _ in range(3 * N):
        a, b = rng.choice(N, 2, replace=False)
        ops.append({"name": "CNOT", "wires": [int(a), int(b)]})
    ops.append({"name": "X", "wires": [5]})
    with NativeState(N) as st:
        _prepare(st, 11)
        before = [st.to_numpy(o, c) for o, c in WINDOWS]
        marg0 = st.probabilities(list(range(20)))
        for run in (_chain(), ops):
            st.apply_xor_affine(*compose_inverse_affine(run, N))
            st.apply_xor_affine(*compose_inverse_affine(list(reversed(run)), N))
        after = [st.to_numpy(o, c) for o, c in WINDOWS]
        marg1 = st.probabilities(list(range(20)))
    for x, y in zip(before, after):
        assert np.array_equal(x, y)
    assert np.array_equal(marg0, marg1)


def _bitrev(y, bits):
    r = np.zeros_like(y)
    for b in range(bits):
        r |= ((y >> b) & 1) << (bits - 1 - b)
    return r


@pytest.mark.parametrize("with_ring_and_chain", [False, True])
@pytest.mark.parametrize("defer", [1, 0])
def test_qft_of_basis_state_matches_closed_form_30q(gpu, with_ring_and_chain, defer):
    from mettleq_amd import _native as nat
    from mettleq_amd.layers import compose_inverse_affine
    from mettleq_amd.native_state import NativeState
    from mettleq_amd.workloads import ring_bonds
    rng = np.random.default_rng(5050)
    k = int(rng.integers(1 << N))
    bonds = ring_bonds(N)
    th = [float(x) for x in rng.uniform(0, math.pi, len(bonds))]
    phase = 0.0
    k2 = k
    if with_ring_and_chain:
        for (a, b), t in zip(bonds, th):
            pa = (k >> (N - 1 - a)) & 1
            pb = (k >> (N - 1 - b)) & 1
            phase += -t * (1 - 2 * (pa ^ pb))
        for i in range(N - 1):  # CNOT i -> i+1 on the basis index
            if (k2 >> (N - 1 - i)) & 1:
                k2 ^= 1 << (N - 2 - i)
    old = (nat.get_tuning(nat.MSV_TUNE_DEFER_DIAG), nat.get_tuning(nat.MSV_TUNE_DEFER_PERM))
    try:
        nat.set_tuning(nat.MSV_TUNE_DEFER_DIAG, defer)
        nat.set_tuning(nat.MSV_TUNE_DEFER_PERM, defer)
        with NativeState(N) as st:
            st.set_basis(k)
            if with_ring_and_chain:
                st.apply_phase_layer([nat.MSV_PHASE_ZZ] * len(bonds), bonds, th)
                st.apply_xor_affine(*compose_inverse_affine(_chain(), N))
            s0 = st.sweep_count()
            st.apply_qft(0, N)
            sweeps = st.sweep_count() - s0
            got = [st.to_numpy(o, c) for o, c in WINDOWS]
    finally:
        nat.set_tuning(nat.MSV_TUNE_DEFER_DIAG, old[0])
        nat.set_tuning(nat.MSV_TUNE_DEFER_PERM, old[1])
    size = 1 << N
    for (o, c), g in zip(WINDOWS, got):
        y = np.arange(o, o + c, dtype=np.int64)
        m = (_bitrev(y, N) * k2) % size
        want = np.exp(1j * (2 * math.pi * m / size + phase)) / math.sqrt(size)
        # amplitudes are 2^-15 in magnitude: 1e-9 is 3e-5 relative
        assert np.max(np.abs(g - want)) < 1e-9
    if with_ring_and_chain and defer:
        assert sweeps == 3  # the ring and the chain ride in the first QFT sweep


def test_complex64_against_complex128_after_a_cycle_30q(gpu):
    from mettleq_amd import _native as nat
    from mettleq_amd.layers import compose_inverse_affine
    from mettleq_amd.native_state import NativeState
    from mettleq_amd.workloads import ring_bonds
    rng = np.random.default_rng(6060)
    mats = _mats(rng, N)
    bonds = ring_bonds(N)
    th = [float(x) for x in rng.uniform(0, math.pi, len(bonds))]
    rows, c = compose_inverse_affine(_chain(), N)
    states = []
    try:
        for dt in ("complex64", "complex128"):
            st = NativeState(N, dt)
            states.append(st)
            st.reset()
            st.apply_1q_layer(mats)
            st.apply_phase_layer([nat.MSV_PHASE_ZZ] * len(bonds), bonds, th)
            st.apply_xor_affine(rows, c)
            st.apply_qft(0, N)
            st.apply_1q_layer(mats)
        f = _fidelity(states[0], states[1])
    finally:
        for st in states:
            st.close()
    assert 1.0 - f < 1e-6


def test_brick_layers_complex64_against_complex128_30q(gpu):
    """U3 layers and CZ bricks (a lone deferred diagonal between layers, each
    layer's last sweep pending): complex64 against the complex128 mode after
    six layers, 1 - F <= 1e-6."""
    from mettleq_amd import _native as nat
    from mettleq_amd.native_state import NativeState
    rng = np.random.default_rng(7070)
    layers = [(_mats(rng, N), [(q, q + 1) for q in range(l % 2, N - 1, 2)]) for l in range(6)]
    states = []
    try:
        for dt in ("complex64", "complex128"):
            st = NativeState(N, dt)
            states.append(st)
            st.reset()
            for mats, bonds in layers:
                st.apply_1q_layer(mats)
                st.apply_phase_layer([nat.MSV_PHASE_CPHASE] * len(bonds), bonds, [math.pi] * len(bonds))
        f = _fidelity(states[0], states[1])
    finally:
        for st in states:
            st.close()
    assert 1.0 - f < 1e-6
