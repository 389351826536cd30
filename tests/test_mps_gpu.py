"""The device MPS engine (mettleq_amd.mps, msv_mps_* in the C ABI) against the
MPS oracle and the dense statevector oracle.

Tolerances: complex64 site tensors and a single-precision device SVD, so
states agree to |<want|got>| >= 1 - 1e-5 and expectation values to 1e-4;
bond dimensions and truncation counts must be equal; discarded weights agree
to 1e-3 relative (they are differences of single-precision squares).
"""
import math

import numpy as np
import pytest

from oracle import mps_oracle as M
from oracle import sv_oracle as O

pytestmark = pytest.mark.gpu


def _rand_u(rng, k):
    z = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def _random_circuit(n, count, rng, max_dist=None):
    ops = []
    for _ in range(count):
        if n == 1 or rng.random() < 0.4:
            ops.append((int(rng.integers(n)), _rand_u(rng, 1)))
        else:
            while True:
                a, b = (int(x) for x in rng.choice(n, 2, replace=False))
                if max_dist is None or abs(a - b) <= max_dist:
                    break
            ops.append(((a, b), _rand_u(rng, 2)))
    return ops


def _run(mps, ops):
    for q, u in ops:
        if np.ndim(q) == 0:
            mps.apply_single(u, q)
        else:
            mps.apply_two(u, q[0], q[1])


def _overlap(a, b):
    return abs(np.vdot(a, b)) / (np.linalg.norm(a) * np.linalg.norm(b))


@pytest.mark.parametrize("n", [1, 2, 5, 9, 12])
def test_mps_exact_regime_matches_dense_oracle(gpu, n):
    from mettleq_amd.mps import MPSOptions, MPSState
    rng = np.random.default_rng(300 + n)
    ops = _random_circuit(n, 5 * n + 3, rng)
    want = O.zero_state(n)
    for q, u in ops:
        want = O.apply_dense(want, n, u, [q] if np.ndim(q) == 0 else list(q))
    with MPSState(n, MPSOptions(dmax=1 << max(1, n // 2), eps=0.0)) as mps:
        _run(mps, ops)
        got = mps.to_statevector()
        assert not mps.truncated()
        assert abs(mps.norm() - 1.0) < 1e-4
    assert _overlap(want, got) > 1 - 1e-5
    assert abs(np.linalg.norm(got) - 1.0) < 1e-5


@pytest.mark.parametrize("n,dmax,eps", [(8, 4, 1e-10), (10, 6, 1e-10), (12, 8, 1e-3)])
def test_mps_truncation_matches_oracle(gpu, n, dmax, eps):
    from mettleq_amd.mps import MPSOptions, MPSState
    rng = np.random.default_rng(700 + n)
    ops = _random_circuit(n, 8 * n, rng, max_dist=3)
    ref = M.MPSOracle(n, dmax=dmax, eps=eps)
    for q, u in ops:
        if np.ndim(q) == 0:
            ref.apply_single(u, q)
        else:
            ref.apply_two(u, q[0], q[1])
    with MPSState(n, MPSOptions(dmax=dmax, eps=eps)) as mps:
        _run(mps, ops)
        assert mps.bond_dims() == ref.bond_dims()
        assert mps.trunc_count() == ref.trunc_events
        assert mps.truncated() == (ref.trunc_events > 0)
        d = mps.truncation_diagnostics()
        assert d["local_discarded_weight_sum"] == pytest.approx(ref.discarded_sum, rel=1e-3, abs=1e-9)
        assert d["maximum_bond_dimension_reached"] == ref.max_bond_ever <= dmax
        got = mps.to_statevector()
    assert _overlap(ref.to_statevector(), got) > 1 - 1e-4


def test_mps_wide_ghz_expectations(gpu):
    """A 48-qubit GHZ state: bonds of 2, and <Z_0 Z_47> = <X^48> = 1,
    <Z_0> = 0, read on the MPS by transfer matrices (no dense state)."""
    from mettleq_amd.mps import MPSOptions, MPSState
    n = 48
    z = np.diag([1.0, -1.0]).astype(complex)
    x = np.array([[0, 1], [1, 0]], dtype=complex)
    with MPSState(n, MPSOptions(dmax=16)) as mps:
        mps.apply_single(O.H(), 0)
        for i in range(n - 1):
            mps.apply_two(O.CNOT(), i, i + 1)
        assert mps.bond_dims() == [2] * (n - 1)
        assert abs(mps.expectation_product({0: z, n - 1: z}) - 1.0) < 1e-4
        assert abs(mps.expectation_product({0: z})) < 1e-4
        assert abs(mps.expectation_product({q: x for q in range(n)}) - 1.0) < 1e-4
        # a long-range gate through the swap network keeps it a GHZ state
        mps.apply_two(O.CZ(), 0, n - 1)
        assert abs(mps.expectation_product({0: z, 20: z}) - 1.0) < 1e-4
        # the swaps only drop numerically zero singular values (rank 2 of 4)
        assert mps.bond_dims() == [2] * (n - 1)
        assert mps.truncation_diagnostics()["local_discarded_weight_sum"] < 1e-9


def test_mps_expectation_matches_dense(gpu):
    from mettleq_amd.mps import MPSOptions, MPSState
    n = 9
    rng = np.random.default_rng(11)
    ops = _random_circuit(n, 40, rng)
    with MPSState(n, MPSOptions(dmax=32, eps=0.0)) as mps:
        _run(mps, ops)
        psi = mps.to_statevector().astype(np.complex128)
        for trial in range(3):
            wires = [int(q) for q in rng.choice(n, 1 + trial, replace=False)]
            opers = {q: _rand_u(rng, 1) for q in wires}
            dense = psi.copy()
            for q, u in opers.items():
                dense = O.apply_dense(dense, n, u, [q])
            want = np.vdot(psi, dense) / np.vdot(psi, psi)
            assert abs(mps.expectation_product(opers) - want) < 1e-4


def test_mps_gate_forms_and_readout(gpu):
    """Diagonal, ZZ-phase and descending-operand gates against the dense
    oracle; probabilities and samples come from the engine's reductions."""
    from mettleq_amd.mps import MPSOptions, MPSState
    n = 7
    rng = np.random.default_rng(5)
    want = O.zero_state(n)
    with MPSState(n, MPSOptions(dmax=64, eps=0.0)) as mps:
        for q in range(n):
            u = _rand_u(rng, 1)
            mps.apply_single(u, q)
            want = O.apply_dense(want, n, u, [q])
        d = np.exp(1j * rng.uniform(0, 2 * math.pi, 4))
        mps.apply_diagonal(d, [5, 1])
        want = O.apply_dense(want, n, np.diag(d), [5, 1])
        mps.apply_zz_phase(0.37, 0, 6)
        zz = np.diag(np.exp(-1j * 0.37 * np.array([1, -1, -1, 1])))
        want = O.apply_dense(want, n, zz, [0, 6])
        u4 = _rand_u(rng, 2)
        mps.apply_two(u4, 6, 2)
        want = O.apply_dense(want, n, u4, [6, 2])
        got = mps.to_statevector()
        assert _overlap(want, got) > 1 - 1e-5
        p = mps.probabilities_array([0, 3])
        pw = np.abs(got.astype(np.complex128)) ** 2
        pw = pw.reshape([2] * n).sum(axis=tuple(q for q in range(n) if q not in (0, 3))).reshape(-1)
        assert np.max(np.abs(p - pw)) < 1e-6
        shots = mps.sample(500, [0, 3], seed=9)
        assert shots.shape == (500,) and shots.max() < 4
        nrm = mps.norm()
        mps.normalize()
        assert abs(mps.norm() - 1.0) < 1e-5 and nrm > 0


def test_mps_invalid_arguments(gpu):
    from mettleq_amd.mps import MPSOptions, MPSState
    with pytest.raises(ValueError):
        MPSState(0)
    with pytest.raises(ValueError):
        MPSState(3, MPSOptions(dmax=0))
    with MPSState(3) as mps:
        with pytest.raises(ValueError):
            mps.apply_single(np.eye(2), 3)
        with pytest.raises(ValueError):
            mps.apply_two(np.eye(4), 1, 1)
        with pytest.raises(ValueError):
            mps.apply_two(np.eye(2), 0, 1)
        with pytest.raises(ValueError):
            mps.expectation_product({0: np.eye(2), 7: np.eye(2)})
        mps.reset()
        assert mps.bond_dims() == [1, 1]
