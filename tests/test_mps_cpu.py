"""The MPS oracle (oracle/mps_oracle.py) pinned on the CPU.

Independent checks: the dense statevector oracle where no truncation
happens (random circuits with adjacent, distant and descending two-qubit
operands); hand-computed truncation cases of the reference's rule
(mps_state.py:208-231); the Eckart-Young closed form of a truncation at the
orthogonality center (1 - F equals the relative discarded weight); GHZ bond
dimensions; transfer-matrix expectations against dense ones.
"""
import math

import numpy as np
import pytest

from oracle import mps_oracle as M
from oracle import sv_oracle as O


def _rand_u(rng, k):
    z = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def _random_circuit(n, count, rng):
    ops = []
    for _ in range(count):
        if rng.random() < 0.4:
            ops.append((rng.integers(n), _rand_u(rng, 1)))
        else:
            a, b = (int(x) for x in rng.choice(n, 2, replace=False))
            ops.append(((a, b), _rand_u(rng, 2)))
    return ops


def _run_dense(n, ops):
    s = O.zero_state(n)
    for q, u in ops:
        s = O.apply_dense(s, n, u, [int(q)] if np.ndim(q) == 0 else list(q))
    return s


def _run_mps(mps, ops):
    for q, u in ops:
        if np.ndim(q) == 0:
            mps.apply_single(u, int(q))
        else:
            mps.apply_two(u, q[0], q[1])


@pytest.mark.parametrize("n", [2, 3, 5, 8])
def test_oracle_exact_regime_matches_dense(n):
    rng = np.random.default_rng(40 + n)
    ops = _random_circuit(n, 6 * n, rng)
    mps = M.MPSOracle(n, dmax=1 << n, eps=0.0)
    _run_mps(mps, ops)
    want = _run_dense(n, ops)
    got = mps.to_statevector()
    assert abs(abs(np.vdot(want, got)) - 1.0) < 1e-5
    assert np.max(np.abs(got - want * (np.vdot(want, got) / abs(np.vdot(want, got))))) < 2e-5
    assert mps.trunc_events == 0
    # bonds never exceed the exact Schmidt bound min(2^i, 2^(n-i))
    assert all(d <= min(2 ** (i + 1), 2 ** (n - i - 1)) for i, d in enumerate(mps.bond_dims()))


def test_truncation_rule_cases():
    s = np.array([1.0, 0.5, 0.25, 1e-12])
    # eps cut: s >= 1e-10 * s0 drops the last one
    r, disc, rel, kept = M.truncate(s, dmax=8, eps=1e-10)
    assert r == 3 and disc == pytest.approx(1e-24) and kept == pytest.approx(math.sqrt(1.3125))
    # dmax cut
    r, disc, rel, kept = M.truncate(s, dmax=2, eps=0.0)
    assert r == 2
    assert disc == pytest.approx(0.0625 + 1e-24)
    assert rel == pytest.approx(0.0625 / 1.3125)
    # at least one value is kept whatever eps says
    r, *_ = M.truncate(np.array([1.0, 0.9]), dmax=4, eps=2.0)
    assert r == 1


def test_truncation_at_center_is_eckart_young():
    """One truncated split at the orthogonality center of a normalized MPS:
    the kept state's infidelity equals the relative discarded weight."""
    rng = np.random.default_rng(7)
    n = 6
    exact = M.MPSOracle(n, dmax=64, eps=0.0)
    ops = _random_circuit(n, 40, rng)
    _run_mps(exact, ops)
    psi = exact.to_statevector()
    u = _rand_u(rng, 2)
    full = M.MPSOracle(n, dmax=64, eps=0.0)
    full.A = [t.copy() for t in exact.A]
    full.center = exact.center
    full.apply_two(u, 2, 3)
    want = full.to_statevector()
    cut = M.MPSOracle(n, dmax=2, eps=0.0)
    cut.A = [t.copy() for t in exact.A]
    cut.center = exact.center
    # the split at bond 2-3 is the only one limited by dmax = 2 there
    cut._two_site(2, u, 0.0)
    assert cut.trunc_events == 1
    got = cut.to_statevector()
    f = abs(np.vdot(want, got)) ** 2
    # relative discarded weight of that split, recomputed from the exact spectrum
    full2 = M.MPSOracle(n, dmax=64, eps=0.0)
    full2.A = [t.copy() for t in exact.A]
    full2.center = exact.center
    full2._move_center(2)
    left = full2.A[2].astype(np.complex128)
    right = full2.A[3].astype(np.complex128)
    th = (left.reshape(-1, left.shape[2]) @ right.reshape(right.shape[0], -1)).reshape(left.shape[0], 4, -1)
    th = np.tensordot(u, th, axes=([1], [1])).transpose(1, 0, 2).reshape(left.shape[0] * 2, -1)
    s = np.linalg.svd(th, compute_uv=False)
    rel = float(np.sum(s[2:] ** 2) / np.sum(s ** 2))
    assert 1.0 - f == pytest.approx(rel, rel=1e-3, abs=1e-7)


def test_ghz_bonds_and_expectations():
    n = 12
    mps = M.MPSOracle(n, dmax=8, eps=1e-10)
    mps.apply_single(O.H(), 0)
    for i in range(n - 1):
        mps.apply_two(O.CNOT(), i, i + 1)
    assert mps.bond_dims() == [2] * (n - 1)
    z = np.diag([1.0, -1.0])
    x = np.array([[0, 1], [1, 0]], dtype=complex)
    assert mps.expectation_product({0: z, n - 1: z}).real == pytest.approx(1.0, abs=1e-6)
    assert mps.expectation_product({0: z}).real == pytest.approx(0.0, abs=1e-6)
    assert mps.expectation_product({q: x for q in range(n)}).real == pytest.approx(1.0, abs=1e-6)


@pytest.mark.parametrize("n", [4, 7])
def test_expectation_product_matches_dense(n):
    rng = np.random.default_rng(90 + n)
    mps = M.MPSOracle(n, dmax=64, eps=0.0)
    _run_mps(mps, _random_circuit(n, 5 * n, rng))
    psi = mps.to_statevector()
    ops = {int(q): _rand_u(rng, 1) for q in rng.choice(n, 3, replace=False)}
    dense = psi.copy()
    for q, u in ops.items():
        dense = O.apply_dense(dense, n, u, [q])
    want = np.vdot(psi, dense)
    assert abs(mps.expectation_product(ops) - want) < 1e-5


def test_descending_operands_conjugate_by_swap():
    """apply_two(U, c, t) with c > t acts as U on (c, t): the dense oracle
    lists c first as the gate's most significant bit."""
    rng = np.random.default_rng(3)
    n = 5
    u = _rand_u(rng, 2)
    mps = M.MPSOracle(n, dmax=32, eps=0.0)
    for q in range(n):
        mps.apply_single(_rand_u(rng, 1), q)
    psi0 = mps.to_statevector()
    mps.apply_two(u, 4, 1)
    want = O.apply_dense(psi0, n, u, [4, 1])
    got = mps.to_statevector()
    assert abs(abs(np.vdot(want, got)) - 1.0) < 1e-5
