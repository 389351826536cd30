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


def test_mps_large_bonds_take_the_global_memory_paths(gpu):
    """Bonds past 64: a 14-qubit brick circuit of random two-qubit gates in
    the exact regime (dmax 128 = the largest Schmidt rank) reaches a bond of
    128, so two-site matrices of 64 x 256 go through the global-memory SVD
    and 128 x 128 center moves through the Jacobi factorization (beyond the
    LDS QR). The state must match the dense oracle."""
    from mettleq_amd.mps import MPSOptions, MPSState
    n = 14
    rng = np.random.default_rng(1400)
    ops = []
    for layer in range(9):
        for q in range(n):
            ops.append((q, _rand_u(rng, 1)))
        for a in range(layer % 2, n - 1, 2):
            ops.append(((a, a + 1), _rand_u(rng, 2)))
    want = O.zero_state(n)
    for q, u in ops:
        want = O.apply_dense(want, n, u, [q] if np.ndim(q) == 0 else list(q))
    with MPSState(n, MPSOptions(dmax=128, eps=0.0)) as mps:
        _run(mps, ops)
        assert max(mps.bond_dims()) == 128
        assert not mps.truncated()
        got = mps.to_statevector()
    assert _overlap(want, got) > 1 - 1e-5


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
        shots = mps.sample_array(500, [0, 3], rng=np.random.default_rng(9))
        assert shots.shape == (500, 2) and set(np.unique(shots)) <= {0, 1}
        counts = mps.sample_counts(200, [3])
        assert sum(counts.values()) == 200 and set(counts) <= {"0", "1"}
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


def test_device_mps_backend_runs_op_lists(gpu):
    """Device(backend="mps") runs the reference's op dicts (device.py:91-128,
    1138-1290) on the MPS; the result matches the dense oracle."""
    from mettleq_amd.device import Device
    from mettleq_amd.mps import MPSOptions
    n = 8
    rng = np.random.default_rng(21)
    names1 = ["H", "X", "S", "T", "RX", "RY", "RZ", "U3"]
    ops = []
    for _ in range(60):
        if rng.random() < 0.5:
            nm = names1[rng.integers(len(names1))]
            params = {"RX": 1, "RY": 1, "RZ": 1, "U3": 3}.get(nm, 0)
            ops.append({"name": nm, "wires": [int(rng.integers(n))],
                        "parameters": [float(x) for x in rng.uniform(0, 6, params)]})
        else:
            nm = ["CNOT", "CZ", "CPHASE", "SWAP", "CRY", "ZZPHASE", "ISWAP"][rng.integers(7)]
            a, b = (int(x) for x in rng.choice(n, 2, replace=False))
            params = [float(rng.uniform(0, 3))] if nm in ("CPHASE", "CRY", "ZZPHASE") else []
            ops.append({"name": nm, "wires": [a, b], "parameters": params})
    want = O.run_ops(n, ops)
    dev = Device(n, backend="mps", mps_opts=MPSOptions(dmax=64, eps=0.0))
    try:
        dev.execute(ops)
        got = dev.sim.to_statevector()
        p = dev.sim.probabilities_array([2, 5])
    finally:
        dev.close()
    assert _overlap(want, got) > 1 - 1e-5
    pw = (np.abs(want) ** 2).reshape([2] * n).sum(axis=tuple(q for q in range(n) if q not in (2, 5))).reshape(-1)
    assert np.max(np.abs(p - pw)) < 1e-5
    with pytest.raises(ValueError):
        Device(4, backend="mps").execute([{"name": "CCX", "wires": [0, 1, 2]}])


def test_sdk_seam_runs_the_mps_method(gpu):
    """execute_operations(method="mps") and the approximate fallback of an
    automatic request (_common.py:118-228, planning.py:108-204): statevector,
    marginals and Pauli words read from the MPS."""
    from mettleq_amd.integrations import _common as C
    n = 6
    ops = [C.core_operation("H", [0])] + [C.core_operation("CNOT", [i, i + 1]) for i in range(n - 1)]
    ops += [C.core_operation("RY", [2], [0.4]), C.core_operation("CZ", [1, 4])]
    want = O.run_ops(n, ops)
    dev = C.execute_operations(n, ops, method="mps")
    try:
        assert dev.backend == "mps"
        assert dev.execution_selection["selected_method"] == "matrix_product_state"
        got = C.statevector_numpy(dev)
        assert _overlap(want, got) > 1 - 1e-5
        p = C.marginal_probabilities(dev, [0, 2])
        pw = (np.abs(want) ** 2).reshape([2] * n).sum(axis=(1, 3, 4, 5)).reshape(-1)
        assert np.max(np.abs(p - pw)) < 1e-5
        zz = C.pauli_product_expectation(dev, {0: "Z", 5: "Z"})
        psi_z = O.apply_dense(O.apply_dense(want, n, np.diag([1, -1]), [0]), n, np.diag([1, -1]), [5])
        assert abs(zz - np.vdot(want, psi_z)) < 1e-4
        C.apply_global_phase(dev, 0.3)
        assert abs(np.vdot(got, C.statevector_numpy(dev)) - np.exp(0.3j)) < 1e-4
    finally:
        dev.close()
    # 40 qubits: the preflight refuses no statevector here, but a local chain
    # above automatic_mps_min_qubits with approximation allowed runs as MPS
    n = 40
    chain = [C.core_operation("H", [0])] + [C.core_operation("CNOT", [i, i + 1]) for i in range(n - 1)]
    dev = C.execute_operations(n, chain, allow_approximation=True, automatic_mps_min_qubits=24)
    try:
        assert dev.backend == "mps"
        vals = C.pauli_product_expectations(dev, [{0: "Z", 39: "Z"}, {q: "X" for q in range(n)}, {7: "Z"}])
        assert abs(vals[0] - 1) < 1e-4 and abs(vals[1] - 1) < 1e-4 and abs(vals[2]) < 1e-4
    finally:
        dev.close()
