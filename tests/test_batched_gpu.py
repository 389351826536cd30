"""The batched estimator path (mettleq_amd.batched) on the GPU: a
parameterized rotation circuit with affine parameter maps, constant angles
and fixed phases, all bindings in one batched-engine launch; energies against
the oracle's circuit run per binding, gradients against FP64 central
differences of the circuit parameters."""
import numpy as np
import pytest

from oracle import sv_oracle as O

pytestmark = pytest.mark.gpu


def _circuit(n):
    from mettleq_amd.batched import ParamRef
    ops, p = [], 0
    for layer in range(2):
        for q in range(n):
            ops.append({"name": "RY", "wires": [q], "parameters": [ParamRef(p, 1.0, 0.0)]})
            ops.append({"name": "RZ", "wires": [q], "parameters": [ParamRef(p + 1, -0.5, 0.3)]})
            p += 2
        ops.append({"name": "RX", "wires": [layer], "parameters": [0.9]})
        for q in range(n):
            ops.append({"name": "CZ", "wires": [q, (q + 1) % n], "parameters": []})
        ops.append({"name": "ZZPHASE", "wires": [0, n - 1], "parameters": [0.21]})
    # a parameter used twice (shared)
    ops.append({"name": "RX", "wires": [2], "parameters": [ParamRef(0, 2.0, 0.0)]})
    return ops, p


def _energy(n, ops, vals, terms, const):
    from mettleq_amd.batched import ParamRef
    bound = [{**op, "parameters": [(x.coeff * vals[x.index] + x.offset) if isinstance(x, ParamRef) else x
                                   for x in op["parameters"]]} for op in ops]
    psi = O.run_ops(n, bound)
    e = const
    for qs, w in terms:
        e += w * float(np.real(O.pauli_expectation(psi, n, {q: "Z" for q in qs})))
    return e


def test_batched_estimator_energies_and_gradients(gpu):
    from mettleq_amd.batched import BatchedEstimator, ansatz_structure, z_terms
    n = 6
    ops, P = _circuit(n)
    st = ansatz_structure(n, ops)
    assert st is not None and st.n_params == P
    words = [{q: "Z", (q + 1) % n: "Z"} for q in range(n)] + [{1: "Z"}, {}]
    coeffs = [1.0] * n + [-0.5, 0.75]
    terms, const = z_terms(words, coeffs)
    vals = np.random.default_rng(7).uniform(-np.pi, np.pi, (40, P))
    est = BatchedEstimator()
    try:
        e = est.evaluate(st, terms, const, vals)
        e2, g = est.evaluate(st, terms, const, vals[:3], grads=True)
        e_only_const = est.evaluate(st, [], 0.4, vals[:2])
    finally:
        est.close()
    for b in (0, 17, 39):
        assert abs(e[b] - _energy(n, ops, vals[b], terms, const)) < 5e-6
    assert np.max(np.abs(e2 - e[:3])) < 1e-6
    h = 1e-4
    for b in range(2):
        for p in (0, 1, P - 1):
            hi, lo = vals[b].copy(), vals[b].copy()
            hi[p] += h
            lo[p] -= h
            fd = (_energy(n, ops, hi, terms, const) - _energy(n, ops, lo, terms, const)) / (2 * h)
            assert abs(g[b, p] - fd) < 1e-4
    assert np.allclose(e_only_const, 0.4, atol=1e-6)
