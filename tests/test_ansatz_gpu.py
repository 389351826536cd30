"""Batched ansatz engine (msv_ansatz_*) vs the complex128 oracle: energies
within 5e-6, adjoint gradients against exact parameter shift (2e-5) and
FP64 central differences (h = 1e-4, 1e-4)."""
import math

import numpy as np
import pytest

from oracle import sv_oracle as O

pytestmark = pytest.mark.gpu

TOL_E = 5e-6
TOL_G = 2e-5


def _random_ansatz(n, layers, rng, shared=False):
    from mettleq_amd.ansatz import LayeredAnsatz
    a = LayeredAnsatz(n)
    gates = []
    p = 0
    for _ in range(layers):
        for q in range(n):
            for _k in range(int(rng.integers(1, 3))):
                name = ["RX", "RY", "RZ"][int(rng.integers(3))]
                a.rotation(name, q, p)
                gates.append((name, q, p))
                p += 1
        if n > 1:
            for q in range(n - 1):
                if rng.random() < 0.6:
                    if rng.random() < 0.5:
                        a.cz(q, q + 1)
                        gates.append(("CPHASE", q, q + 1, math.pi))
                    else:
                        th = float(rng.uniform(0, 3))
                        a.phase("ZZPHASE", q + 1, q, th)
                        gates.append(("ZZPHASE", q + 1, q, th))
    if shared and p > 0:
        a.rotation("RY", n - 1, 0)
        gates.append(("RY", n - 1, 0))
    terms = [((q, (q + 1) % n), float(rng.uniform(-1, 1))) for q in range(n)] if n > 1 else [((0,), 1.0)]
    terms.append(((0,), 0.3))
    a.observable(terms)
    return a, gates, terms


@pytest.mark.parametrize("n,layers,batch", [(1, 2, 3), (3, 2, 2), (4, 3, 5), (6, 2, 3), (9, 2, 4), (13, 2, 2)])
def test_energies_and_gradients_match_oracle(gpu, n, layers, batch):
    rng = np.random.default_rng(100 * n + batch)
    a, gates, terms = _random_ansatz(n, layers, rng, shared=True)
    with a:
        params = rng.uniform(-math.pi, math.pi, (batch, a.n_params))
        e, g = a.evaluate(params)
        for b in range(batch):
            want_e = O.ansatz_energy(n, gates, params[b], terms)
            assert abs(e[b] - want_e) < TOL_E, (b, e[b], want_e)
            want_g = O.ansatz_param_shift(n, gates, params[b], terms)
            assert np.max(np.abs(g[b] - want_g)) < TOL_G, (b, np.max(np.abs(g[b] - want_g)))


def test_wide_angles_match_oracle(gpu):
    """Angles far outside [-pi, pi] and at the points where the half angle
    is reduced by a multiple of pi (+-pi, +-2 pi, 3 pi, ...): the dropped
    global signs must leave energies and gradients unchanged."""
    rng = np.random.default_rng(77)
    n = 7
    a, gates, terms = _random_ansatz(n, 3, rng)
    with a:
        params = rng.uniform(-9 * math.pi, 9 * math.pi, (4, a.n_params))
        special = np.array([math.pi, -math.pi, 2 * math.pi, -2 * math.pi, 3 * math.pi, 0.0, 1e3, -47.25])
        params[0, :] = special[np.arange(a.n_params) % len(special)]
        params[1, :] = np.nextafter(special, np.inf)[np.arange(a.n_params) % len(special)]
        e, g = a.evaluate(params)
        for b in range(params.shape[0]):
            want_e = O.ansatz_energy(n, gates, params[b], terms)
            assert abs(e[b] - want_e) < TOL_E, (b, e[b], want_e)
            want_g = O.ansatz_param_shift(n, gates, params[b], terms)
            assert np.max(np.abs(g[b] - want_g)) < TOL_G, (b, np.max(np.abs(g[b] - want_g)))


@pytest.mark.parametrize("seq", [("RY",), ("RX",), ("RZ", "RY"), ("RY", "RZ"), ("RX", "RZ"), ("RZ", "RX"),
                                 ("RZ", "RY", "RZ")])
def test_compiled_stage_programs_match_oracle(gpu, seq):
    """Layered ansaetze whose every qubit runs the same rotation sequence
    per layer, with CZ rings: every sweep runs a kernel compiled for its
    stage program (forward: seq; adjoint: reversed), checked against the
    oracle. 14 qubits: a 10-bit low sweep and a high sweep per layer."""
    from mettleq_amd.ansatz import LayeredAnsatz
    n, layers, batch = 14, 2, 3
    rng = np.random.default_rng(len(seq) * 31 + sum(map(ord, "".join(seq))))
    a = LayeredAnsatz(n)
    gates, p = [], 0
    for _ in range(layers):
        for q in range(n):
            for name in seq:
                a.rotation(name, q, p)
                gates.append((name, q, p))
                p += 1
        for q in range(n):
            a.cz(q, (q + 1) % n)
            gates.append(("CPHASE", q, (q + 1) % n, math.pi))
    terms = [((q, (q + 1) % n), float(rng.uniform(-1, 1))) for q in range(n)] + [((3,), 0.4)]
    a.observable(terms)
    with a:
        params = rng.uniform(-2 * math.pi, 2 * math.pi, (batch, a.n_params))
        e, g = a.evaluate(params)
        info = a.info(batch)
        assert info["compiled_program_sweeps"] == info["forward_sweeps"] + info["backward_sweeps"], info
        e_only = a.evaluate(params, grads=False)
        assert np.array_equal(e_only, e)
        for b in range(batch):
            want_e = O.ansatz_energy(n, gates, params[b], terms)
            assert abs(e[b] - want_e) < TOL_E, (b, e[b], want_e)
            want_g = O.ansatz_param_shift(n, gates, params[b], terms)
            assert np.max(np.abs(g[b] - want_g)) < TOL_G, (b, np.max(np.abs(g[b] - want_g)))


def test_rz_diagonal_layouts_match_oracle(gpu):
    """RZ rotations that end their qubits' rotations in a block run as one
    diagonal per sweep (forward after the last segment, adjoint before the
    first). Blocks here: RZ only (sweeps with a diagonal and no stage),
    RX/RY followed by RZ on some qubits and ending on RX/RY on others, and a
    run of two RZ on one qubit sharing one parameter (the first stays a
    stage rotation, the second joins the diagonal); CZ rings and ZZ phases
    between blocks; 14 qubits, so every block has a low and a high sweep
    and the diagonal holds register and thread bits."""
    from mettleq_amd.ansatz import LayeredAnsatz
    n, batch = 14, 2
    rng = np.random.default_rng(1414)
    blocks = [
        [("RZ",)] * n,
        [("RX", "RZ") if q % 3 else ("RY",) for q in range(n)],
        [("RY", "RZ", "RZ") if q % 2 else ("RZ", "RX") for q in range(n)],
        [("RZ",) if q < 7 else ("RY", "RZ") for q in range(n)],
    ]
    a = LayeredAnsatz(n)
    gates, p = [], 0
    for k, block in enumerate(blocks):
        for q, seq in enumerate(block):
            shared = None
            for name in seq:
                if name == "RZ" and len(seq) == 3:
                    # both RZ of the run take one parameter
                    if shared is None:
                        shared = p
                        p += 1
                    idx = shared
                else:
                    idx = p
                    p += 1
                a.rotation(name, q, idx)
                gates.append((name, q, idx))
        for q in range(n):
            if k % 2 == 0:
                a.cz(q, (q + 1) % n)
                gates.append(("CPHASE", q, (q + 1) % n, math.pi))
            else:
                th = float(rng.uniform(0, 3))
                a.phase("ZZPHASE", (q + 1) % n, q, th)
                gates.append(("ZZPHASE", (q + 1) % n, q, th))
    terms = [((q, (q + 1) % n), float(rng.uniform(-1, 1))) for q in range(n)] + [((5,), 0.4), ((12,), -0.3)]
    a.observable(terms)
    with a:
        params = rng.uniform(-2 * math.pi, 2 * math.pi, (batch, a.n_params))
        e, g = a.evaluate(params)
        for b in range(batch):
            want_e = O.ansatz_energy(n, gates, params[b], terms)
            assert abs(e[b] - want_e) < TOL_E, (b, e[b], want_e)
            want_g = O.ansatz_param_shift(n, gates, params[b], terms)
            assert np.max(np.abs(g[b] - want_g)) < TOL_G, (b, np.max(np.abs(g[b] - want_g)))


def test_gradients_match_central_differences(gpu):
    rng = np.random.default_rng(5)
    n = 5
    a, gates, terms = _random_ansatz(n, 2, rng)
    with a:
        params = rng.uniform(-math.pi, math.pi, (2, a.n_params))
        _, g = a.evaluate(params)
        for b in range(2):
            fd = O.ansatz_central_diff(n, gates, params[b], terms, h=1e-4)
            assert np.max(np.abs(g[b] - fd)) < 1e-4


def test_hardware_efficient_16q_config_shape(gpu):
    """16 qubits, 8 layers of RY, RZ + CZ ring (256 parameters), ZZ-ring
    observable: a batch of 8 bindings checked binding by binding."""
    from mettleq_amd.ansatz import hardware_efficient, zz_ring_terms
    n, L, batch = 16, 8, 8
    rng = np.random.default_rng(20261203)
    params = rng.uniform(-math.pi, math.pi, (batch, 2 * n * L))
    gates = []
    for l in range(L):
        for q in range(n):
            gates.append(("RY", q, 2 * (l * n + q)))
            gates.append(("RZ", q, 2 * (l * n + q) + 1))
        for q in range(n):
            gates.append(("CPHASE", q, (q + 1) % n, math.pi))
    terms = zz_ring_terms(n)
    with hardware_efficient(n, L) as a:
        a.observable(terms)
        e, g = a.evaluate(params)
        assert a.n_params == 256
        for b in range(batch):
            assert abs(e[b] - O.ansatz_energy(n, gates, params[b], terms)) < TOL_E
        # parameter shift on a sample of parameters of binding 0
        for p in (0, 1, 37, 128, 200, 255):
            vals = []
            for sh in (math.pi / 2, -math.pi / 2):
                q = params[0].copy()
                q[p] += sh
                vals.append(O.ansatz_energy(n, gates, q, terms))
            assert abs(g[0, p] - 0.5 * (vals[0] - vals[1])) < TOL_G, p


def test_forward_only_and_replanning_across_batch_sizes(gpu):
    rng = np.random.default_rng(9)
    n = 7
    a, gates, terms = _random_ansatz(n, 2, rng)
    with a:
        for batch in (4, 9, 4, 1):
            params = rng.uniform(-math.pi, math.pi, (batch, a.n_params))
            e_only = a.evaluate(params, grads=False)
            e, g = a.evaluate(params)
            assert np.array_equal(e_only, e)
            for b in range(batch):
                assert abs(e[b] - O.ansatz_energy(n, gates, params[b], terms)) < TOL_E
            assert np.max(np.abs(g[0] - O.ansatz_param_shift(n, gates, params[0], terms))) < TOL_G


def test_gradients_are_bit_reproducible(gpu):
    rng = np.random.default_rng(3)
    a, _, _ = _random_ansatz(8, 2, rng)
    with a:
        params = rng.uniform(-math.pi, math.pi, (6, a.n_params))
        e1, g1 = a.evaluate(params)
        e2, g2 = a.evaluate(params)
        assert np.array_equal(e1, e2) and np.array_equal(g1, g2)


def test_torch_autograd_function(gpu):
    import torch
    from mettleq_amd.autograd import ansatz_energy
    rng = np.random.default_rng(11)
    a, gates, terms = _random_ansatz(5, 2, rng)
    with a:
        params = torch.tensor(rng.uniform(-math.pi, math.pi, (3, a.n_params)), dtype=torch.float64,
                              requires_grad=True)
        energy = ansatz_energy(params, a)
        weights = torch.tensor([1.0, -2.0, 0.5], dtype=torch.float64)
        (energy * weights).sum().backward()
        for b in range(3):
            want = O.ansatz_param_shift(5, gates, params[b].detach().numpy(), terms)
            assert np.max(np.abs(params.grad[b].numpy() - weights[b].item() * want)) < 3 * TOL_G
        # inference: no gradient requested, the forward pass alone, same energies
        with torch.no_grad():
            e_inf = ansatz_energy(params.detach(), a)
        assert not e_inf.requires_grad
        assert np.max(np.abs(e_inf.numpy() - energy.detach().numpy())) < 1e-12


def test_invalid_arguments(gpu):
    from mettleq_amd import _native as nat
    from mettleq_amd.ansatz import LayeredAnsatz
    with LayeredAnsatz(3) as a:
        with pytest.raises(ValueError):
            a.rotation("RY", 3, 0)
        with pytest.raises(ValueError):
            a.phase("CPHASE", 1, 1, 1.0)
        a.rotation("RX", 0, 0)
        with pytest.raises(ValueError):
            a.evaluate(np.zeros((2, 1)))  # observable not set
        a.observable([((0,), 1.0)])
        with pytest.raises(ValueError):
            a.evaluate(np.zeros((2, 2)))
        with pytest.raises(ValueError):
            a.evaluate(np.full((1, 1), np.nan))
    assert nat.MSV_ROT_Z == 2
