"""Host logic on the CPU: fusion patterns, gate routing and decompositions.

The fused op list, executed with the documented semantics of each native
entry point (tests/emulator.py), must equal the unfused dense oracle -- the
A/B structure of the reference's kernel tests (tests/mlxQMetalKernelsTest.py)
with the device replaced by its specification.
"""
import math

import numpy as np
import pytest

from mettleq_amd.fusion import fuse_operations, gates_for_op
from mettleq_amd.workloads import brick_20q, circuit_ops, layered_30q
from oracle import sv_oracle as O
from tests import emulator as E


def _names(fused):
    return [f["name"] for f in fused]


def _check(n, ops, atol=1e-10):
    fused = fuse_operations(ops, n)
    got = E.apply_fused(O.zero_state(n), n, fused)
    want = O.run_ops(n, ops)
    assert np.max(np.abs(got - want)) < atol
    return fused


def test_layered_workload_fuses_to_structured_layers():
    n = 8
    layers = layered_30q(n, 20)
    ops = circuit_ops(layers, n)
    fused = _check(n, ops)
    names = _names(fused)
    assert names.count("_QFTLAYER") == 2
    assert names.count("_ZZWEIGHTED") == 6
    assert names.count("_XORLAYER") == 6
    assert names.count("_U2LISTLAYER") == 6
    assert "_GATE" not in names


def test_brick_workload():
    n = 7
    ops = circuit_ops(brick_20q(n, 6), n)
    fused = _check(n, ops)
    assert "_DIAGLAYER" in _names(fused)


def test_uniform_layers_and_identity_collapse():
    n = 5
    ops = [{"name": "RX", "wires": [q], "parameters": [0.3]} for q in range(n)]
    assert _names(_check(n, ops)) == ["_RXLAYER"]
    ops = [{"name": "H", "wires": [q], "parameters": []} for q in range(n)]
    assert _names(_check(n, ops)) == ["_U2LAYER"]
    ops = [{"name": "X", "wires": [q], "parameters": []} for q in range(n)]
    assert _names(_check(n, ops)) == ["_XORLAYER"]
    # H.H = I on every wire: the window disappears (device.py:716-726)
    ops = [{"name": "H", "wires": [q], "parameters": []} for q in range(n)] * 2
    assert _names(_check(n, ops)) == []


def test_small_angles_are_not_dropped():
    """RZ(1e-6) must survive the identity collapse (atol 1e-12)."""
    n = 3
    ops = [{"name": "RZ", "wires": [0], "parameters": [1e-6]},
           {"name": "RX", "wires": [1], "parameters": [1e-7]}]
    fused = _check(n, ops)
    assert fused and fused[0]["name"] == "_U2LISTLAYER"


def test_qft_stage_patterns_forward_and_inverse():
    n = 6
    fwd = []
    for j in range(1, 5):
        fwd.append({"name": "H", "wires": [j], "parameters": []})
        for k in range(j + 1, 5):
            fwd.append({"name": "CPHASE", "wires": [k, j], "parameters": [math.pi / 2 ** (k - j)]})
    inv = []
    for j in reversed(range(1, 5)):
        for k in reversed(range(j + 1, 5)):
            inv.append({"name": "CPHASE", "wires": [k, j], "parameters": [-math.pi / 2 ** (k - j)]})
        inv.append({"name": "H", "wires": [j], "parameters": []})
    fused = _check(n, [{"name": "X", "wires": [2], "parameters": []}] + fwd + inv)
    stages = [f for f in fused if f["name"] == "_QFTSTAGE"]
    assert len(stages) >= 5  # the first H merges with the leading X window
    assert sum(1 for f in stages if f["inverse"]) == 3  # j = 4 has no ladder: a lone H


def test_cz_cphase_and_zz_runs():
    n = 6
    ops = [{"name": "CZ", "wires": [0, 1]}, {"name": "CZ", "wires": [2, 3]},
           {"name": "CP", "wires": [4, 5], "parameters": [0.3]}]
    for op in ops:
        op.setdefault("parameters", [])
    fused = _check(n, [{"name": "H", "wires": [q], "parameters": []} for q in range(n)] + ops)
    assert "_DIAGWEIGHTED" in _names(fused)
    zz = [{"name": "ZZPHASE", "wires": [i, (i + 1) % n], "parameters": [0.2]} for i in range(n)]
    fused = _check(n, [{"name": "H", "wires": [q], "parameters": []} for q in range(n)] + zz)
    assert "_ZZLAYER" in _names(fused)


def test_xx_yy_layers():
    n = 5
    h = [{"name": "RY", "wires": [q], "parameters": [0.1 * (q + 1)]} for q in range(n)]
    for name, pat in (("XXPHASE", "_XXLAYER"), ("YYPHASE", "_YYLAYER")):
        ops = h + [{"name": name, "wires": [i, i + 1], "parameters": [0.4]} for i in range(n - 1)]
        assert pat in _names(_check(n, ops))


def test_controlled_groups():
    n = 5
    ops = [{"name": "H", "wires": [q], "parameters": []} for q in range(n)]
    ops += [{"name": "CRX", "wires": [0, 1], "parameters": [0.3]},
            {"name": "CRY", "wires": [2, 3], "parameters": [-0.7]},
            {"name": "CH", "wires": [4, 0], "parameters": []},
            {"name": "CRZ", "wires": [1, 4], "parameters": [1.1]}]
    assert "_CULISTLAYER" in _names(_check(n, ops))


@pytest.mark.parametrize("seed", range(6))
def test_random_circuits_fused_equal_unfused(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 8))
    names = ["H", "X", "Y", "Z", "S", "SDG", "T", "TDG", "SX", "RX", "RY", "RZ", "U1", "U2", "U3",
             "CNOT", "CZ", "CPHASE", "CRX", "CRY", "CRZ", "CH", "SWAP", "ISWAP", "ZZPHASE",
             "XXPHASE", "YYPHASE", "TOFFOLI", "FREDKIN", "I"]
    two = {"CNOT", "CZ", "CPHASE", "CRX", "CRY", "CRZ", "CH", "SWAP", "ISWAP", "ZZPHASE", "XXPHASE",
           "YYPHASE"}
    npar = {"RX": 1, "RY": 1, "RZ": 1, "U1": 1, "U2": 2, "U3": 3, "CPHASE": 1, "CRX": 1, "CRY": 1,
            "CRZ": 1, "ZZPHASE": 1, "XXPHASE": 1, "YYPHASE": 1}
    ops = []
    for _ in range(80):
        nm = names[int(rng.integers(len(names)))]
        ar = 3 if nm in ("TOFFOLI", "FREDKIN") else (2 if nm in two else 1)
        wires = [int(x) for x in rng.choice(n, ar, replace=False)]
        ops.append({"name": nm, "wires": wires, "parameters": rng.uniform(-4, 4, npar.get(nm, 0)).tolist()})
    _check(n, ops)


def test_gate_routing_uses_diagonal_and_controlled_kinds():
    from mettleq_amd import _native as nat
    n = 4
    assert [g.kind for g in gates_for_op({"name": "RZ", "wires": [1], "parameters": [0.2]}, n)] == [nat.MSV_G_DIAG1]
    assert [g.kind for g in gates_for_op({"name": "CZ", "wires": [1, 2]}, n)] == [nat.MSV_G_DIAG2]
    g = gates_for_op({"name": "CNOT", "wires": [3, 0]}, n)
    assert g[0].kind == nat.MSV_G_U and g[0].ncontrols == 1 and g[0].q[0] == 0 and g[0].q[1] == 3
    g = gates_for_op({"name": "TOFFOLI", "wires": [0, 1, 2]}, n)
    assert g[0].ncontrols == 2
    with pytest.raises(ValueError):
        gates_for_op({"name": "CNOT", "wires": [1, 1]}, n)
    with pytest.raises(ValueError):
        gates_for_op({"name": "H", "wires": [9]}, n)


@pytest.mark.parametrize("seed", range(8))
def test_two_qubit_dense_decomposition(seed):
    from mettleq_amd.sim import _two_qubit_gates
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    u, _ = np.linalg.qr(a)
    n = 3
    qa, qb = [int(x) for x in rng.choice(n, 2, replace=False)]
    psi = rng.normal(size=8) + 1j * rng.normal(size=8)
    got = E.apply_gates(psi, n, _two_qubit_gates(u, qa, qb))
    want = O.apply_dense(psi, n, u, [qa, qb])
    assert np.max(np.abs(got - want)) < 1e-10


def test_pauli_pair_gate_decompositions():
    from mettleq_amd.sim import _pauli_pair_gates
    n = 3
    psi = np.random.default_rng(1).normal(size=8) + 0j
    for name in ("XXPHASE", "YYPHASE"):
        got = E.apply_gates(psi, n, _pauli_pair_gates(name, 0.37, 2, 0))
        want = O.apply_dense(psi, n, O.pauli_pair_phase(name, 0.37), [2, 0])
        assert np.max(np.abs(got - want)) < 1e-12


def test_layers_compose_inverse_affine_matches_oracle():
    from mettleq_amd.layers import compose_inverse_affine
    rng = np.random.default_rng(5)
    n = 7
    ops = []
    for _ in range(30):
        k = int(rng.integers(3))
        if k == 0:
            ops.append({"name": "X", "wires": [int(rng.integers(n))]})
        else:
            a, b = (int(x) for x in rng.choice(n, 2, replace=False))
            ops.append({"name": ["CNOT", "SWAP"][k - 1], "wires": [a, b]})
    assert compose_inverse_affine(ops, n) == O.compose_inverse_affine(ops, n)


def _chain_phase_rx_ops(n, phase_name):
    # the reference's case (mlxQMetalKernelsTest.py:153-191): a uniform
    # CPHASE ring or ZZ chain, then RX on every wire
    bonds = [(q, q + 1) for q in range(n - 1)]
    if phase_name == "CPHASE":
        bonds.append((n - 1, 0))
    return ([{"name": phase_name, "wires": list(b), "parameters": [0.23]} for b in bonds]
            + [{"name": "RX", "wires": [w], "parameters": [-0.31]} for w in range(n)])


@pytest.mark.parametrize("phase_name", ["CPHASE", "ZZPHASE"])
def test_chain_phase_rx_amortized_layer(monkeypatch, phase_name):
    """A diagonal chain/ring layer followed by an RX layer fuses into one
    _CHAINPHASE_RXLAYER (device.py:1029-1087), with the reference's masks;
    METTLEQ_CHAINPHASE_RX_FUSION=0 keeps the two layers. Both match the
    gate-by-gate oracle."""
    n = 8
    ops = _chain_phase_rx_ops(n, phase_name)
    monkeypatch.setenv("METTLEQ_CHAINPHASE_RX_FUSION", "1")
    fused = _check(n, ops)
    assert _names(fused) == ["_CHAINPHASE_RXLAYER"]
    f = fused[0]
    assert f["rx_theta"] == -0.31
    if phase_name == "ZZPHASE":
        assert (f["phase_kind"], f["n_bonds"], f["edge_mask"], f["wrap_mask"]) == ("zz", 7, 127, 0)
    else:
        assert (f["phase_kind"], f["n_bonds"], f["edge_mask"], f["wrap_mask"]) == ("cphase", 8, 127, 129)
    monkeypatch.setenv("METTLEQ_CHAINPHASE_RX_FUSION", "0")
    unfused = _check(n, ops)
    assert _names(unfused) == ["_ZZLAYER" if phase_name == "ZZPHASE" else "_DIAGLAYER", "_RXLAYER"]
    # both forms reach the native planner as the same two calls: the diagonal
    # stays deferred and the RX layer's sweep reads it
    from mettleq_amd.dispatch import planned_run
    assert planned_run(fused, n)["kernels"] == planned_run(unfused, n)["kernels"]


def test_chain_phase_rx_not_amortized_for_other_bonds():
    n = 6
    ops = ([{"name": "CZ", "wires": [0, 2]}, {"name": "CZ", "wires": [3, 4]}]
           + [{"name": "RX", "wires": [w], "parameters": [0.4]} for w in range(n)])
    fused = _check(n, ops)
    assert "_CHAINPHASE_RXLAYER" not in _names(fused)
