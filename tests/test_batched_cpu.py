"""The batched estimator path's circuit matching and parameter mapping (CPU):
which circuits reach the batched engine, and how circuit parameters map to
engine slots and back (mettleq_amd.batched)."""
import math

import numpy as np

from mettleq_amd.batched import ParamRef, ansatz_structure, z_terms
from oracle import sv_oracle as O


def _hea_ops(n, layers, offset=0.0):
    ops, p = [], 0
    for _ in range(layers):
        for q in range(n):
            ops.append({"name": "RY", "wires": [q], "parameters": [ParamRef(p)]})
            ops.append({"name": "RZ", "wires": [q], "parameters": [ParamRef(p + 1, 2.0, offset)]})
            p += 2
        for q in range(n):
            ops.append({"name": "CZ", "wires": [q, (q + 1) % n], "parameters": []})
    return ops, p


def test_rotation_ansatz_matches_and_maps_parameters():
    ops, P = _hea_ops(4, 2, offset=0.25)
    ops.append({"name": "RX", "wires": [1], "parameters": [0.7]})  # constant angle
    ops.append({"name": "ZZPHASE", "wires": [0, 2], "parameters": [0.3]})
    st = ansatz_structure(4, ops)
    assert st is not None and st.n_params == P
    assert len(st.slot_index) == P + 1
    vals = np.random.default_rng(0).uniform(-3, 3, (5, P))
    slots = st.slot_values(vals)
    assert np.allclose(slots[:, 0], vals[:, 0])
    assert np.allclose(slots[:, 1], 2.0 * vals[:, 1] + 0.25)
    assert np.allclose(slots[:, -1], 0.7)
    # gradients: d/dp = sum over slots of coeff * d/dslot
    g = np.arange(len(st.slot_index), dtype=float)[None, :].repeat(5, 0)
    pg = st.param_grads(g)
    assert np.allclose(pg[:, 0], 0.0) and np.allclose(pg[:, 1], 2.0 * 1.0)
    phases = [x for x in st.gates if x[0] in ("CPHASE", "ZZPHASE")]
    assert ("CPHASE", 0, 1, math.pi) in phases and ("ZZPHASE", 0, 2, 0.3) in phases


def test_energy_of_the_mapped_structure_equals_the_circuit():
    """The oracle energy of the engine gate list at the mapped slot values
    equals the oracle energy of the original circuit with the parameters
    substituted (the mapping loses nothing)."""
    n = 4
    ops, P = _hea_ops(n, 2, offset=-0.4)
    st = ansatz_structure(n, ops)
    vals = np.random.default_rng(1).uniform(-3, 3, P)
    slots = st.slot_values(vals)[0]
    bound = []
    for op in ops:
        ps = [(p.coeff * vals[p.index] + p.offset) if isinstance(p, ParamRef) else p for p in op["parameters"]]
        bound.append({**op, "parameters": ps})
    terms = [((q, (q + 1) % n), 1.0) for q in range(n)]
    psi = O.run_ops(n, bound)
    want = float(np.real(sum(w * O.pauli_expectation(psi, n, {a: "Z", b: "Z"}) for (a, b), w in terms)))
    assert abs(O.ansatz_energy(n, st.gates, slots, terms) - want) < 1e-12


def test_circuits_outside_the_form_stay_on_the_per_binding_path():
    ref = [{"name": "RX", "wires": [0], "parameters": [ParamRef(0)]}]
    assert ansatz_structure(2, ref + [{"name": "CNOT", "wires": [0, 1], "parameters": []}]) is None
    assert ansatz_structure(2, ref + [{"name": "H", "wires": [1], "parameters": []}]) is None
    assert ansatz_structure(2, [{"name": "CPHASE", "wires": [0, 1], "parameters": [ParamRef(0)]}]) is None
    assert ansatz_structure(2, [{"name": "CZ", "wires": [0, 1], "parameters": []}]) is None  # nothing to bind


def test_observable_matching():
    terms, const = z_terms([{0: "Z", 2: "Z"}, {1: "Z"}, {}], [0.5, -1.0 + 0j, 0.25])
    assert terms == [((0, 2), 0.5), ((1,), -1.0)] and const == 0.25
    assert z_terms([{0: "X"}], [1.0]) is None
    assert z_terms([{0: "Z"}], [1j]) is None
