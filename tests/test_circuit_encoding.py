"""Compact circuit encoding (CPU) and native circuit ingestion (GPU) against
the oracle and the Python gate translation."""
import math

import numpy as np
import pytest

from mettleq_amd.circuit import CODES, encode_ops
from oracle import sv_oracle as O

ALL_OPS = [
    ("I", 1, 0), ("X", 1, 0), ("Y", 1, 0), ("Z", 1, 0), ("H", 1, 0), ("S", 1, 0), ("SDG", 1, 0),
    ("T", 1, 0), ("TDG", 1, 0), ("SX", 1, 0), ("RX", 1, 1), ("RY", 1, 1), ("RZ", 1, 1), ("U1", 1, 1),
    ("U2", 1, 2), ("U3", 1, 3), ("CNOT", 2, 0), ("CZ", 2, 0), ("CPHASE", 2, 1), ("CRX", 2, 1),
    ("CRY", 2, 1), ("CRZ", 2, 1), ("CH", 2, 0), ("SWAP", 2, 0), ("ISWAP", 2, 0), ("ZZPHASE", 2, 1),
    ("XXPHASE", 2, 1), ("YYPHASE", 2, 1), ("TOFFOLI", 3, 0), ("FREDKIN", 3, 0),
]


def _random_ops(n, count, seed):
    rng = np.random.default_rng(seed)
    ops = []
    for _ in range(count):
        name, ar, npar = ALL_OPS[int(rng.integers(len(ALL_OPS)))]
        w = [int(x) for x in rng.choice(n, size=ar, replace=False)]
        ops.append({"name": name, "wires": w, "parameters": [float(x) for x in rng.uniform(-3, 3, npar)]})
    return ops


def test_encoding_layout_and_aliases():
    codes, wires, params = encode_ops([
        {"name": "cx", "wires": [0, 2], "parameters": []},
        {"name": "U3", "wires": [1], "parameters": [0.1, 0.2, 0.3]},
        {"name": "ccx", "wires": [0, 1, 2], "parameters": []},
        {"name": "P", "wires": [2], "parameters": [0.5]},
    ])
    assert list(codes) == [CODES["CNOT"], CODES["U3"], CODES["TOFFOLI"], CODES["U1"]]
    assert list(wires) == [0, 2, -1, 1, -1, -1, 0, 1, 2, 2, -1, -1]
    assert list(params[3:6]) == [0.1, 0.2, 0.3]
    assert len(CODES) == 30


@pytest.mark.parametrize("op", [
    {"name": "FOO", "wires": [0], "parameters": []},
    {"name": "CNOT", "wires": [0], "parameters": []},
    {"name": "RX", "wires": [0], "parameters": []},
])
def test_encoding_rejects_bad_ops(op):
    with pytest.raises(ValueError):
        encode_ops([op])


@pytest.mark.gpu
@pytest.mark.parametrize("n,count,seed", [(3, 60, 1), (7, 200, 2), (14, 150, 3)])
def test_native_circuit_matches_oracle_and_python_translation(gpu, n, count, seed):
    from mettleq_amd.fusion import gates_for_op
    from mettleq_amd.native_state import NativeState
    ops = _random_ops(n, count, seed)
    want = O.run_ops(n, ops)
    with NativeState(n) as a, NativeState(n) as b:
        a.apply_circuit(*encode_ops(ops))
        gates = []
        for op in ops:
            gates.extend(gates_for_op(op, n))
        b.apply_gates(gates)
        ga, gb = a.to_numpy(), b.to_numpy()
    assert np.max(np.abs(ga - want)) < 5e-6
    assert np.max(np.abs(ga - gb)) < 1e-6


@pytest.mark.gpu
def test_native_circuit_rejects_bad_input(gpu):
    from mettleq_amd.native_state import NativeState
    with NativeState(3) as s:
        with pytest.raises(ValueError):
            s.apply_circuit(np.array([99]), np.array([0, -1, -1]), np.zeros(3))
        with pytest.raises(ValueError):
            s.apply_circuit(np.array([CODES["CNOT"]]), np.array([1, 1, -1]), np.zeros(3))
        with pytest.raises(ValueError):
            s.apply_circuit(np.array([CODES["RX"]]), np.array([5, -1, -1]), np.zeros(3))
        with pytest.raises(ValueError):
            s.apply_circuit(np.array([CODES["RX"]]), np.array([0, -1, -1]), np.array([math.nan, 0, 0]))
