"""Compact circuit encoding for msv_apply_circuit: parallel arrays of gate
codes, wires (3 per op, -1 unused) and angles (3 per op).

Names and aliases follow the reference's canonical op names
(integrations/_common.py:29-115, gates.py:5-234).
"""
from __future__ import annotations

from typing import Dict, Iterable, Tuple

import numpy as np

CODES: Dict[str, int] = {
    "I": 0, "X": 1, "Y": 2, "Z": 3, "H": 4, "S": 5, "SDG": 6, "T": 7, "TDG": 8, "SX": 9,
    "RX": 10, "RY": 11, "RZ": 12, "U1": 13, "U2": 14, "U3": 15,
    "CNOT": 16, "CZ": 17, "CPHASE": 18, "CRX": 19, "CRY": 20, "CRZ": 21, "CH": 22,
    "SWAP": 23, "ISWAP": 24, "ZZPHASE": 25, "XXPHASE": 26, "YYPHASE": 27,
    "TOFFOLI": 28, "FREDKIN": 29,
}
ALIASES = {"CX": "CNOT", "CP": "CPHASE", "P": "U1", "PHASESHIFT": "U1", "SDAG": "SDG", "S†": "SDG",
           "TDAG": "TDG", "T†": "TDG", "CCX": "TOFFOLI", "CSWAP": "FREDKIN", "ID": "I"}
ARITY = {c: (1 if c <= 15 else 2 if c <= 27 else 3) for c in CODES.values()}
N_PARAMS = {"RX": 1, "RY": 1, "RZ": 1, "U1": 1, "U2": 2, "U3": 3, "CPHASE": 1, "CRX": 1, "CRY": 1,
            "CRZ": 1, "ZZPHASE": 1, "XXPHASE": 1, "YYPHASE": 1}


def encode_ops(ops: Iterable[dict]) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """op dicts {"name", "wires", "parameters"} -> (codes, wires, params)."""
    ops = list(ops)
    codes = np.empty(len(ops), dtype=np.int32)
    wires = np.full((len(ops), 3), -1, dtype=np.int32)
    params = np.zeros((len(ops), 3), dtype=np.float64)
    for k, op in enumerate(ops):
        name = str(op["name"]).upper()
        name = ALIASES.get(name, name)
        if name not in CODES:
            raise ValueError(f"Unsupported operation: {op['name']}")
        code = CODES[name]
        w = op.get("wires", [])
        if len(w) != ARITY[code]:
            raise ValueError(f"{name} acts on {ARITY[code]} wire(s), got {len(w)}")
        p = op.get("parameters", []) or []
        if len(p) != N_PARAMS.get(name, 0):
            raise ValueError(f"{name} takes {N_PARAMS.get(name, 0)} parameter(s), got {len(p)}")
        codes[k] = code
        wires[k, : len(w)] = w
        params[k, : len(p)] = p
    return codes, wires.reshape(-1), params.reshape(-1)
