"""Seeded synthetic workloads of the benchmark configurations.

All draws use ``np.random.default_rng`` (PCG64) from the base seed
20261203; per-item seeds are base + index.

layered_30q(n=30, layers=40): layers 9, 19, 29, 39 are the full QFT
(no terminal swaps, wires 0..n-1); the other 36 cycle
  U3 on every qubit (theta, phi, lambda ~ U[0, 2pi) in (qubit, param) order)
  -> RZZ ring: bonds (i, i+1), i < n-1, plus (n-1, 0); theta ~ U[0, pi) is
     Qiskit rzz(theta) = exp(-i theta/2 ZZ), i.e. ZZPHASE(theta/2)
  -> CNOT chain: control i -> target i+1 for i = 0..n-2 in increasing order.

brick_20q(n=20, layers=10): U3 on every qubit then CZ on (q, q+1) for
q = l%2, l%2+2, ... <= n-2.
"""
from __future__ import annotations

import math
from typing import Dict, List

import numpy as np

BASE_SEED = 20261203


def layered_30q(n: int = 30, n_layers: int = 40, seed: int = BASE_SEED) -> List[Dict]:
    rng = np.random.default_rng(seed)
    layers: List[Dict] = []
    k = 0
    for l in range(n_layers):
        if l % 10 == 9:
            layers.append({"type": "qft"})
            continue
        kind = ("u3", "rzz", "cnot")[k % 3]
        k += 1
        if kind == "u3":
            layers.append({"type": "u3", "angles": rng.uniform(0.0, 2.0 * math.pi, size=(n, 3))})
        elif kind == "rzz":
            layers.append({"type": "rzz", "thetas": rng.uniform(0.0, math.pi, size=n)})
        else:
            layers.append({"type": "cnot"})
    return layers


def brick_20q(n: int = 20, n_layers: int = 10, seed: int = BASE_SEED) -> List[Dict]:
    rng = np.random.default_rng(seed)
    layers: List[Dict] = []
    for l in range(n_layers):
        layers.append({"type": "u3", "angles": rng.uniform(0.0, 2.0 * math.pi, size=(n, 3))})
        layers.append({"type": "cz", "bonds": [(q, q + 1) for q in range(l % 2, n - 1, 2)]})
    return layers


def ring_bonds(n: int):
    bonds = [(i, i + 1) for i in range(n - 1)]
    if n > 2:
        bonds.append((n - 1, 0))
    return bonds


def layer_ops(layer: Dict, n: int) -> List[Dict]:
    """Canonical op dicts of one layer (the gate stream an SDK would send)."""
    t = layer["type"]
    if t == "u3":
        return [{"name": "U3", "wires": [q], "parameters": [float(x) for x in layer["angles"][q]]}
                for q in range(n)]
    if t == "rzz":
        return [{"name": "ZZPHASE", "wires": [a, b], "parameters": [float(th) / 2.0]}
                for (a, b), th in zip(ring_bonds(n), layer["thetas"])]
    if t == "cnot":
        return [{"name": "CNOT", "wires": [i, i + 1], "parameters": []} for i in range(n - 1)]
    if t == "cz":
        return [{"name": "CZ", "wires": [a, b], "parameters": []} for a, b in layer["bonds"]]
    if t == "qft":
        ops = []
        for j in range(n):
            ops.append({"name": "H", "wires": [j], "parameters": []})
            for k in range(j + 1, n):
                ops.append({"name": "CPHASE", "wires": [k, j], "parameters": [math.pi / (1 << (k - j))]})
        return ops
    raise ValueError(t)


def circuit_ops(layers: List[Dict], n: int) -> List[Dict]:
    out: List[Dict] = []
    for layer in layers:
        out.extend(layer_ops(layer, n))
    return out
