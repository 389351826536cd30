"""Structured-layer entry points with the reference's ``shaders`` signatures.

Each function mirrors one host wrapper of the reference's Metal layer
kernels (paths relative to the reference's ``src/mettleq/shaders``) but acts
in place on a device-resident :class:`NativeState` through the C ABI:

===========================  ==================================  ==========================
reference                    here                                native entry point
===========================  ==================================  ==========================
u2_list_layer_all 512-615    u2_list_layer_all                   msv_apply_1q_layer
u2_layer_all 422-509         u2_layer_all                        msv_apply_1q_layer
rx_layer_all 618-673         rx_layer_all                        msv_apply_1q_layer
hadamard_layer_all 319-387   hadamard_layer_all                  msv_apply_1q_layer
s_phase_layer 390-419        s_phase_layer                       msv_apply_diagonal
diag.py diag_pair_layer      diag_pair_layer                     msv_apply_phase_layer
diag.py diag_weighted_layer  diag_weighted_layer                 msv_apply_phase_layer
zz.py zz_chain_layer         zz_chain_layer                      msv_apply_phase_layer
zz.py zz_weighted_layer      zz_weighted_layer                   msv_apply_phase_layer
xor_affine.py 46-116         compose_inverse_affine, xor_affine_gather  msv_apply_xor_affine
qft.py qft_stage_all         qft_stage_all                       msv_apply_qft
qft.py qft_stage_sub         qft_stage_sub                       msv_apply_qft_stage
controlled_single.py 89-134  controlled_u_groups                 msv_apply_gates
pauli_pair.py 32-64          xx_layer, yy_layer                  1q layer + phase layer
chain_phase.py 91-142        chain_phase_u2_layer_all            phase layer + 1q layer
===========================  ==================================  ==========================

Validation and error types follow the reference (ValueError for invalid
bonds, wires and parameters: diag.py:128-132, zz.py:102-106).
"""
from __future__ import annotations

import math
from typing import Iterable, List, Sequence, Tuple

import numpy as np

from . import _native as nat
from .native_state import NativeState, gate_u

_H = np.array([[1, 1], [1, -1]], dtype=np.complex128) / math.sqrt(2.0)


def _check_bonds(n: int, bonds, what: str) -> None:
    for a, b in bonds:
        if a == b or not (0 <= a < n and 0 <= b < n):
            raise ValueError(f"{what}: invalid bond ({a}, {b})")


def u2_list_layer_all(state: NativeState, n: int, mats, active=None) -> NativeState:
    """Different 2x2 per qubit: ``mats`` is (n, 4) [u00, u01, u10, u11] or
    (n, 2, 2); ``active`` lists the qubits with non-identity matrices."""
    m = np.asarray(mats, dtype=np.complex128).reshape(n, 2, 2)
    state.apply_1q_layer(list(m), active=active)
    return state


def u2_layer_all(state: NativeState, n: int, u2x2) -> NativeState:
    """The same 2x2 on every qubit."""
    u = np.asarray(u2x2, dtype=np.complex128).reshape(2, 2)
    state.apply_1q_layer([u] * n)
    return state


def rx_layer_all(state: NativeState, n: int, theta: float) -> NativeState:
    c, s = math.cos(theta / 2.0), math.sin(theta / 2.0)
    u = np.array([[c, -1j * s], [-1j * s, c]], dtype=np.complex128)
    return u2_layer_all(state, n, u)


def hadamard_layer_all(state: NativeState, n: int) -> NativeState:
    return u2_layer_all(state, n, _H)


def s_phase_layer(state: NativeState, n: int, dagger: bool = False) -> NativeState:
    ph = -1j if dagger else 1j
    state.apply_diagonal(one=[(q, 1.0, ph) for q in range(n)])
    return state


def diag_pair_layer(state: NativeState, theta: float, n: int, bonds) -> NativeState:
    """Equal-angle CPHASE(theta) over ``bonds`` (CZ when theta = pi)."""
    bonds = [tuple(b) for b in bonds]
    _check_bonds(n, bonds, "diag_pair_layer")
    if bonds:
        state.apply_phase_layer([nat.MSV_PHASE_CPHASE] * len(bonds), bonds, [theta] * len(bonds))
    return state


def diag_weighted_layer(state: NativeState, n: int, bonds, thetas) -> NativeState:
    bonds = [tuple(b) for b in bonds]
    if len(bonds) != len(thetas):
        raise ValueError("diag_weighted_layer: bonds/thetas length mismatch")
    _check_bonds(n, bonds, "diag_weighted_layer")
    if bonds:
        state.apply_phase_layer([nat.MSV_PHASE_CPHASE] * len(bonds), bonds, list(thetas))
    return state


def zz_chain_layer(state: NativeState, theta: float, n: int, bonds) -> NativeState:
    """exp(-i*theta*sum Z_a Z_b) over ``bonds`` (one angle)."""
    bonds = [tuple(b) for b in bonds]
    _check_bonds(n, bonds, "zz_chain_layer")
    if bonds:
        state.apply_phase_layer([nat.MSV_PHASE_ZZ] * len(bonds), bonds, [theta] * len(bonds))
    return state


def zz_weighted_layer(state: NativeState, n: int, bonds, thetas) -> NativeState:
    bonds = [tuple(b) for b in bonds]
    if len(bonds) != len(thetas):
        raise ValueError("zz_weighted_layer: bonds/thetas length mismatch")
    _check_bonds(n, bonds, "zz_weighted_layer")
    if bonds:
        state.apply_phase_layer([nat.MSV_PHASE_ZZ] * len(bonds), bonds, list(thetas))
    return state


def compose_inverse_affine(ops: Iterable[dict], n: int) -> Tuple[List[int], int]:
    """Inverse GF(2) affine map (rows, c) of an X/CNOT/SWAP run.

    X, CNOT and SWAP send basis index x to Mx ^ c, so the run is a pure
    amplitude permutation out[y] = state[f^{-1}(y)]. Each gate is an
    involution, so f^{-1} is the run replayed in reverse onto the identity
    map (the construction of xor_affine.py:46-76); rows[k] is the y-mask
    whose parity is source qubit k.
    """
    rows = [1 << (n - 1 - k) for k in range(n)]
    c = 0
    for op in reversed(list(ops)):
        name = str(op.get("name", "")).upper()
        w = list(op.get("wires", []))
        if name == "X":
            c ^= 1 << (n - 1 - w[0])
        elif name in ("CNOT", "CX"):
            a, b = w
            rows[b] ^= rows[a]
            if (c >> (n - 1 - a)) & 1:
                c ^= 1 << (n - 1 - b)
        elif name == "SWAP":
            a, b = w
            rows[a], rows[b] = rows[b], rows[a]
            if ((c >> (n - 1 - a)) & 1) != ((c >> (n - 1 - b)) & 1):
                c ^= (1 << (n - 1 - a)) | (1 << (n - 1 - b))
        else:
            raise ValueError(f"Not a GF(2) affine gate: {name}")
    return rows, c


def xor_affine_gather(state: NativeState, n: int, rows: Sequence[int], c: int) -> NativeState:
    state.apply_xor_affine(rows, c)
    return state


def qft_stage_all(state: NativeState, n: int) -> NativeState:
    """Full-register QFT without terminal swaps."""
    state.apply_qft(0, n, inverse=False)
    return state


def qft_stage_sub(state: NativeState, n: int, j: int, m: int, inverse: bool = False) -> NativeState:
    if not (0 <= j < m <= n):
        raise ValueError(f"qft_stage_sub: invalid j={j}, m={m}, n={n}")
    state.apply_qft_stage(j, m, inverse=inverse)
    return state


def controlled_u_groups(state: NativeState, n: int, groups) -> NativeState:
    """Groups of one or two controlled 2x2 gates {control, target, matrix}."""
    gates = []
    for group in groups:
        if len(group) not in (1, 2):
            raise ValueError("controlled-unitary groups must contain one or two gates")
        for g in group:
            m = np.asarray(g["matrix"], dtype=np.complex128).reshape(2, 2)
            gates.append(gate_u(int(g["target"]), m, [int(g["control"])]))
    state.apply_gates(gates)
    return state


def xx_layer(state: NativeState, theta: float, n: int, bonds) -> NativeState:
    """exp(-i*theta*sum X_a X_b) = H^n . ZZ-layer . H^n (X = H Z H)."""
    hadamard_layer_all(state, n)
    zz_chain_layer(state, theta, n, bonds)
    return hadamard_layer_all(state, n)


def yy_layer(state: NativeState, theta: float, n: int, bonds) -> NativeState:
    """exp(-i*theta*sum Y_a Y_b) with V = S H (V Z V^dag = Y): apply V^dag =
    H Sdag, the ZZ layer, then V (H first, then S). The per-qubit H Sdag and
    S H products run as single 1q-layer sweeps."""
    sdg = np.diag([1, -1j])
    s = np.diag([1, 1j])
    state.apply_1q_layer([_H @ sdg] * n)
    zz_chain_layer(state, theta, n, bonds)
    state.apply_1q_layer([s @ _H] * n)
    return state


def chain_phase_u2_layer_all(state: NativeState, n: int, u2x2, *, kind: str, theta: float,
                             bonds) -> NativeState:
    """A nearest-neighbour ZZ or CPHASE layer followed by the same 2x2 on every
    qubit (chain_phase.py:91-142)."""
    if kind == "zz":
        zz_chain_layer(state, theta, n, bonds)
    elif kind == "cphase":
        diag_pair_layer(state, theta, n, bonds)
    else:
        raise ValueError("kind must be 'zz' or 'cphase'")
    return u2_layer_all(state, n, u2x2)
