"""Gate matrices in the reference engine's conventions (gates.py:5-234,
device.py:32-70), as complex128 NumPy arrays built on the host.

Conventions that differ from Qiskit/PennyLane and must be kept:
  U3(t, p, l) = RZ(p) . RY(t) . RZ(l)   (gates.py:85-87; Qiskit U times
                                         exp(-i(p+l)/2), a global phase)
  U2(p, l)    = RZ(p) . RY(pi/2) . RZ(l) (gates.py:80-82)
  RZ(t)       = diag(e^{-it/2}, e^{it/2}) (gates.py:64-68)
  ZZ/XX/YYPHASE(t) = exp(-i t P(x)P)     (sim.py:252-264): SDK angles are
                                         halved on entry
"""
from __future__ import annotations

import math
from typing import Sequence

import numpy as np

C = np.complex128


def _rz(t):
    h = t / 2.0
    return np.array([[complex(math.cos(-h), math.sin(-h)), 0], [0, complex(math.cos(h), math.sin(h))]], dtype=C)


def _ry(t):
    c, s = math.cos(t / 2.0), math.sin(t / 2.0)
    return np.array([[c, -s], [s, c]], dtype=C)


def _rx(t):
    c, s = math.cos(t / 2.0), math.sin(t / 2.0)
    return np.array([[c, -1j * s], [-1j * s, c]], dtype=C)


def _phase(p):
    return np.array([[1, 0], [0, complex(math.cos(p), math.sin(p))]], dtype=C)


_SQ = 1.0 / math.sqrt(2.0)
_T = complex(math.cos(math.pi / 4.0), math.sin(math.pi / 4.0))
_TD = complex(math.cos(math.pi / 4.0), -math.sin(math.pi / 4.0))

CONST_1Q = {
    "I": np.eye(2, dtype=C),
    "X": np.array([[0, 1], [1, 0]], dtype=C),
    "Y": np.array([[0, -1j], [1j, 0]], dtype=C),
    "Z": np.array([[1, 0], [0, -1]], dtype=C),
    "H": np.array([[_SQ, _SQ], [_SQ, -_SQ]], dtype=C),
    "S": np.array([[1, 0], [0, 1j]], dtype=C),
    "SDG": np.array([[1, 0], [0, -1j]], dtype=C),
    "T": np.array([[1, 0], [0, _T]], dtype=C),
    "TDG": np.array([[1, 0], [0, _TD]], dtype=C),
    "SX": np.array([[complex(0.5, 0.5), complex(0.5, -0.5)], [complex(0.5, -0.5), complex(0.5, 0.5)]], dtype=C),
}

# one-qubit gates whose matrix is diagonal (routed to the phase path,
# device.py:1153-1170)
DIAGONAL_1Q = {"Z", "S", "SDG", "T", "TDG", "RZ", "U1"}
PARAM_1Q = {"RX": 1, "RY": 1, "RZ": 1, "U1": 1, "U2": 2, "U3": 3}


def single_qubit_matrix(name: str, params: Sequence[float] = ()) -> np.ndarray:
    name = name.upper()
    if name in CONST_1Q:
        return CONST_1Q[name]
    p = list(params)
    if name == "RX":
        return _rx(p[0])
    if name == "RY":
        return _ry(p[0])
    if name == "RZ":
        return _rz(p[0])
    if name == "U1":
        return _phase(p[0])
    if name == "U2":
        return _rz(p[0]) @ _ry(math.pi / 2.0) @ _rz(p[1])
    if name == "U3":
        return _rz(p[1]) @ _ry(p[0]) @ _rz(p[2])
    raise ValueError(f"Unsupported single-qubit op: {name}")


def cphase_entry(theta: float) -> complex:
    """exp(i*theta), exactly -1 for theta = pi (diag.py:89-98)."""
    if abs(theta - math.pi) < 1e-12:
        return -1.0 + 0.0j
    return complex(math.cos(theta), math.sin(theta))


def two_qubit_diagonal(name: str, params: Sequence[float] = ()):
    """Diagonal of a diagonal two-qubit gate (first wire = MSB), else None."""
    name = name.upper()
    if name == "CZ":
        return (1, 1, 1, -1)
    if name in ("CP", "CPHASE"):
        return (1, 1, 1, complex(math.cos(params[0]), math.sin(params[0])))
    if name == "CRZ":
        d = np.diag(_rz(params[0]))
        return (1, 1, d[0], d[1])
    if name == "ZZPHASE":
        t = params[0]
        em, ep = complex(math.cos(t), -math.sin(t)), complex(math.cos(t), math.sin(t))
        return (em, ep, ep, em)
    return None


CONTROLLED_TARGETS = {"CNOT": "X", "CX": "X", "CH": "H"}


def controlled_target_matrix(name: str, params: Sequence[float] = ()):
    name = name.upper()
    if name in CONTROLLED_TARGETS:
        return CONST_1Q[CONTROLLED_TARGETS[name]]
    if name == "CRX":
        return _rx(params[0])
    if name == "CRY":
        return _ry(params[0])
    if name == "CRZ":
        return _rz(params[0])
    return None


def pauli_pair_phase(name: str, theta: float) -> np.ndarray:
    """Dense exp(-i*theta*P(x)P) (device.py:53-70)."""
    c = math.cos(theta)
    ms = -1j * math.sin(theta)
    z = 0
    if name == "ZZPHASE":
        return np.diag(two_qubit_diagonal("ZZPHASE", [theta])).astype(C)
    if name == "XXPHASE":
        return np.array([[c, z, z, ms], [z, c, ms, z], [z, ms, c, z], [ms, z, z, c]], dtype=C)
    if name == "YYPHASE":
        return np.array([[c, z, z, -ms], [z, c, ms, z], [z, ms, c, z], [-ms, z, z, c]], dtype=C)
    raise ValueError(f"Unknown Pauli-pair phase op: {name}")


# exact-identity collapse threshold of fused windows (device.py:716-733)
IDENTITY_ATOL = 1e-12
