"""ctypes front end of oracle/sv_cpu.c (TEST INFRASTRUCTURE / CPU baseline).

``CpuEngine`` replays canonical op dicts the way the reference routes them on
a CPU device (device.py:1138-1290 with Metal disabled, execution.py:167-176):
phase gates through apply_diagonal, controlled gates through
apply_controlled_single, SWAP as an axis swap, everything else through
apply_single, and uniform-angle ZZPHASE runs fused into one diagonal
(device.py:939-958, sim.py:271-300). Only tests, the smoke check and the
CPU-baseline leg of bench.py use it.
"""
from __future__ import annotations

import ctypes
import math
import subprocess
from pathlib import Path
from typing import Iterable, Sequence

import numpy as np

from . import sv_oracle as O

HERE = Path(__file__).resolve().parent
LIB = HERE / "libsv_cpu.so"

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
        u64 = ctypes.c_uint64
        L.cpu_apply_1q.argtypes = [vp, i, i, d]
        L.cpu_apply_diag1.argtypes = [vp, i, i, d]
        L.cpu_apply_diag2.argtypes = [vp, i, i, i, d]
        L.cpu_apply_cu.argtypes = [vp, i, i, i, d]
        L.cpu_apply_swap.argtypes = [vp, i, i, i]
        L.cpu_qft_ladder.argtypes = [vp, i, ctypes.POINTER(ctypes.c_int), i, i, ctypes.c_double]
        L.cpu_block_sums.argtypes = [vp, i, d]
        L.cpu_sample.argtypes = [vp, i, u64, u64, u64, ctypes.POINTER(u64)]
        L.cpu_sample.restype = i
        L.cpu_philox_u01.argtypes = [u64, u64, u64]
        L.cpu_philox_u01.restype = ctypes.c_double
        L.cpu_philox4x32_10.argtypes = [ctypes.POINTER(ctypes.c_uint32)] * 3
        L.cpu_num_threads.restype = i
        L.cpu_set_threads.argtypes = [i]
        _lib = L
    return _lib


def _m8(m) -> np.ndarray:
    a = np.asarray(m, dtype=np.complex128).reshape(-1)
    out = np.empty(2 * a.size)
    out[0::2] = a.real
    out[1::2] = a.imag
    return out


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


_DIAG1 = {"Z", "S", "SDG", "T", "TDG", "RZ", "U1"}
_CTRL = {"CNOT": O.X, "CX": O.X, "CH": O.H}


class CpuEngine:
    """complex64 statevector on the host, gate-by-gate like the reference CPU path."""

    def __init__(self, n: int, state: np.ndarray | None = None):
        self.n = int(n)
        if state is None:
            self.state = np.zeros(1 << self.n, dtype=np.complex64)
            self.state[0] = 1
        else:
            self.state = np.ascontiguousarray(state, dtype=np.complex64).copy()
        self.passes = 0

    @property
    def _p(self):
        return self.state.ctypes.data_as(ctypes.c_void_p)

    def apply_single(self, m, q):
        lib().cpu_apply_1q(self._p, self.n, int(q), _dp(_m8(m)))
        self.passes += 1

    def apply_diag1(self, d, q):
        lib().cpu_apply_diag1(self._p, self.n, int(q), _dp(_m8(d)))
        self.passes += 1

    def apply_diag2(self, d, a, b):
        lib().cpu_apply_diag2(self._p, self.n, int(a), int(b), _dp(_m8(d)))
        self.passes += 1

    def apply_cu(self, m, c, t):
        lib().cpu_apply_cu(self._p, self.n, int(c), int(t), _dp(_m8(m)))
        self.passes += 1

    def apply_swap(self, a, b):
        lib().cpu_apply_swap(self._p, self.n, int(a), int(b))
        self.passes += 1

    def qft_ladder(self, wires: Sequence[int], j: int, sign: float):
        w = (ctypes.c_int * len(wires))(*[int(x) for x in wires])
        lib().cpu_qft_ladder(self._p, self.n, w, len(wires), int(j), float(sign))
        self.passes += 1

    def apply_op(self, op: dict) -> None:
        """Reference routing (device.py:1153-1290)."""
        name = str(op["name"]).upper()
        wires = list(op.get("wires", []))
        params = list(op.get("parameters", []) or [])
        if name == "I":
            return
        if len(wires) == 1:
            if name in _DIAG1:
                self.apply_diag1(np.diag(O.gate_matrix(name, params)), wires[0])
            else:
                self.apply_single(O.gate_matrix(name, params), wires[0])
            return
        if len(wires) == 2:
            a, b = wires
            if name in ("CZ", "CP", "CPHASE", "CRZ", "ZZPHASE"):
                self.apply_diag2(np.diag(O.gate_matrix(name, params)), a, b)
            elif name in _CTRL:
                self.apply_cu(_CTRL[name](), a, b)
            elif name in ("CRX", "CRY"):
                self.apply_cu(O.RX(params[0]) if name == "CRX" else O.RY(params[0]), a, b)
            elif name == "SWAP":
                self.apply_swap(a, b)
            else:
                # XX/YY/ISWAP and other dense pairs: exact dense application
                self.state = O.apply_dense(self.state.astype(np.complex128), self.n,
                                           O.gate_matrix(name, params), wires).astype(np.complex64)
                self.passes += 1
            return
        self.state = O.apply_dense(self.state.astype(np.complex128), self.n,
                                   O.gate_matrix(name, params), wires).astype(np.complex64)
        self.passes += 1

    def run(self, ops: Iterable[dict]) -> np.ndarray:
        ops = list(ops)
        i = 0
        while i < len(ops):
            op = ops[i]
            if str(op["name"]).upper() == "ZZPHASE":
                j = i
                while j < len(ops) and str(ops[j]["name"]).upper() == "ZZPHASE":
                    j += 1
                run = ops[i:j]
                thetas = [r["parameters"][0] for r in run]
                if len(run) >= 2 and all(t == thetas[0] for t in thetas):
                    # one cached phase vector per uniform layer (sim.py:271-300)
                    d = O.zz_layer_diag(self.n, [tuple(r["wires"]) for r in run], thetas)
                    self.state = (self.state * d.astype(np.complex64)).astype(np.complex64)
                    self.passes += 1
                    i = j
                    continue
            self.apply_op(op)
            i += 1
        return self.state

    def qft(self, wires: Sequence[int], inverse: bool = False) -> None:
        """sim.qft / sim.iqft with the fused per-stage ladder (sim.py:522-563)."""
        w = list(wires)
        h = O.H()
        if not inverse:
            for j in range(len(w)):
                self.apply_single(h, w[j])
                self.qft_ladder(w, j, 1.0)
        else:
            for j in reversed(range(len(w))):
                self.qft_ladder(w, j, -1.0)
                self.apply_single(h, w[j])

    def sample_indices(self, shots: int, seed: int, stream_id: int = 0) -> np.ndarray:
        out = np.zeros(max(int(shots), 1), dtype=np.uint64)
        rc = lib().cpu_sample(self._p, self.n, int(shots), int(seed), int(stream_id),
                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        if rc != 0:
            raise ValueError("zero-norm state")
        return out[: int(shots)]

    def block_sums(self) -> np.ndarray:
        nb = ((1 << self.n) + O.BLOCK - 1) // O.BLOCK
        out = np.zeros(nb)
        lib().cpu_block_sums(self._p, self.n, _dp(out))
        return out
