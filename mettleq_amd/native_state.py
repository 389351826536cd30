"""Python handle of one device-resident statevector (C ABI ``msv_state``).

Methods map one-to-one onto the entry points of include/mettleq_sv.h; array
arguments are converted to the C layouts here. Conventions follow the
reference: qubit q is index bit n-1-q (sim.py:10-14).
"""
from __future__ import annotations

import ctypes
from typing import Iterable, Optional, Sequence

import numpy as np

from . import _native as nat

DTYPES = {"complex64": nat.MSV_C64, "complex128": nat.MSV_C128}
NP_DTYPES = {nat.MSV_C64: np.complex64, nat.MSV_C128: np.complex128}


def _mat8(m) -> np.ndarray:
    a = np.asarray(m, dtype=np.complex128).reshape(2, 2)
    out = np.empty(8, dtype=np.float64)
    out[0::2] = a.reshape(4).real
    out[1::2] = a.reshape(4).imag
    return out


class NativeState:
    """An n-qubit amplitude buffer on one GPU, initialised to |0...0>."""

    def __init__(self, n_qubits: int, dtype: str = "complex64", device: int = 0, virtual: bool = False):
        if dtype not in DTYPES:
            raise ValueError(f"dtype must be one of {sorted(DTYPES)}")
        self.n = int(n_qubits)
        self.dtype = dtype
        self.code = DTYPES[dtype]
        self.device = int(device)
        self.virtual = bool(virtual)
        self._h = ctypes.c_void_p()
        create = nat.lib().msv_state_create_virtual if self.virtual else nat.lib().msv_state_create
        nat.check(create(self.n, self.code, self.device, ctypes.byref(self._h)), "state_create")

    @classmethod
    def planner(cls, n_qubits: int, dtype: str = "complex64", device: int = 0) -> "NativeState":
        """A virtual state (msv_state_create_virtual): the native planner of
        an n-qubit state with no device memory; layer calls plan, defer and
        log their kernels without launching them."""
        return cls(n_qubits, dtype, device, virtual=True)

    def dispatch_log(self, reset: bool = False) -> list:
        """Kernel kinds launched since the last reset, in order (nat.KERNEL_KINDS names)."""
        cap = 1 << 16
        buf = np.zeros(cap, dtype=np.int32)
        total = ctypes.c_uint64(0)
        nat.check(nat.lib().msv_state_dispatch_log(self.handle, nat.as_ptr(buf, ctypes.c_int32), cap,
                                                   ctypes.byref(total), int(bool(reset))), "dispatch_log")
        m = min(int(total.value), cap)
        return [nat.KERNEL_KINDS.get(int(k), f"kind{int(k)}") for k in buf[:m]]

    # -- lifecycle ---------------------------------------------------------
    @property
    def handle(self) -> ctypes.c_void_p:
        if not self._h:
            raise ValueError("state has been closed")
        return self._h

    def close(self) -> None:
        if self._h:
            nat.lib().msv_state_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):  # pragma: no cover - interpreter shutdown ordering
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def nbytes(self) -> int:
        return (1 << self.n) * (8 if self.code == nat.MSV_C64 else 16)

    def reset(self) -> None:
        nat.check(nat.lib().msv_state_reset(self.handle), "state_reset")

    def set_basis(self, index: int) -> None:
        nat.check(nat.lib().msv_state_set_basis(self.handle, int(index)), "set_basis")

    def sync(self) -> None:
        nat.check(nat.lib().msv_state_sync(self.handle), "sync")

    def release_scratch(self) -> None:
        nat.check(nat.lib().msv_state_release_scratch(self.handle), "release_scratch")

    def device_ptr(self) -> int:
        return int(nat.lib().msv_state_device_ptr(self.handle) or 0)

    def sweep_count(self, reset: bool = False) -> int:
        out = ctypes.c_uint64(0)
        nat.check(nat.lib().msv_state_sweep_count(self.handle, ctypes.byref(out), int(reset)), "sweeps")
        return int(out.value)

    def pending(self) -> int:
        """Bit mask of deferred work (nat.MSV_PENDING_*)."""
        out = ctypes.c_int(0)
        nat.check(nat.lib().msv_state_pending(self.handle, ctypes.byref(out)), "pending")
        return int(out.value)

    def timer_start(self) -> None:
        nat.check(nat.lib().msv_state_timer_start(self.handle), "timer_start")

    def timer_stop(self) -> float:
        out = ctypes.c_double(0.0)
        nat.check(nat.lib().msv_state_timer_stop(self.handle, ctypes.byref(out)), "timer_stop")
        return float(out.value)

    # -- data movement ----------------------------------------------------
    def to_numpy(self, offset: int = 0, count: Optional[int] = None) -> np.ndarray:
        count = (1 << self.n) - offset if count is None else int(count)
        out = np.empty(count, dtype=NP_DTYPES[self.code])
        nat.check(nat.lib().msv_state_download(self.handle, out.ctypes.data_as(ctypes.c_void_p),
                                               int(offset), int(count)), "download")
        return out

    def from_numpy(self, amps: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(amps, dtype=NP_DTYPES[self.code])
        nat.check(nat.lib().msv_state_upload(self.handle, a.ctypes.data_as(ctypes.c_void_p),
                                             int(offset), int(a.size)), "upload")

    def copy_from(self, other: "NativeState") -> None:
        nat.check(nat.lib().msv_state_copy(self.handle, other.handle), "state_copy")

    def scale(self, factor: complex) -> None:
        f = complex(factor)
        nat.check(nat.lib().msv_state_scale(self.handle, f.real, f.imag), "scale")

    # -- layers -------------------------------------------------------------
    def apply_1q_layer(self, mats: Sequence, active: Optional[Iterable[int]] = None) -> None:
        """mats[q] (2x2) acts on qubit q; qubits outside ``active`` are skipped."""
        if len(mats) != self.n:
            raise ValueError("need one 2x2 matrix per qubit")
        buf = np.concatenate([_mat8(m) for m in mats])
        act_ptr = None
        if active is not None:
            act = np.zeros(self.n, dtype=np.uint8)
            for q in active:
                act[int(q)] = 1
            act_ptr = nat.as_ptr(act, ctypes.c_uint8)
        nat.check(nat.lib().msv_apply_1q_layer(self.handle, nat.as_ptr(buf, ctypes.c_double), act_ptr),
                  "apply_1q_layer")

    def apply_phase_layer(self, kinds: Sequence[int], bonds: Sequence[Sequence[int]],
                          thetas: Sequence[float]) -> None:
        k = np.ascontiguousarray(kinds, dtype=np.int32)
        b = np.ascontiguousarray(np.asarray(bonds, dtype=np.int32).reshape(-1))
        t = np.ascontiguousarray(thetas, dtype=np.float64)
        if not (len(k) == len(t) == len(b) // 2):
            raise ValueError("phase_layer: bonds/thetas length mismatch")
        nat.check(nat.lib().msv_apply_phase_layer(
            self.handle, nat.as_ptr(k, ctypes.c_int32), nat.as_ptr(b, ctypes.c_int32),
            nat.as_ptr(t, ctypes.c_double), int(len(k))), "apply_phase_layer")

    def apply_diagonal(self, one=(), two=()) -> None:
        """one: [(q, d0, d1)], two: [(a, b, d00, d01, d10, d11)] with a the MSB."""
        q1 = np.asarray([int(t[0]) for t in one], dtype=np.int32)
        d1 = np.asarray([[complex(v).real if i % 2 == 0 else complex(v).imag
                          for v in t[1:3] for i in (0, 1)] for t in one], dtype=np.float64).reshape(-1)
        q2 = np.asarray([[int(t[0]), int(t[1])] for t in two], dtype=np.int32).reshape(-1)
        d2 = np.asarray([[complex(v).real if i % 2 == 0 else complex(v).imag
                          for v in t[2:6] for i in (0, 1)] for t in two], dtype=np.float64).reshape(-1)
        nat.check(nat.lib().msv_apply_diagonal(
            self.handle, nat.as_ptr(q1, ctypes.c_int32), nat.as_ptr(d1, ctypes.c_double), len(one),
            nat.as_ptr(q2, ctypes.c_int32), nat.as_ptr(d2, ctypes.c_double), len(two)),
            "apply_diagonal")

    def apply_xor_affine(self, rows: Sequence[int], c: int) -> None:
        r = np.ascontiguousarray([int(x) for x in rows], dtype=np.uint64)
        if len(r) != self.n:
            raise ValueError("xor_affine: need one row per qubit")
        nat.check(nat.lib().msv_apply_xor_affine(self.handle, nat.as_ptr(r, ctypes.c_uint64), int(c)),
                  "apply_xor_affine")

    def apply_qft(self, first: int = 0, m: Optional[int] = None, inverse: bool = False) -> None:
        m = self.n if m is None else int(m)
        nat.check(nat.lib().msv_apply_qft(self.handle, int(first), m, int(bool(inverse))), "apply_qft")

    def apply_qft_stage(self, j: int, m: int, inverse: bool = False) -> None:
        nat.check(nat.lib().msv_apply_qft_stage(self.handle, int(j), int(m), int(bool(inverse))),
                  "apply_qft_stage")

    def apply_circuit(self, codes: np.ndarray, wires: np.ndarray, params: np.ndarray) -> None:
        """Compact circuit (mettleq_amd.circuit.encode_ops): gate matrices are
        built and fused natively."""
        c = np.ascontiguousarray(codes, dtype=np.int32)
        w = np.ascontiguousarray(wires, dtype=np.int32)
        p = np.ascontiguousarray(params, dtype=np.float64)
        if w.size != 3 * c.size or p.size != 3 * c.size:
            raise ValueError("circuit arrays must hold 3 wires and 3 angles per op")
        nat.check(nat.lib().msv_apply_circuit(self.handle, nat.as_ptr(c, ctypes.c_int32),
                                              nat.as_ptr(w, ctypes.c_int32), nat.as_ptr(p, ctypes.c_double),
                                              int(c.size)), "apply_circuit")

    def apply_gates(self, gates: Sequence[nat.msv_gate]) -> None:
        if not gates:
            return
        arr = (nat.msv_gate * len(gates))(*gates)
        nat.check(nat.lib().msv_apply_gates(self.handle, arr, len(gates)), "apply_gates")

    # -- reductions ------------------------------------------------------------
    def probabilities(self, wires: Optional[Sequence[int]] = None) -> np.ndarray:
        if wires is None:
            out = np.empty(1 << self.n, dtype=np.float64)
            nat.check(nat.lib().msv_probabilities(self.handle, None, self.n,
                                                  nat.as_ptr(out, ctypes.c_double)), "probabilities")
            return out
        w = np.ascontiguousarray([int(x) for x in wires], dtype=np.int32)
        out = np.empty(1 << len(w), dtype=np.float64)
        nat.check(nat.lib().msv_probabilities(self.handle, nat.as_ptr(w, ctypes.c_int32), len(w),
                                              nat.as_ptr(out, ctypes.c_double)), "probabilities")
        return out

    def sample(self, shots: int, wires: Optional[Sequence[int]] = None, seed: int = 0,
               stream_id: int = 0) -> np.ndarray:
        """Packed outcomes (first wire = MSB), one uint64 per shot."""
        shots = int(shots)
        if shots < 0:
            raise ValueError("shots must be non-negative")
        out = np.zeros(max(shots, 1), dtype=np.uint64)
        if wires is None:
            wp, nw = None, self.n
        else:
            w = np.ascontiguousarray([int(x) for x in wires], dtype=np.int32)
            wp, nw = nat.as_ptr(w, ctypes.c_int32), len(w)
        nat.check(nat.lib().msv_sample(self.handle, wp, nw, shots, int(seed) & ((1 << 64) - 1),
                                       int(stream_id) & ((1 << 64) - 1),
                                       nat.as_ptr(out, ctypes.c_uint64)), "sample")
        return out[:shots]

    def expval_pauli(self, words: Sequence[dict]) -> np.ndarray:
        """<P> for words {qubit: 'X'|'Y'|'Z'|'I'} (complex; real for Hermitian P)."""
        xs, zs = [], []
        for word in words:
            x = z = 0
            for q, p in word.items():
                p = str(p).upper()
                q = int(q)
                if not (0 <= q < self.n):
                    raise ValueError("Observable wire is outside the device")
                if p in ("X", "Y"):
                    x |= 1 << q
                if p in ("Z", "Y"):
                    z |= 1 << q
                if p not in ("X", "Y", "Z", "I"):
                    raise ValueError("Pauli products support only I, X, Y, and Z")
            xs.append(x)
            zs.append(z)
        if not xs:
            return np.zeros(0, dtype=np.complex128)
        xa = np.ascontiguousarray(xs, dtype=np.uint64)
        za = np.ascontiguousarray(zs, dtype=np.uint64)
        re = np.zeros(len(xs))
        im = np.zeros(len(xs))
        nat.check(nat.lib().msv_expval_pauli(self.handle, nat.as_ptr(xa, ctypes.c_uint64),
                                             nat.as_ptr(za, ctypes.c_uint64), len(xs),
                                             nat.as_ptr(re, ctypes.c_double),
                                             nat.as_ptr(im, ctypes.c_double)), "expval_pauli")
        return re + 1j * im

    def inner(self, other: "NativeState") -> complex:
        re = ctypes.c_double(0.0)
        im = ctypes.c_double(0.0)
        nat.check(nat.lib().msv_inner(self.handle, other.handle, ctypes.byref(re), ctypes.byref(im)),
                  "inner")
        return complex(re.value, im.value)

    def norm2(self) -> float:
        out = ctypes.c_double(0.0)
        nat.check(nat.lib().msv_norm2(self.handle, ctypes.byref(out)), "norm2")
        return float(out.value)


def make_gate(kind: int, qubits: Sequence[int], m=None, ncontrols: int = 0) -> nat.msv_gate:
    g = nat.msv_gate()
    g.kind = int(kind)
    g.ncontrols = int(ncontrols)
    for i, q in enumerate(qubits):
        g.q[i] = int(q)
    if m is not None:
        vals = np.asarray(m, dtype=np.complex128).reshape(-1)
        for i, v in enumerate(vals[:4]):
            g.m[2 * i] = float(v.real)
            g.m[2 * i + 1] = float(v.imag)
    return g


def gate_u(target: int, u, controls: Sequence[int] = ()) -> nat.msv_gate:
    return make_gate(nat.MSV_G_U, [target, *controls], u, len(controls))


def gate_diag1(q: int, d0: complex, d1: complex) -> nat.msv_gate:
    return make_gate(nat.MSV_G_DIAG1, [q], [d0, d1])


def gate_diag2(a: int, b: int, d) -> nat.msv_gate:
    return make_gate(nat.MSV_G_DIAG2, [a, b], d)


def gate_swap(a: int, b: int, controls: Sequence[int] = ()) -> nat.msv_gate:
    return make_gate(nat.MSV_G_SWAP, [a, b, *controls], None, len(controls))
