"""CPU restatement of the reference's MPS method (test infrastructure only).

NumPy, complex64 site tensors and single-precision SVDs like the reference's. Only tests/ use it, as the checker of mettleq_amd.mps; the
product never imports it. Each step follows src/mettleq/mps_state.py:

* reset: |0...0> as (1, 2, 1) tensors [1, 0], center 0 (mps_state.py:288-332);
* center moves by reduced QR without truncation (mps_state.py:347-399);
* two-site update: move the center to the left site, merge, apply the 4x4
  gate, SVD, keep r = min(rank, max(1, min(dmax, #{s >= eps s_0}))),
  renormalize the kept singular values, left = U, right = S Vh, center at
  the right site (mps_state.py:193-247, 460-480, 536-544, 652-678);
* swaps use eps floored at 1e-6 when positive (mps_state.py:84, 711-742);
* apply_two: the lower operand walks up through swaps to the site below the
  upper one, the gate (conjugated by SWAP for descending operands) acts
  there, and the swaps run back (mps_state.py:744-768);
* to_statevector contracts and normalizes (mps_state.py:1417-1440);
* expectation_product by left-to-right transfer matrices
  (mps_state.py:1269-1330).
"""
from __future__ import annotations

import math
from typing import Dict, List

import numpy as np

SWAP4 = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=np.complex128)
SWAP_EPS_FLOOR = 1e-6


def truncate(s: np.ndarray, dmax: int, eps: float):
    """Kept rank, discarded weight, relative discarded weight and the kept
    norm for singular values s (descending) (mps_state.py:208-231)."""
    s = np.asarray(s, dtype=np.float64)
    r = int(s.size)
    r_eps = r
    if r > 0:
        r_eps = int(np.count_nonzero(s >= eps * s[0]))
    r_keep = min(r, max(1, min(dmax, r_eps)))
    w = s * s
    total = float(np.sum(w))
    discarded = float(np.sum(w[r_keep:]))
    kept = float(np.sum(w[:r_keep]))
    return r_keep, discarded, (discarded / total if total > 0 else 0.0), math.sqrt(kept)


class MPSOracle:
    def __init__(self, n: int, dmax: int = 64, eps: float = 1e-10, renormalize: bool = True):
        self.n, self.dmax, self.eps, self.renormalize = int(n), int(dmax), float(eps), bool(renormalize)
        self.A = [np.array([1, 0], dtype=np.complex64).reshape(1, 2, 1) for _ in range(self.n)]
        self.center = 0
        self.trunc_events = 0
        self.discarded_sum = 0.0
        self.max_bond_ever = 1

    def bond_dims(self) -> List[int]:
        return [int(t.shape[2]) for t in self.A[:-1]]

    def _move_right(self, i: int) -> None:
        t = self.A[i].astype(np.complex128)
        dl, p, dr = t.shape
        q, r = np.linalg.qr(t.reshape(dl * p, dr))
        nxt = self.A[i + 1].astype(np.complex128)
        self.A[i] = q.reshape(dl, p, q.shape[1]).astype(np.complex64)
        self.A[i + 1] = (r @ nxt.reshape(nxt.shape[0], -1)).reshape(r.shape[0], 2, nxt.shape[2]).astype(np.complex64)
        self.center = i + 1

    def _move_left(self, i: int) -> None:
        t = self.A[i].astype(np.complex128)
        dl, p, dr = t.shape
        q, r = np.linalg.qr(t.reshape(dl, p * dr).T)
        prev = self.A[i - 1].astype(np.complex128)
        self.A[i - 1] = (prev.reshape(-1, prev.shape[2]) @ r.T).reshape(prev.shape[0], 2, r.shape[0]).astype(
            np.complex64)
        self.A[i] = q.T.reshape(q.shape[1], p, dr).astype(np.complex64)
        self.center = i - 1

    def _move_center(self, target: int) -> None:
        while self.center < target:
            self._move_right(self.center)
        while self.center > target:
            self._move_left(self.center)

    def _two_site(self, i: int, gate: np.ndarray, eps: float) -> None:
        self._move_center(i)
        left = self.A[i].astype(np.complex128)
        right = self.A[i + 1].astype(np.complex128)
        dl, dr = left.shape[0], right.shape[2]
        theta = (left.reshape(dl * 2, -1) @ right.reshape(right.shape[0], -1)).reshape(dl, 4, dr)
        theta = np.tensordot(np.asarray(gate, dtype=np.complex128), theta, axes=([1], [1])).transpose(1, 0, 2)
        # the reference takes the SVD of the complex64 matrix (cgesdd, single
        # precision: mps_state.py:200-202), so noise-level singular values
        # meet the eps and dmax cuts the way they do there
        mat = theta.reshape(dl * 2, 2 * dr).astype(np.complex64)
        u, s, vh = np.linalg.svd(mat, full_matrices=False)
        u, s, vh = u.astype(np.complex128), s.astype(np.float64), vh.astype(np.complex128)
        r, disc, _rel, kept = truncate(s, self.dmax, eps)
        if r < s.size:
            self.trunc_events += 1
            self.discarded_sum += disc
        sv = s[:r] / (kept if self.renormalize else 1.0)
        self.A[i] = u[:, :r].reshape(dl, 2, r).astype(np.complex64)
        self.A[i + 1] = (sv[:, None] * vh[:r, :]).reshape(r, 2, dr).astype(np.complex64)
        self.center = i + 1
        self.max_bond_ever = max(self.max_bond_ever, r)

    def apply_single(self, u, q: int) -> None:
        t = np.tensordot(np.asarray(u, dtype=np.complex128), self.A[q].astype(np.complex128), axes=([1], [1]))
        self.A[q] = t.transpose(1, 0, 2).astype(np.complex64)

    def apply_two(self, u4, c: int, t: int) -> None:
        i, j = sorted((int(c), int(t)))
        g = np.asarray(u4, dtype=np.complex128)
        if c > t:
            g = SWAP4 @ g @ SWAP4
        swap_eps = max(self.eps, SWAP_EPS_FLOOR) if self.eps > 0 else self.eps
        for k in range(i, j - 1):
            self._two_site(k, SWAP4, swap_eps)
        self._two_site(j - 1, g, self.eps)
        for k in range(j - 2, i - 1, -1):
            self._two_site(k, SWAP4, swap_eps)

    def norm(self) -> float:
        return float(np.linalg.norm(self.A[self.center].astype(np.complex128).reshape(-1)))

    def to_statevector(self) -> np.ndarray:
        psi = self.A[0].astype(np.complex128)
        for t in self.A[1:]:
            psi = np.tensordot(psi, t.astype(np.complex128), axes=([psi.ndim - 1], [0]))
        v = psi.reshape(1 << self.n)
        return v / np.linalg.norm(v)

    def expectation_product(self, operators: Dict[int, np.ndarray]) -> complex:
        env = np.ones((1, 1), dtype=np.complex128)
        eye = np.eye(2, dtype=np.complex128)
        for w, t in enumerate(self.A):
            o = np.asarray(operators.get(w, eye), dtype=np.complex128)
            a = t.astype(np.complex128)
            # env'[b, b'] = sum conj(A[a, s, b]) env[a, a'] O[s, t] A[a', t, b']
            x = np.einsum("ab,bsc->asc", env, a)
            y = np.einsum("st,atc->asc", o, x)
            env = np.einsum("asb,asc->bc", a.conj(), y)
        nrm = self.norm()
        return complex(env.reshape(-1)[0]) / (nrm * nrm)
