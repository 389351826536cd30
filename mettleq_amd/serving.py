"""Serving loop: circuits of uneven width admitted as they arrive, run
concurrently on resident device states, results returned per request.

The reference executes one circuit per call on the caller's thread
(integrations/_common.py:118-228, qiskit.py:461-528: a synchronous JobV1
per `run`). Here a server keeps a pool of device states per width
(allocated once and kept resident: freeing device memory synchronizes the
whole device), admits each request by the complex64 statevector preflight
(execution.statevector_preflight: the device's free memory now, minus the
bytes admitted by other workers whose states are not allocated yet), and runs up to `workers` requests at a
time, each on its own state and HIP stream, so that their kernels overlap
on the GPU. Requests that arrive while others run are taken as soon as a
worker and a state of their width are free (continuous batching, not a
queue in front of a single executor). `FIFO` mode (workers=1) is the plain
one-circuit-at-a-time path.

Every request's gates go through the native translation and planner
(msv_apply_circuit), its
shots through the counter-based sampler keyed by the request's seed, so a
result does not depend on what else ran beside it: the kernels involve no
atomics or cross-request state.
"""
from __future__ import annotations

import math
import queue
import threading
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

from .workloads import BASE_SEED

STATE_BYTES_PER_AMP = 8  # complex64


@dataclass
class Request:
    id: int
    n: int
    ops: List[Dict]
    shots: int
    seed: int
    arrival: float = 0.0  # seconds from the start of the stream


@dataclass
class Result:
    id: int
    n: int
    outcomes: np.ndarray          # distinct packed outcomes (qubit 0 = MSB), ascending
    counts: np.ndarray            # their counts
    arrival: float
    start: float
    done: float
    probabilities: Optional[np.ndarray] = None
    error: Optional[str] = None

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def qasm_request(source, id: int = 0, shots: int = 1024, seed: int = BASE_SEED,
                 arrival: float = 0.0) -> Request:
    """A request from OpenQASM 2.0 (source text, or a path to a .qasm file)
    through the strict importer (mettleq_amd.qasm; reference qasm.py:290-563):
    the QASM wire format of the serving loop."""
    from pathlib import Path
    from .qasm import parse_qasm, parse_qasm_file
    if isinstance(source, Path) or (isinstance(source, str) and source.rstrip().endswith(".qasm")
                                    and "\n" not in source):
        n, ops = parse_qasm_file(source)
    else:
        n, ops = parse_qasm(str(source))
    return Request(id=int(id), n=int(n), ops=ops, shots=int(shots), seed=int(seed), arrival=float(arrival))


# ---------------------------------------------------------------------------
# workload: Poisson arrivals of random circuits
# ---------------------------------------------------------------------------


def random_circuit(n: int, n_gates: int, rng: np.random.Generator) -> List[Dict]:
    """50% U3 (random wire, angles U[0, 2pi)), 25% CZ or RZZ (50/50, random
    distinct pair, RZZ theta ~ U[0, pi) as Qiskit rzz = ZZPHASE(theta/2)),
    25% CNOT (random ordered distinct pair)."""
    ops: List[Dict] = []
    for _ in range(n_gates):
        u = rng.random()
        if u < 0.5:
            q = int(rng.integers(n))
            ops.append({"name": "U3", "wires": [q], "parameters": [float(x) for x in rng.uniform(0, 2 * math.pi, 3)]})
            continue
        a, b = (int(x) for x in rng.choice(n, size=2, replace=False))
        if u < 0.75:
            if rng.random() < 0.5:
                ops.append({"name": "CZ", "wires": [a, b], "parameters": []})
            else:
                ops.append({"name": "ZZPHASE", "wires": [a, b],
                            "parameters": [float(rng.uniform(0, math.pi)) / 2.0]})
        else:
            ops.append({"name": "CNOT", "wires": [a, b], "parameters": []})
    return ops


def poisson_stream(n_requests: int, rate: float, seed: int = BASE_SEED, widths=(12, 28),
                   gates=(10, 200), shots=(1000, 10000)) -> List[Request]:
    """Arrivals: exponential gaps at `rate` per second from seed base + 5.
    Request i: width ~ U{widths}, gate count ~ U{gates}, shots =
    round(exp(U[ln lo, ln hi])), all from seed base + 1_000_000 + i; its
    sample seed is base + i."""
    arr = np.random.default_rng(seed + 5)
    t = np.cumsum(arr.exponential(1.0 / rate, size=n_requests))
    reqs = []
    for i in range(n_requests):
        rng = np.random.default_rng(seed + 1_000_000 + i)
        n = int(rng.integers(widths[0], widths[1] + 1))
        g = int(rng.integers(gates[0], gates[1] + 1))
        s = int(round(math.exp(rng.uniform(math.log(shots[0]), math.log(shots[1])))))
        reqs.append(Request(id=i, n=n, ops=random_circuit(n, g, rng), shots=s, seed=seed + i,
                            arrival=float(t[i])))
    return reqs


# ---------------------------------------------------------------------------
# resident state pool with preflight admission
# ---------------------------------------------------------------------------


class StatePool:
    """Device states kept resident per width. `acquire(n)` returns a free
    state of width n, allocating one when the preflight admits it, or blocks
    until one is released. The preflight's working set is the device's free
    memory, which already excludes the resident states, so only admissions
    whose allocation is still in flight are passed as reserved bytes."""

    def __init__(self, budget_bytes: Optional[int] = None, factory: Optional[Callable[[int], object]] = None,
                 preflight: Optional[Callable[[int, int], dict]] = None):
        self._free: Dict[int, List[object]] = {}
        self._resident = 0
        self._inflight = 0  # admitted, allocation not finished
        self._cv = threading.Condition()
        self.budget = budget_bytes
        self._factory = factory or _native_state
        self._preflight = preflight or _preflight_in_place
        self.peak_resident = 0
        self.allocations = 0
        self.evictions = 0

    @property
    def resident_bytes(self) -> int:
        return self._resident

    def acquire(self, n: int, timeout: Optional[float] = None):
        need = (1 << n) * STATE_BYTES_PER_AMP
        deadline = None if timeout is None else time.monotonic() + timeout
        with self._cv:
            while True:
                free = self._free.get(n)
                if free:
                    return free.pop()
                within_budget = self.budget is None or self._resident + need <= self.budget
                report = self._preflight(n, self._inflight) if within_budget else {"allowed": False}
                if report.get("allowed"):
                    self._resident += need
                    self._inflight += need
                    self.peak_resident = max(self.peak_resident, self._resident)
                    self.allocations += 1
                    break
                if self._resident == 0:
                    from .execution import StatevectorMemoryError
                    raise StatevectorMemoryError(
                        f"{n}-qubit request refused by the statevector preflight before allocation", report)
                # make room from idle states of other widths (largest first)
                idle = sorted((w for w, lst in self._free.items() if lst and w != n), reverse=True)
                if idle:
                    st = self._free[idle[0]].pop()
                    self._resident -= (1 << st.n) * STATE_BYTES_PER_AMP
                    self.evictions += 1
                    close = getattr(st, "close", None)
                    if close:
                        close()
                    continue
                remaining = None if deadline is None else deadline - time.monotonic()
                if remaining is not None and remaining <= 0:
                    raise TimeoutError(f"no {n}-qubit state became free")
                self._cv.wait(remaining)
        try:
            st = self._factory(n)
        except Exception:
            with self._cv:
                self._resident -= need
                self._inflight -= need
                self._cv.notify_all()
            raise
        with self._cv:
            self._inflight -= need
        return st

    def release(self, state) -> None:
        with self._cv:
            self._free.setdefault(state.n, []).append(state)
            self._cv.notify_all()

    def close(self) -> None:
        with self._cv:
            for lst in self._free.values():
                for st in lst:
                    close = getattr(st, "close", None)
                    if close:
                        close()
            self._free.clear()
            self._resident = 0


def _native_state(n: int):
    from .native_state import NativeState
    return NativeState(n, "complex64")


def _preflight_in_place(n: int, reserved: int) -> dict:
    from .execution import statevector_preflight
    return statevector_preflight(n, dtype="complex64", mode="in_place", reserved_bytes=reserved)


# ---------------------------------------------------------------------------
# execution of one request on a state
# ---------------------------------------------------------------------------


def run_request(state, req: Request, keep_probabilities: bool = False):
    """Reset, apply the request's gates, sample its shots (and optionally
    read the probability vector). Returns (outcomes, counts, probs)."""
    from .circuit import encode_ops
    state.reset()
    state.apply_circuit(*encode_ops(req.ops))
    probs = state.probabilities() if keep_probabilities else None
    packed = state.sample(req.shots, seed=req.seed)
    outcomes, counts = np.unique(packed, return_counts=True)
    return outcomes, counts, probs


# ---------------------------------------------------------------------------
# server
# ---------------------------------------------------------------------------


class Server:
    """`workers` concurrent requests over a shared StatePool; workers=1 is
    the plain FIFO path."""

    def __init__(self, workers: int = 8, pool: Optional[StatePool] = None,
                 executor: Callable = run_request, keep_probabilities: bool = False):
        if workers < 1:
            raise ValueError("workers must be >= 1")
        self.workers = workers
        self.pool = pool or StatePool()
        self._exec = executor
        self._keep = keep_probabilities

    def serve(self, requests: Sequence[Request], realtime: bool = True) -> List[Result]:
        """Feed `requests` at their arrival times (realtime) or all at once,
        return results in request order."""
        q: "queue.Queue[Optional[Request]]" = queue.Queue()
        results: Dict[int, Result] = {}
        lock = threading.Lock()
        errors: List[BaseException] = []
        t0 = time.perf_counter()

        def worker():
            while True:
                req = q.get()
                if req is None:
                    return
                start = time.perf_counter() - t0
                try:
                    st = self.pool.acquire(req.n)
                    try:
                        outcomes, counts, probs = self._exec(st, req, self._keep)
                    finally:
                        self.pool.release(st)
                    res = Result(req.id, req.n, outcomes, counts, req.arrival, start,
                                 time.perf_counter() - t0, probs)
                except Exception as exc:  # reported per request
                    res = Result(req.id, req.n, np.zeros(0, np.uint64), np.zeros(0, np.int64), req.arrival,
                                 start, time.perf_counter() - t0, None, f"{type(exc).__name__}: {exc}")
                    with lock:
                        errors.append(exc)
                with lock:
                    results[req.id] = res

        threads = [threading.Thread(target=worker, daemon=True) for _ in range(self.workers)]
        for th in threads:
            th.start()
        ordered = sorted(requests, key=lambda r: r.arrival)
        for req in ordered:
            if realtime:
                delay = req.arrival - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
            q.put(req)
        for _ in threads:
            q.put(None)
        for th in threads:
            th.join()
        return [results[r.id] for r in requests]


def summarize(results: Sequence[Result]) -> dict:
    ok = [r for r in results if r.error is None]
    if not ok:
        return {"completed": 0, "failed": len(results)}
    lat = np.array([r.latency for r in ok])
    span = max(r.done for r in ok) - min(r.arrival for r in ok)
    return {
        "completed": len(ok),
        "failed": len(results) - len(ok),
        "circuits_per_s": round(len(ok) / span, 2) if span > 0 else None,
        "p50_latency_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
        "p99_latency_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
        "mean_latency_ms": round(1e3 * float(lat.mean()), 3),
        "span_s": round(span, 3),
    }
