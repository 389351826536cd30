"""Serving loop logic on the CPU: the seeded request stream, preflight
admission and eviction in the state pool, and the scheduler with a
stand-in executor (no GPU)."""
import threading
import time

import numpy as np
import pytest

from mettleq_amd.execution import StatevectorMemoryError
from mettleq_amd.serving import Request, Server, StatePool, poisson_stream, random_circuit, summarize


def test_stream_is_seeded_and_within_ranges():
    a = poisson_stream(300, rate=200.0)
    b = poisson_stream(300, rate=200.0)
    assert [r.arrival for r in a] == [r.arrival for r in b]
    assert [r.ops for r in a[:5]] == [r.ops for r in b[:5]]
    assert all(12 <= r.n <= 28 for r in a)
    assert all(10 <= len(r.ops) <= 200 for r in a)
    assert all(1000 <= r.shots <= 10000 for r in a)
    assert [r.seed for r in a[:3]] == [20261203, 20261204, 20261205]
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert abs(gaps.mean() - 1 / 200.0) < 0.25 / 200.0


def test_gate_mix():
    rng = np.random.default_rng(1)
    ops = random_circuit(20, 20000, rng)
    names = [o["name"] for o in ops]
    frac = {k: names.count(k) / len(names) for k in ("U3", "CZ", "ZZPHASE", "CNOT")}
    assert abs(frac["U3"] - 0.5) < 0.02
    assert abs(frac["CZ"] + frac["ZZPHASE"] - 0.25) < 0.02
    assert abs(frac["CZ"] - frac["ZZPHASE"]) < 0.02
    assert abs(frac["CNOT"] - 0.25) < 0.02
    for o in ops:
        if len(o["wires"]) == 2:
            assert o["wires"][0] != o["wires"][1]


class FakeState:
    live = 0
    lock = threading.Lock()

    def __init__(self, n):
        self.n = n
        self.closed = False
        with FakeState.lock:
            FakeState.live += 1

    def close(self):
        self.closed = True
        with FakeState.lock:
            FakeState.live -= 1


def _allow(n, reserved):
    return {"allowed": True}


def test_pool_reuses_and_respects_budget():
    pool = StatePool(budget_bytes=3 * (1 << 10) * 8, factory=FakeState, preflight=_allow)
    s1 = pool.acquire(10)
    s2 = pool.acquire(10)
    pool.release(s1)
    assert pool.acquire(10) is s1  # reused, not reallocated
    assert pool.allocations == 2
    s3 = pool.acquire(10)
    assert pool.resident_bytes == 3 * (1 << 10) * 8
    got = []
    t = threading.Thread(target=lambda: got.append(pool.acquire(10)))
    t.start()
    time.sleep(0.05)
    assert not got  # blocked: budget full
    pool.release(s2)
    t.join(1.0)
    assert got == [s2]
    pool.release(s3)


def test_pool_evicts_idle_states_of_other_widths():
    pool = StatePool(budget_bytes=(1 << 12) * 8, factory=FakeState, preflight=_allow)
    small = [pool.acquire(10) for _ in range(4)]
    for s in small:
        pool.release(s)
    big = pool.acquire(12)
    assert big.n == 12 and pool.evictions == 4 and all(s.closed for s in small)


def test_pool_refuses_what_cannot_fit():
    def refuse(n, reserved):
        return {"allowed": False, "decision": "refused"}
    pool = StatePool(factory=FakeState, preflight=refuse)
    with pytest.raises(StatevectorMemoryError):
        pool.acquire(30)


def test_server_concurrency_and_summary():
    def fake_exec(state, req, keep):
        time.sleep(0.01)
        out = np.array([req.seed % 7], dtype=np.uint64)
        return out, np.array([req.shots]), None

    reqs = [Request(i, 10 + i % 3, [], 100 + i, 1000 + i, arrival=0.0) for i in range(24)]
    pool = StatePool(factory=FakeState, preflight=_allow)
    t0 = time.perf_counter()
    fifo = Server(1, pool, fake_exec).serve(reqs, realtime=False)
    t_fifo = time.perf_counter() - t0
    t0 = time.perf_counter()
    conc = Server(6, pool, fake_exec).serve(reqs, realtime=False)
    t_conc = time.perf_counter() - t0
    assert t_conc < t_fifo / 2
    for a, b in zip(fifo, conc):
        assert a.id == b.id and np.array_equal(a.outcomes, b.outcomes) and np.array_equal(a.counts, b.counts)
    s = summarize(conc)
    assert s["completed"] == 24 and s["failed"] == 0
    assert s["p50_latency_ms"] <= s["p99_latency_ms"]


def test_server_reports_errors_per_request():
    def bad_exec(state, req, keep):
        if req.id == 3:
            raise ValueError("boom")
        return np.zeros(1, np.uint64), np.ones(1, np.int64), None
    reqs = [Request(i, 10, [], 10, i) for i in range(6)]
    res = Server(2, StatePool(factory=FakeState, preflight=_allow), bad_exec).serve(reqs, realtime=False)
    assert res[3].error and "boom" in res[3].error
    assert all(r.error is None for i, r in enumerate(res) if i != 3)
