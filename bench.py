#!/usr/bin/env python3
"""Benchmark: structured-layer HBM throughput at 30 qubits, complex64.

Workload (one "step"): reset to |0...0> and run the 40-layer 30-qubit circuit
of mettleq_amd.workloads.layered_30q (U3-all / RZZ ring / CNOT chain cycle,
full QFT every 10th layer) on one GPU, inputs resident in HBM.

value = sum over layers of (bytes moved by the layer's full-state sweeps)
        / (GPU time of the timed steps), in GB/s; one sweep moves
        2 * 2^n * 8 bytes (read + write of the complex64 state).
roofline = the dominant kernel, k_rot<float, 4, false> (the rotation sweeps of
        the U3 layers): algorithmic bytes per launch / mean launch duration
        from HIP events on the state's stream, over back-to-back U3 layers
        that read no deferred work (3 launches each); the other sweep kernels,
        including the gather variants of k_rot and k_qft that read deferred
        layers (priced as an RZZ ring + CNOT chain + layer loop minus the
        layer's plain launches), are listed beside it.
cpu_baseline = the oracle's gate-by-gate CPU restatement (oracle/sv_cpu.c)
        timed on this host on a bounded sample (one layer of each type at
        --cpu-qubits), reported in the same unit.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK_GBPS = 8000.0  # MI355X HBM3E


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--layers", type=int, default=40)
    ap.add_argument("--cpu-qubits", type=int, default=27)
    ap.add_argument("--cpu-repeats", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-oracle", action="store_true", help="skip the complex128 fidelity check")
    ap.add_argument("--json-out", default=None)
    ap.add_argument("--workload", default="layered30",
                    choices=["layered30", "brick20", "memlimit", "vqe16", "serve", "mps"],
                    help="layered30 (default): the 30-qubit layered circuit; brick20: 20-qubit brick "
                         "circuit with probabilities and shots vs the CPU restatement; memlimit: 34q "
                         "complex64 and 33q complex128 through the preflight; vqe16: 1024 bindings of a "
                         "16-qubit ansatz with adjoint gradients; serve: Poisson request stream; mps: 40-qubit "
                         "brick circuit by the MPS method")
    ap.add_argument("--no-extra", action="store_true",
                    help="layered30 only: skip the short brick20 / vqe16 / serve summaries")
    ap.add_argument("--rates", default="50,100,200,500,1000,2000")
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--traffic-from", default=os.path.join(ROOT, "profiles", "k_rot_30q_layered_circuit_pmc.md"),
                    help="rocprofv3 FETCH_SIZE/WRITE_SIZE summary of the dominant kernel (HBM bytes per launch)")
    return ap.parse_args()


def apply_layer(state, layer, n, lay):
    from mettleq_amd import gates as G
    from mettleq_amd import _native as nat
    from mettleq_amd.workloads import ring_bonds
    t = layer["type"]
    if t == "u3":
        mats = [G.single_qubit_matrix("U3", layer["angles"][q]) for q in range(n)]
        state.apply_1q_layer(mats)
    elif t == "rzz":
        bonds = ring_bonds(n)
        state.apply_phase_layer([nat.MSV_PHASE_ZZ] * len(bonds), bonds,
                                [float(x) / 2.0 for x in layer["thetas"][:len(bonds)]])
    elif t == "cnot":
        rows, c = lay["chain"]
        state.apply_xor_affine(rows, c)
    elif t == "cz":
        state.apply_phase_layer([nat.MSV_PHASE_CPHASE] * len(layer["bonds"]), layer["bonds"],
                                [math.pi] * len(layer["bonds"]))
    elif t == "qft":
        state.apply_qft(0, n)
    else:
        raise ValueError(t)


def run_circuit(state, layers, n, lay):
    state.reset()
    for layer in layers:
        apply_layer(state, layer, n, lay)


def read_traffic(path):
    """HBM bytes per k_rot launch from a committed rocprofv3 PMC summary: the
    line `traffic_bytes_per_launch: N` (FETCH_SIZE and WRITE_SIZE of the same
    bench command, each scaled by what the counter reads for the copy probe)."""
    try:
        with open(path) as fh:
            for line in fh:
                if line.strip().startswith("traffic_bytes_per_launch:"):
                    return int(float(line.split(":", 1)[1].strip())), os.path.basename(path)
    except OSError:
        pass
    return None, None


def host_cpu_info() -> dict:
    """Model name, logical CPUs and physical cores of this host (Linux)."""
    model, cores = None, set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None and core is not None:
                    cores.add((phys, core))
                    phys = core = None
        if phys is not None and core is not None:
            cores.add((phys, core))
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count(), "physical_cores": len(cores) or None}


def cpu_baseline(n_cpu: int, layers, repeats: int = 1):
    """Gate-by-gate CPU restatement on one layer of each type at n_cpu qubits."""
    from oracle.sv_cpu import CpuEngine, lib
    from mettleq_amd.workloads import layer_ops
    lib()
    threads = int(lib().cpu_num_threads())
    seen = {}
    for layer in layers:
        seen.setdefault(layer["type"], layer)
    eng = CpuEngine(n_cpu)
    rng = np.random.default_rng(7)
    v = (rng.normal(size=1 << n_cpu) + 1j * rng.normal(size=1 << n_cpu)).astype(np.complex64)
    eng.state[:] = v / np.linalg.norm(v)
    per = {}
    total_t = 0.0
    total_passes = 0
    for t, layer in seen.items():
        src = dict(layer)
        if t == "u3":
            src = {"type": "u3", "angles": layer["angles"][:n_cpu]}
        if t == "rzz":
            src = {"type": "rzz", "thetas": layer["thetas"][:n_cpu]}
        ops = layer_ops(src, n_cpu)
        p0 = eng.passes
        t0 = time.perf_counter()
        for _ in range(repeats):
            eng.run(ops)
        dt = (time.perf_counter() - t0) / repeats
        passes = (eng.passes - p0) // repeats
        per[t] = {"ms": dt * 1e3, "passes": passes}
        total_t += dt
        total_passes += passes
    state_bytes = (1 << n_cpu) * 8
    gbps = total_passes * 2 * state_bytes / total_t / 1e9
    # the same restatement on one thread, on a sample 8x smaller
    n1 = max(4, n_cpu - 3)
    lib().cpu_set_threads(1)
    try:
        eng1 = CpuEngine(n1)
        eng1.state[:] = eng.state[: 1 << n1] / np.linalg.norm(eng.state[: 1 << n1])
        t1 = 0.0
        p1 = 0
        for t, layer in seen.items():
            src = dict(layer)
            if t == "u3":
                src = {"type": "u3", "angles": layer["angles"][:n1]}
            if t == "rzz":
                src = {"type": "rzz", "thetas": layer["thetas"][:n1]}
            ops = layer_ops(src, n1)
            q0 = eng1.passes
            t0 = time.perf_counter()
            eng1.run(ops)
            t1 += time.perf_counter() - t0
            p1 += eng1.passes - q0
    finally:
        lib().cpu_set_threads(threads)
    single = {"value": round(p1 * 2 * (1 << n1) * 8 / t1 / 1e9, 3), "unit": "GB/s", "cores": 1,
              "sample": f"one layer of each type at {n1} qubits complex64, gate by gate, once ({t1:.1f} s)"}
    return {
        "value": round(gbps, 3),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"one layer of each type ({', '.join(sorted(per))}) at {n_cpu} qubits complex64, "
                   f"gate by gate, {repeats}x ({total_passes} full-state passes per repeat, "
                   f"{total_t * repeats:.1f} s in all)"),
        "layer_ms": {k: round(v["ms"], 2) for k, v in per.items()},
        "layer_ms_scaled_to_30q": {k: round(v["ms"] * (1 << (30 - n_cpu)), 1) for k, v in per.items()},
        "single_thread": single,
        "host": host_cpu_info(),
    }


# ---------------------------------------------------------------------------
# other configurations (one JSON line each with --workload)
# ---------------------------------------------------------------------------


def bench_brick20(args, quick=False):
    """20 qubits complex64, 10 layers of U3-all + CZ brick: full FP64
    probability vector and 1000 shots (seed base + 1), GPU engine vs the
    gate-by-gate CPU restatement of the reference's CPU path."""
    from mettleq_amd.device import Device
    from mettleq_amd.workloads import BASE_SEED, brick_20q, circuit_ops
    from oracle.sv_cpu import CpuEngine, lib
    import oracle.sv_oracle as O
    n = 20
    ops = circuit_ops(brick_20q(n, 10), n)
    dev = Device(n)
    for _ in range(2):
        dev.reset()
        dev.execute(ops)
        dev.synchronize()
    st = dev.native
    reps = 3 if quick else 10
    ts = []
    for _ in range(reps):
        dev.reset()
        st.sync()
        t0 = time.perf_counter()
        dev.execute(ops)
        probs = st.probabilities()
        shots = st.sample(1000, seed=BASE_SEED + 1)
        ts.append(time.perf_counter() - t0)
    gpu_state = st.to_numpy()
    lib()
    cpu = CpuEngine(n)
    t0 = time.perf_counter()
    cpu.run(ops)
    cpu_probs = O.amp_probs_c64(cpu.state)
    cpu_shots = cpu.sample_indices(1000, BASE_SEED + 1)
    cpu_s = time.perf_counter() - t0
    # bit-exact sampling pin: the C restatement sampling the GPU's amplitudes
    cross = CpuEngine(n, gpu_state).sample_indices(1000, BASE_SEED + 1)
    dev.close()
    return {
        "metric": "20q brick circuit: probabilities + 1000 shots, ms per circuit",
        "workload": "20-qubit complex64, 10 layers U3-all + CZ brick, full probabilities + 1000 shots",
        "gpu_ms": round(1e3 * float(np.median(ts)), 3),
        "cpu_ms": round(1e3 * cpu_s, 1),
        "cpu_threads": int(lib().cpu_num_threads()),
        "probabilities_max_abs_diff_vs_cpu": float(np.max(np.abs(probs - cpu_probs))),
        "shots_bit_exact_vs_cpu_sampler_on_gpu_state": bool(np.array_equal(shots, cross)),
        "shots_equal_to_cpu_run": float(np.mean(shots == cpu_shots)),
        "norm": float(probs.sum()),
    }


def bench_mps(args, quick=False):
    """MPS method: brick layers (U3-all + CZ brick) on 40 qubits, depth 16 (12 quick),
    dmax 64, eps 1e-10, on the device engine (mettleq_amd.mps) and on the
    NumPy restatement of the reference's CPU MPS (oracle/mps_oracle.py, the
    reference's algorithm with LAPACK SVDs), same circuit; <Z_0 Z_1> from both."""
    from mettleq_amd import gates as G
    from mettleq_amd.mps import MPSOptions, MPSState
    from mettleq_amd.workloads import brick_20q
    from oracle.mps_oracle import MPSOracle
    n, depth, dmax = 40, (12 if quick else 16), 64
    layers = brick_20q(n, depth)
    cz = np.diag([1, 1, 1, -1]).astype(complex)

    def run(mps):
        for layer in layers:
            if layer["type"] == "u3":
                for q in range(n):
                    mps.apply_single(G.single_qubit_matrix("U3", layer["angles"][q]), q)
            else:
                for a, b in layer["bonds"]:
                    mps.apply_two(cz, a, b)

    z = np.diag([1.0, -1.0])
    with MPSState(n, MPSOptions(dmax=dmax, eps=1e-10)) as warm:
        run(warm)
    with MPSState(n, MPSOptions(dmax=dmax, eps=1e-10)) as mps:
        t0 = time.perf_counter()
        run(mps)
        zz = mps.expectation_product({0: z, 1: z})
        gpu_s = time.perf_counter() - t0
        d = mps.truncation_diagnostics()
        bonds = mps.bond_dims()
    ref = MPSOracle(n, dmax=dmax, eps=1e-10)
    t0 = time.perf_counter()
    run(ref)
    zz_ref = ref.expectation_product({0: z, 1: z})
    cpu_s = time.perf_counter() - t0
    updates = sum(len(l["bonds"]) for l in layers if l["type"] == "cz")
    return {
        "metric": "MPS brick circuit: seconds per circuit (device engine vs the reference's CPU algorithm)",
        "workload": f"{n} qubits, {depth} layers U3-all + CZ brick, dmax {dmax}, eps 1e-10",
        "gpu_s": round(gpu_s, 3),
        "cpu_s": round(cpu_s, 3),
        "two_qubit_gates": updates,
        "svd_calls": d["svd_calls"],
        "max_bond": d["maximum_bond_dimension_reached"],
        "bond_mid": bonds[n // 2],
        "truncation_events": d["events"],
        "truncation_events_cpu": ref.trunc_events,
        "relative_discarded_weight_sum": d["relative_discarded_weight_sum"],
        "discarded_weight_sum": d["local_discarded_weight_sum"],
        "discarded_weight_sum_cpu": float(ref.discarded_sum),
        "zz01_gpu": round(float(zz.real), 6),
        "zz01_cpu": round(float(zz_ref.real), 6),
        "cpu_kind": "port (NumPy/LAPACK restatement of the reference's CPU MPS)",
    }


def bench_memlimit(args):
    """34q complex64 (128 GiB) and 33q complex128 (128 GiB) through the
    preflight, in place; 1-F between 33q complex64 and the 33q complex128
    oracle, both resident. Layer mix of the 20q brick circuit, depth 20."""
    from mettleq_amd import _native as nat
    from mettleq_amd.execution import statevector_preflight
    from mettleq_amd.native_state import NativeState
    from mettleq_amd.workloads import brick_20q
    out = {"metric": "memory-limit run: seconds per circuit and 1-F (33q c64 vs c128)",
           "workload": "brick layers (U3-all + CZ brick), depth 20: 34q complex64, 33q complex128 and 33q complex64"}
    pf = {}
    for n, dt, mode in ((34, "complex64", "in_place"), (34, "complex64", "out_of_place"),
                        (33, "complex128", "in_place"), (35, "complex64", "out_of_place"),
                        (36, "complex64", "in_place")):
        r = statevector_preflight(n, dtype=dt, mode=mode)
        pf[f"{n}q_{dt}_{mode}"] = r["decision"]
    out["preflight"] = pf

    def run(n, dtype, depth=20):
        layers = brick_20q(n, depth)
        st = NativeState(n, dtype)
        st.sync()
        s0 = st.sweep_count()
        st.timer_start()
        for layer in layers:
            if layer["type"] == "u3":
                from mettleq_amd import gates as G
                st.apply_1q_layer([G.single_qubit_matrix("U3", a) for a in layer["angles"]])
            else:
                st.apply_phase_layer([nat.MSV_PHASE_CPHASE] * len(layer["bonds"]), layer["bonds"],
                                     [math.pi] * len(layer["bonds"]))
        ms = st.timer_stop()
        sweeps = st.sweep_count() - s0
        return st, ms, sweeps

    st34, ms34, sw34 = run(34, "complex64")
    gb34 = sw34 * 2 * st34.nbytes / (ms34 * 1e-3) / 1e9
    out["34q_complex64"] = {"s": round(ms34 / 1e3, 3), "sweeps": sw34, "GBps": round(gb34, 1),
                            "norm2": st34.norm2()}
    st34.close()
    st128, ms128, sw128 = run(33, "complex128")
    st64, ms64, sw64 = run(33, "complex64")
    ip = st64.inner(st128)
    fid = abs(ip) ** 2 / (st64.norm2() * st128.norm2())
    out["33q_complex128"] = {"s": round(ms128 / 1e3, 3), "sweeps": sw128,
                             "GBps": round(sw128 * 2 * st128.nbytes / (ms128 * 1e-3) / 1e9, 1)}
    out["33q_complex64"] = {"s": round(ms64 / 1e3, 3), "sweeps": sw64,
                            "GBps": round(sw64 * 2 * st64.nbytes / (ms64 * 1e-3) / 1e9, 1)}
    out["one_minus_fidelity_33q_c64_vs_c128"] = float(1.0 - fid)
    st64.close()
    st128.close()
    return out


def _vqe_plain_gradients(params_row, n, L):
    """The plain per-circuit path on the same GPU: every gradient entry by
    parameter shift, each shifted circuit executed alone through Device
    (fusion + dispatch) and its <H> read with the Pauli-word reduction."""
    from mettleq_amd.device import Device
    words = [{i: "Z", (i + 1) % n: "Z"} for i in range(n)]

    def ops_for(th):
        ops = []
        for l in range(L):
            for q in range(n):
                ops.append({"name": "RY", "wires": [q], "parameters": [float(th[2 * (l * n + q)])]})
                ops.append({"name": "RZ", "wires": [q], "parameters": [float(th[2 * (l * n + q) + 1])]})
            for q in range(n):
                ops.append({"name": "CZ", "wires": [q, (q + 1) % n], "parameters": []})
        return ops

    dev = Device(n)
    grads = np.zeros(len(params_row))
    for p in range(len(params_row)):
        vals = []
        for sh in (math.pi / 2, -math.pi / 2):
            th = params_row.copy()
            th[p] += sh
            dev.reset()
            dev.execute(ops_for(th))
            vals.append(float(np.real(dev.native.expval_pauli(words)).sum()))
        grads[p] = 0.5 * (vals[0] - vals[1])
    dev.close()
    return grads


def bench_vqe16(args, quick=False):
    """1024 bindings of the 16-qubit hardware-efficient ansatz (8 layers of
    RY, RZ per qubit + CZ ring, 256 parameters), H = sum Z_i Z_{i+1}:
    energies + all gradients per graph launch, vs the plain per-circuit
    parameter-shift path on the same GPU."""
    from mettleq_amd.ansatz import hardware_efficient, zz_ring_terms
    from mettleq_amd.workloads import BASE_SEED
    n, L, B = 16, 8, 1024
    rng = np.random.default_rng(BASE_SEED + 4)
    params = rng.uniform(-math.pi, math.pi, (B, 2 * n * L))
    with hardware_efficient(n, L) as a:
        a.observable(zz_ring_terms(n))
        t0 = time.perf_counter()
        e, g = a.evaluate(params)
        first = time.perf_counter() - t0
        ts = []
        for _ in range(3 if quick else 10):
            t0 = time.perf_counter()
            e, g = a.evaluate(params)
            ts.append(time.perf_counter() - t0)
        fwd = []
        for _ in range(3):
            t0 = time.perf_counter()
            a.evaluate(params, grads=False)
            fwd.append(time.perf_counter() - t0)
        info = a.info(B)
        sw = a.sweep_times(B, grads=True, reps=3)
        a.evaluate(params)
    t_adj = float(np.median(ts))
    fwd_frac = sw["forward_bytes"] / (sw["forward_ms"] * 1e-3) / (PEAK_GBPS * 1e9)
    bwd_frac = sw["backward_bytes"] / (sw["backward_ms"] * 1e-3) / (PEAK_GBPS * 1e9)
    n_plain = 1 if quick else 2
    t0 = time.perf_counter()
    plain = [_vqe_plain_gradients(params[b].copy(), n, L) for b in range(n_plain)]
    t_plain = (time.perf_counter() - t0) / n_plain
    diff = max(float(np.max(np.abs(plain[b] - g[b]))) for b in range(n_plain))
    return {
        "metric": "gradients/s, 16q hardware-efficient ansatz, 1024 bindings x 256 parameters",
        "value": round(B * params.shape[1] / t_adj, 1),
        "unit": "gradients/s",
        "ms_per_evaluation": round(1e3 * t_adj, 3),
        "forward_only_ms": round(1e3 * float(np.median(fwd)), 3),
        "first_call_s_autotune_and_graph_capture": round(first, 3),
        "plan": info,
        "sweeps": {"forward_ms": round(sw["forward_ms"], 3), "backward_ms": round(sw["backward_ms"], 3),
                   "forward_GBps": round(sw["forward_bytes"] / (sw["forward_ms"] * 1e-3) / 1e9, 1),
                   "backward_GBps": round(sw["backward_bytes"] / (sw["backward_ms"] * 1e-3) / 1e9, 1),
                   "forward_hbm_frac": round(fwd_frac, 3), "backward_hbm_frac": round(bwd_frac, 3),
                   "bytes": "forward sweep 2 x B x 2^16 x 8 B (phi read + write); backward 4 x (phi and lambda)"},
        "plain_per_circuit_gradients_per_s": round(params.shape[1] / t_plain, 1),
        "plain_bindings_timed": n_plain,
        "speedup_vs_plain": round((B * params.shape[1] / t_adj) / (params.shape[1] / t_plain), 1),
        "max_abs_diff_adjoint_vs_plain_parameter_shift": diff,
        "mean_energy": float(np.mean(e)),
    }


def bench_serve(args, rates=None, n_requests=None, quick=False):
    """Poisson request streams at each rate: the continuous-batching server
    vs the plain FIFO one-at-a-time path, circuits/s and p50/p99 latency."""
    from mettleq_amd.serving import Server, StatePool, poisson_stream, summarize
    rates = rates or [float(x) for x in args.rates.split(",")]
    n_requests = n_requests or args.requests
    pool = StatePool()
    # warm the pool and the planner once with a short stream of every width
    warm = poisson_stream(17 * 2, 1e6, widths=(12, 28), gates=(10, 20), shots=(1000, 1000))
    Server(args.workers, pool).serve(warm, realtime=False)
    out = {"metric": "served circuits/s and p50/p99 latency (arrival to result)",
           "workload": (f"Poisson arrivals, {n_requests} requests per rate, width U{{12..28}}, "
                        "10-200 gates (50% U3, 25% CZ/RZZ, 25% CNOT), shots log-uniform 1e3-1e4"),
           "workers": args.workers, "rates": {}}
    for rate in rates:
        reqs = poisson_stream(n_requests, rate)
        cont = summarize(Server(args.workers, pool).serve(reqs, realtime=True))
        fifo = summarize(Server(1, pool).serve(reqs, realtime=True))
        out["rates"][str(int(rate))] = {"continuous": cont, "fifo": fifo}
        log(f"serve rate {rate}: continuous {cont}, fifo {fifo}")
    out["pool"] = {"peak_resident_GiB": round(pool.peak_resident / 2**30, 2), "allocations": pool.allocations,
                   "evictions": pool.evictions}
    pool.close()
    return out


def extra_summaries(args):
    """Short runs of the other configurations for the default bench line."""
    res = {}
    for name, fn in (("memlimit", lambda: bench_memlimit(args)),
                     ("brick20", lambda: bench_brick20(args, quick=True)),
                     ("vqe16_adjoint", lambda: bench_vqe16(args, quick=True)),
                     ("serve_sweep", lambda: bench_serve(args, rates=[50.0, 200.0, 1000.0, 2000.0],
                                                         n_requests=2000, quick=True)),
                     ("mps40", lambda: bench_mps(args, quick=True))):
        try:
            t0 = time.perf_counter()
            res[name] = fn()
            log(f"{name} summary done in {time.perf_counter() - t0:.1f} s")
        except Exception as exc:  # keep the headline line even if a side leg fails
            res[name] = {"error": f"{type(exc).__name__}: {exc}"}
            log(f"{name} summary failed: {exc}")
    return res


def main():
    args = parse()
    if args.workload != "layered30":
        if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
            print("bench.py measures one GPU; run with --gpus 1.", file=sys.stderr)
            sys.exit(2)
        out = {"brick20": bench_brick20, "memlimit": bench_memlimit, "vqe16": bench_vqe16,
               "serve": bench_serve, "mps": bench_mps}[args.workload](args)
        line = json.dumps(out)
        if args.json_out:
            with open(args.json_out, "w") as fh:
                fh.write(line + "\n")
        print(line, flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: every box in this setup has a single MI355X and the "
              "engine's multi-GPU sharding is designed, not a measured path. "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    from mettleq_amd import _native as nat
    from mettleq_amd.native_state import NativeState
    from mettleq_amd.layers import compose_inverse_affine
    from mettleq_amd.workloads import layered_30q

    n = args.qubits
    layers = layered_30q(n, args.layers)
    chain = [{"name": "CNOT", "wires": [i, i + 1]} for i in range(n - 1)]
    lay = {"chain": compose_inverse_affine(chain, n)}
    info = nat.device_info(0)
    log(f"device {info['name']} {info['arch']} CUs={info['compute_units']} "
        f"HBM={info['total_bytes']/2**30:.1f} GiB free={info['free_bytes']/2**30:.1f} GiB")

    # achievable copy bandwidth (calibration)
    copy_bytes = 1 << 32
    copy_ms = nat.probe_copy_ms(0, copy_bytes, 10)
    copy_gbps = 2 * copy_bytes / (copy_ms * 1e-3) / 1e9
    log(f"copy probe: {copy_ms:.3f} ms per 4 GiB copy -> {copy_gbps:.0f} GB/s")

    state = NativeState(n, "complex64")
    state_bytes = state.nbytes
    sweep_bytes = 2 * state_bytes

    for w in range(args.warmup):
        run_circuit(state, layers, n, lay)
        state.sync()
        log(f"warmup {w + 1}/{args.warmup} done")

    # per-layer breakdown (one instrumented pass). A U3 or QFT layer that
    # follows deferred layers (S = 0) reads them in its first sweep:
    # "u3_fused" / "qft_fused" (gather variants of k_rot / k_qft). A U3 layer
    # that finds the previous U3 layer's last rotations pending runs them in
    # its first sweep, across the deferred ring and chain: "u3_cross"
    # (k_rot_relabel). A QFT that finds pending rotations runs them first
    # (one plain k_rot sweep).
    state.sync()
    state.reset()  # pending until the first layer (a U3 layer writes its product state)
    per_type = {}
    labels = []
    launches = {}  # sweep kernel -> launches per step
    step_sweeps = 0
    for i, layer in enumerate(layers):
        pend = state.pending()
        s0 = state.sweep_count()
        state.timer_start()
        apply_layer(state, layer, n, lay)
        ms = state.timer_stop()
        S = state.sweep_count() - s0
        step_sweeps += S
        t = layer["type"]
        tail = bool(pend & nat.MSV_PENDING_TAIL)
        read = bool(pend & (nat.MSV_PENDING_DIAG | nat.MSV_PENDING_PERM))
        label = t
        if t == "u3" and S == 0 and i == 0:
            label = "u3_on_basis"  # the reset state's product state, one write pass
        elif t == "u3" and tail and S > 0:
            label = "u3_cross"
        elif t in ("u3", "qft") and read and S > 0:
            label = t + "_fused"
        labels.append(label)
        d = per_type.setdefault(label, {"ms": [], "S": []})
        d["ms"].append(ms)
        d["S"].append(S)

        def add(k, c):
            if c > 0:
                launches[k] = launches.get(k, 0) + c

        if label == "u3":
            add("k_rot", S)
        elif label == "u3_fused":
            add("k_rot_gather", 1)
            add("k_rot", S - 1)
        elif label == "u3_cross":
            add("k_rot_relabel", 1)
            add("k_rot", S - 1)
        elif t == "qft" and S > 0:
            rest = S
            if tail:
                add("k_rot", 1)  # the pending rotations run first
                rest -= 1
            if read:
                add("k_qft_gather", 1)
                rest -= 1
            add("k_qft", rest)
        elif t in ("rzz", "cnot", "cz") and S > 0:
            add("k_phase" if t != "cnot" else "k_xor_gather", S)
    log("per-layer: " + ", ".join(f"{k}: {np.mean(v['ms']):.3f} ms S={max(v['S'])}" for k, v in per_type.items()))

    # mean launch of each sweep kernel, from short back-to-back loops (the
    # queue stays full, so host-side set-up between launches is hidden)
    first = {t: next((l for l in layers if l["type"] == t), None) for t in ("u3", "rzz", "cnot", "qft")}

    def steady(seq, reps=4, tail=1):
        old = nat.get_tuning(nat.MSV_TUNE_DEFER_TAIL)
        nat.set_tuning(nat.MSV_TUNE_DEFER_TAIL, tail)
        try:
            state.sync()
            s0 = state.sweep_count()
            state.timer_start()
            for _ in range(reps):
                for l in seq:
                    apply_layer(state, l, n, lay)
            ms = state.timer_stop()
            return ms / reps, (state.sweep_count() - s0) / reps
        finally:
            nat.set_tuning(nat.MSV_TUNE_DEFER_TAIL, old)

    launch_ms = {}
    ring_chain = first["rzz"] is not None and first["cnot"] is not None
    if first["u3"] is not None:
        ms, S = steady([first["u3"]], tail=0)
        launch_ms["k_rot"] = ms / S
        if ring_chain:
            ms, S = steady([first["rzz"], first["cnot"], first["u3"]], tail=0)
            launch_ms["k_rot_gather"] = ms - (S - 1) * launch_ms["k_rot"]
        if "u3_cross" in per_type:
            # from the instrumented pass: the crossing sweep's cost depends on
            # the tail it finds (8 bits after a layer's first plan, 9 after a
            # crossing), so it is taken from the step's own U3 layers
            v = per_type["u3_cross"]
            launch_ms["k_rot_relabel"] = float(np.mean(v["ms"])) - (max(v["S"]) - 1) * launch_ms["k_rot"]
    if first["qft"] is not None:
        ms, S = steady([first["qft"]])
        launch_ms["k_qft"] = ms / S
        if ring_chain:
            ms, S = steady([first["rzz"], first["cnot"], first["qft"]])
            launch_ms["k_qft_gather"] = ms - (S - 1) * launch_ms["k_qft"]
    log("steady-state launch ms: " + ", ".join(f"{k}: {v:.4f}" for k, v in launch_ms.items()))

    # timed region
    state.sync()
    sweeps0 = state.sweep_count()
    state.timer_start()
    t0 = time.perf_counter()
    for k in range(args.steps):
        run_circuit(state, layers, n, lay)
    gpu_ms = state.timer_stop()
    wall = time.perf_counter() - t0
    sweeps = state.sweep_count() - sweeps0
    log(f"timed {args.steps} steps: {gpu_ms:.1f} ms GPU, {wall * 1e3:.1f} ms wall, {sweeps} sweeps")
    ms_per_step = gpu_ms / args.steps
    # value: bytes of the sweeps performed (the reset's product state is
    # written in a half sweep, not counted)
    layer_sweeps_per_step = step_sweeps
    s_min = {"u3": 2, "qft": 2, "rzz": 1, "cnot": 1, "cz": 1}
    s_min_per_step = sum(s_min[l["type"]] for l in layers)
    value = layer_sweeps_per_step * sweep_bytes / (ms_per_step * 1e-3) / 1e9

    # per sweep kernel: launches per step and mean launch duration (HIP events
    # on the state's stream around back-to-back layers, as above)
    names = {"k_rot": "k_rot<float, 4, false>",
             "k_rot_gather": "k_rot<float, 4, true> (first sweep reads the deferred RZZ ring + CNOT chain)",
             "k_rot_relabel": "k_rot_relabel<float, 4> (pending rotations, deferred ring + chain, first rotations)",
             "k_qft": "k_qft<float, 4, false, false>",
             "k_qft_gather": "k_qft<float, 4, false, true> (first sweep reads the deferred RZZ ring + CNOT chain)",
             "k_phase": "k_phase", "k_xor_gather": "k_xor_gather"}
    kernels = {}
    for k, cnt in launches.items():
        avg = launch_ms.get(k)
        if avg is None:
            continue
        kernels[names[k]] = {"launches": cnt, "ms": round(avg * cnt, 3), "avg_launch_ms": round(avg, 4),
                             "GBps": round(sweep_bytes / (avg * 1e-3) / 1e9, 1),
                             "frac": round(sweep_bytes / (avg * 1e-3) / 1e9 / PEAK_GBPS, 4)}
    dom = max(kernels, key=lambda k: kernels[k]["ms"])
    tile_avg_ms = kernels[dom]["avg_launch_ms"]
    achieved = sweep_bytes / (tile_avg_ms * 1e-3) / 1e9

    # fidelity against the complex128 oracle mode and final reductions
    extras = {}
    if not args.no_oracle:
        ref = NativeState(n, "complex128")
        run_circuit(ref, layers, n, lay)
        run_circuit(state, layers, n, lay)
        ip = state.inner(ref)
        nrm_a = state.norm2()
        nrm_b = ref.norm2()
        fid = abs(ip) ** 2 / (nrm_a * nrm_b)
        extras["one_minus_fidelity_vs_c128"] = float(1.0 - fid)
        extras["norm2_c64"] = nrm_a
        marg = state.probabilities(list(range(min(20, n))))
        top = np.argsort(marg)[::-1][:16]
        extras["marginal_w0_19_top16"] = {int(i): float(marg[i]) for i in top}
        ref.close()
        log(f"1-F vs complex128 = {extras['one_minus_fidelity_vs_c128']:.3e}")

    cpu = None
    if not args.no_cpu:
        log(f"cpu baseline at {args.cpu_qubits} qubits ...")
        cpu = cpu_baseline(args.cpu_qubits, layers, args.cpu_repeats)
        log(f"cpu baseline: {cpu['value']} GB/s with {cpu['cores']} threads")

    traffic, traffic_src = read_traffic(args.traffic_from) if n == 30 else (None, None)
    if not args.no_extra:
        state.close()
        extras["configs"] = extra_summaries(args)
    out = {
        "metric": "layer GB/s & %HBM peak at 30q c64; served circuits/s + p50/p99 latency",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic (seeded circuit, base seed 20261203)",
        "config": {
            "workload": f"{n}-qubit complex64 statevector, {args.layers} layers: U3-all / RZZ ring / "
                        "CNOT chain cycle with the full QFT every 10th layer",
            "qubits": n,
            "layers": args.layers,
            "state_bytes": state_bytes,
            "parallelism": "single GPU",
        },
        "pct_hbm_peak": round(100.0 * value / PEAK_GBPS, 1),
        "layers": {k: {"ms": round(float(np.mean(v["ms"])), 3), "S": max(v["S"]),
                       "GBps": (round(max(v["S"]) * sweep_bytes / (np.mean(v["ms"]) * 1e-3) / 1e9, 1)
                                if max(v["S"]) else None),
                       "pct_peak": (round(100 * max(v["S"]) * sweep_bytes / (np.mean(v["ms"]) * 1e-3) / 1e9
                                          / PEAK_GBPS, 1) if max(v["S"]) else None)}
                   for k, v in per_type.items()},
        "sweeps_per_step": layer_sweeps_per_step,
        "layer_equivalent_GBps": round(s_min_per_step * sweep_bytes / (ms_per_step * 1e-3) / 1e9, 1),
        "note": ("S = sweeps a layer launches. S = 0: the layer's work is deferred into the next "
                 "layer's first sweep (an RZZ ring and a CNOT chain are read through by the first sweep "
                 "of the following U3 layer or QFT: u3_fused / qft_fused), or, for the first U3 layer "
                 "(u3_on_basis), written as the product state of the reset |0...0> in one write pass. "
                 "A U3 layer leaves its last sweep's rotations pending; the next U3 layer runs them in "
                 "its first sweep across the deferred ring and chain (u3_cross), a QFT runs them first "
                 "(one sweep, inside qft_fused). value counts only the sweeps performed; "
                 "layer_equivalent_GBps prices each layer at a fixed count of sweeps (diagonal 1, "
                 "XOR 1, U3 2, QFT 2) over the same step time, so it rises as the step gets faster"),
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_GBPS, 4),
            "traffic": traffic,
            "traffic_source": traffic_src,
            "avg_launch_ms": round(tile_avg_ms, 4),
            "algorithmic_bytes_per_launch": sweep_bytes,
        },
        "kernels": kernels,
        "copy_probe_GBps": round(copy_gbps, 1),
        "cpu_baseline": cpu,
        **extras,
    }
    line = json.dumps(out)
    if args.json_out:
        with open(args.json_out, "w") as fh:
            fh.write(line + "\n")
    print(line, flush=True)
    state.close()


if __name__ == "__main__":
    main()
