#!/usr/bin/env python3
"""Benchmark: structured-layer HBM throughput at 30 qubits, complex64.

Workload (one "step"): reset to |0...0> and run the 40-layer 30-qubit circuit
of mettleq_amd.workloads.layered_30q (U3-all / RZZ ring / CNOT chain cycle,
full QFT every 10th layer) on one GPU, inputs resident in HBM.

value = sum over layers of (bytes moved by the layer's full-state sweeps)
        / (GPU time of the timed steps), in GB/s; one sweep moves
        2 * 2^n * 8 bytes (read + write of the complex64 state).
roofline = the tile-sweep kernel (1q layers and QFT, the dominant kernel):
        algorithmic bytes per launch / mean launch duration from HIP events.
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
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-oracle", action="store_true", help="skip the complex128 fidelity check")
    ap.add_argument("--json-out", default=None)
    ap.add_argument("--traffic-from", default=os.path.join(ROOT, "profiles", "k_tile_30q_layered_circuit_pmc.md"),
                    help="rocprofv3 FETCH_SIZE/WRITE_SIZE summary of the tile kernel (HBM bytes per launch)")
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
    """HBM bytes per k_tile launch from a committed rocprofv3 PMC summary: the
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
    }


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: every box in this setup has a single MI355X and the "
              "engine's multi-GPU sharding is a design (see DESIGN.md), not a measured path. "
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

    # per-layer breakdown (one instrumented pass)
    state.reset()
    state.sync()
    per_type = {}
    for layer in layers:
        s0 = state.sweep_count()
        state.timer_start()
        apply_layer(state, layer, n, lay)
        ms = state.timer_stop()
        S = state.sweep_count() - s0
        d = per_type.setdefault(layer["type"], {"ms": [], "S": S})
        d["ms"].append(ms)
    log("per-layer: " + ", ".join(f"{k}: {np.mean(v['ms']):.3f} ms S={v['S']}" for k, v in per_type.items()))

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
    # value: bytes of all sweeps (the reset init writes the state once: 1/2 sweep)
    layer_sweeps_per_step = sum(per_type[l["type"]]["S"] for l in layers)
    layer_ms_per_step = sum(np.mean(per_type[l["type"]]["ms"]) for l in layers)
    value = layer_sweeps_per_step * sweep_bytes / (ms_per_step * 1e-3) / 1e9

    # dominant kernel: the tile sweep (u3 layers and QFT)
    tile_ms = sum(np.sum(per_type[t]["ms"]) for t in ("u3", "qft") if t in per_type)
    tile_launches = sum(per_type[t]["S"] * len(per_type[t]["ms"]) for t in ("u3", "qft") if t in per_type)
    tile_avg_ms = tile_ms / max(1, tile_launches)
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
        "layers": {k: {"ms": round(float(np.mean(v["ms"])), 3), "S": v["S"],
                       "GBps": round(v["S"] * sweep_bytes / (np.mean(v["ms"]) * 1e-3) / 1e9, 1),
                       "pct_peak": round(100 * v["S"] * sweep_bytes / (np.mean(v["ms"]) * 1e-3) / 1e9 / PEAK_GBPS, 1)}
                   for k, v in per_type.items()},
        "roofline": {
            "bound": "hbm",
            "kernel": "k_tile (1q-layer and QFT sweeps)",
            "achieved": round(achieved, 1),
            "peak": PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_GBPS, 4),
            "traffic": traffic,
            "traffic_source": traffic_src,
            "avg_launch_ms": round(tile_avg_ms, 4),
            "algorithmic_bytes_per_launch": sweep_bytes,
        },
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
