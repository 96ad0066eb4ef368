"""Benchmark of the HIP Monte Carlo SEIR step loop on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload config1]

A "step" is one full simulation of the workload: reset (initial draw) plus
all T iterations of the step loop (fused exposure / draw / countdown / counts
kernel per iteration) over M replicas.  Inputs (packed contact slices, state)
are resident in HBM before timing; ingestion is excluded, per-step counts are
included.  value = N * M * T * K / wall seconds, in replica*node*steps/s.

Prints one JSON line (the driver's contract) with a roofline object for the
step kernel (algorithmic bytes per launch / HIP-event time per launch) and a
CPU baseline: the oracle (a C restatement of the reference --cpu-only loop)
timed on a bounded sample of the same workload on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def step_kernel_bytes(info: dict) -> float:
    """Algorithmic HBM bytes of one step-kernel launch, as the library
    defines them (replay_info.step_bytes): per 64-replica word the
    susceptible word read + write, the own infectious word read, the next
    infectious word write, the iteration's infection word write and one ring
    read per timer offset; the slice epoch's entries (16 B) and row pointers;
    the per-replica prefix update (12 B)."""
    return float(info["step_bytes"])


def cpu_baseline(w, threads: int) -> dict:
    """Oracle on the host: same network, M and T truncated to a bounded sample."""
    from oracle import oracle_py
    if not os.path.exists(oracle_py.LIB_PATH):
        oracle_py.build()
    p = w.params
    M_s = min(p.M, 64)
    T_s = min(p.iterations, 240)
    g = oracle_py.OracleGraph.from_contacts(w.n_nodes, w.src, w.tgt, w.start, w.end)
    op = oracle_py.make_params(M=M_s, step_size=p.step_size, iterations=T_s,
                               initial_infected=p.initial_infected, infect_prob=p.infect_prob,
                               exposed_duration=p.exposed_duration,
                               infectious_duration=p.infectious_duration,
                               resistant_duration=p.resistant_duration, start_time=p.start_time,
                               static_network_duration=p.static_network_duration, seed=p.seed,
                               threads=threads, use_libm_exp=1)
    _, _, secs = g.simulate(op)
    return {"value": w.n_nodes * M_s * T_s / secs, "unit": "replica*node*steps/s",
            "cores": threads, "kind": "port",
            "sample": f"{w.name}: N={w.n_nodes}, M={M_s} of {p.M}, first {T_s} of "
                      f"{p.iterations} iterations; loop {secs:.2f} s (glibc expf, OpenMP rows "
                      f"when nnz > 20000)"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="config1")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): replicas and networks "
              "are independent, so the path shards with no per-step exchange; multi-GPU runs "
              "are not part of this build.", file=sys.stderr)
        return 2

    from replay_amd import Simulation, workloads
    t0 = time.time()
    w = workloads.CONFIGS[args.workload]()
    gen_s = time.time() - t0
    t0 = time.time()
    sim = Simulation.from_contacts(w.n_nodes, w.src, w.tgt, w.start, w.end, w.params)
    open_s = time.time() - t0
    info = sim.info()
    N, M, T = info["n_nodes"], info["M"], w.params.iterations
    print(f"[bench] {w.name}: N={N} M={M} T={T} contacts={info['n_contacts']} "
          f"windows={info['n_windows']} epochs={info['n_epochs']} entries={info['n_entries']} "
          f"gen {gen_s:.1f}s open {open_s:.1f}s", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        sim.reset()
        sim.run()
    sim.sync()
    kernel_ms = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        sim.reset()
        kernel_ms.append(sim.run_timed(T))
    sim.sync()
    wall = time.perf_counter() - t_start
    counts = sim.counts()
    sim.close()

    value = N * M * T * args.steps / wall
    per_launch_s = (sum(kernel_ms) / len(kernel_ms)) / 1e3 / T
    bytes_launch = step_kernel_bytes(info)
    achieved = bytes_launch / per_launch_s / 1e9
    out = {
        "metric": "replica*node*steps/s (plain fused step loop, one network)",
        "value": value,
        "unit": "replica*node*steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": w.name, "nodes": N, "replicas": M, "iterations": T,
                   "contacts": info["n_contacts"], "step_s": w.params.step_size,
                   "window_s": w.params.static_network_duration,
                   "windows_per_epoch": info["windows_per_epoch"],
                   "final_recovered_mean": float(counts[-1, :, 2].mean())},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "step_kernel", "bytes_per_launch": bytes_launch,
                     "us_per_launch": per_launch_s * 1e6},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(w, threads)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
