"""Benchmark of the HIP Monte Carlo SEIR step loop on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload config3|config1|config2|config4]

A "step" is one full simulation of the workload: reset (initial draw) plus
all T iterations of the step loop (exposure / probability / draw kernel and
timer-update kernel per iteration) over every replica, and the per-iteration
SEIR counts.  Inputs (packed contact slices, state) are resident in HBM
before timing; ingestion is excluded.  value = sum_k N_k * M * T * K / wall
seconds, in replica*node*steps/s.

config3 (default): 64 uneven networks.  The batched form (one context, one
step launch for all networks, the T launches captured in a hipGraph, kernel
variant picked by the autotuner) is timed against the plain per-network
launch loop (one context per network, eager per-iteration launches, one
network at a time) on the same GPU; both must give identical counts.
config1/config2: one network, autotuned variant.  config4: the
differentiable marginal form, forward + adjoint per calibration step.

Prints one JSON line with a roofline object (algorithmic bytes per
iteration -- the window SpMM's bytes in fused form, from the window's
nonzeros counted on the device -- over the HIP-event time per iteration;
measured HBM traffic where a committed rocprofv3 summary holds it) and a
CPU baseline: the oracle (a C
restatement of the reference --cpu-only loop) timed on a bounded sample of
the same workload on the host cores.
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
# MI355X vector fp32 adds without packing: 256 CUs x 4 SIMDs x 16 lanes x
# 2.4 GHz = 39.3 T lane-adds/s (the published 157.3 TFLOP/s counts packed FMAs)
VALU_LANE_ADDS_T = 39.3


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model() -> str:
    """The host CPU's model name and logical CPU count, from /proc/cpuinfo."""
    try:
        names = [ln.split(":", 1)[1].strip() for ln in open("/proc/cpuinfo")
                 if ln.startswith("model name")]
        return f"{names[0]} ({len(names)} logical CPUs visible)" if names else "unknown"
    except OSError:
        return "unknown"


def cpu_baseline(name, n_nodes, src, tgt, start, end, p, threads: int, M_s: int, T_s: int,
                 note: str = "") -> dict:
    """Oracle on the host: same network, M and T truncated to a bounded
    sample, with `threads` OpenMP threads (the reference's default is every
    core, cu:716-723) and again with one thread (SURVEY's baseline plan asks
    for both, with the CPU model)."""
    from oracle import oracle_py
    if not os.path.exists(oracle_py.LIB_PATH):
        oracle_py.build()
    M_s = min(p.M, M_s)
    T_s = min(p.iterations, T_s)
    g = oracle_py.OracleGraph.from_contacts(n_nodes, src, tgt, start, end)

    def timed(nt):
        op = oracle_py.make_params(M=M_s, step_size=p.step_size, iterations=T_s,
                                   initial_infected=p.initial_infected,
                                   infect_prob=p.infect_prob,
                                   exposed_duration=p.exposed_duration,
                                   infectious_duration=p.infectious_duration,
                                   resistant_duration=p.resistant_duration,
                                   start_time=p.start_time,
                                   static_network_duration=p.static_network_duration,
                                   seed=p.seed, threads=nt, use_libm_exp=1)
        return g.simulate(op)[2]

    secs = timed(threads)
    secs1 = timed(1)
    work = n_nodes * M_s * T_s
    return {"value": work / secs, "unit": "replica*node*steps/s",
            "cores": threads, "kind": "port",
            "value_1_thread": work / secs1,
            "cpu_model": cpu_model(),
            "sample": f"{name}: {note}N={n_nodes}, M={M_s} of {p.M}, first {T_s} of "
                      f"{p.iterations} iterations; loop {secs:.2f} s on {threads} threads, "
                      f"{secs1:.2f} s on 1 (glibc expf, OpenMP rows when nnz > 20000)"}


def measured_traffic(workload: str):
    """HBM bytes per step launch measured by rocprofv3 --pmc for this workload
    (FETCH_SIZE and WRITE_SIZE, scaled by a copy of known size), from the
    committed profile summary; None when there is none."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def mean_nnz(sim, T: int, samples: int = 64):
    """Mean nonzeros of the window CSR over the run's iterations (every
    iteration, or `samples` evenly spaced ones when there are more than 256),
    counted on the device; and which iterations were counted.  The count
    walks each row with one thread, which takes seconds per window on hub
    rows, hence the sample."""
    if T <= 256:
        ks = list(range(T))
    else:
        ks = sorted({round(i * (T - 1) / (samples - 1)) for i in range(samples)})
    nnz = sum(sim.slice_nnz(k) for k in ks) / len(ks)
    return nnz, ("every iteration" if T <= 256 else f"{len(ks)} evenly spaced iterations")


def roofline(info: dict, per_launch_s: float, workload: str, nnz: float, nnz_note: str) -> dict:
    """Algorithmic bytes of one iteration by the SpMM definition in its fused
    form (the N x M exposure matrix is never written): 8 nnz for the window's
    CSR + 4 (N + 1) row pointers + N M / 8 for the infectious-mask read + the
    state read and written.  The state term is step_bytes less its entry
    term; step_bytes (what this design streams, the slice epoch's packed
    entries at entry_bytes each) is reported beside it."""
    design = float(info["step_bytes"])
    ent = info["n_entries"] / max(info["n_epochs"], 1)
    state = design - int(info["entry_bytes"] * ent)
    b = 8.0 * nnz + state
    achieved = b / per_launch_s / 1e9
    tr = measured_traffic(workload)
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": tr["hbm_bytes_per_step"] if tr else None,
            "traffic_source": tr["source"] if tr else None,
            "kernel": "step (expo_kernel + update_kernel)",
            "bytes_per_launch": b, "us_per_launch": per_launch_s * 1e6,
            "bytes_definition": "SpMM bytes of the window in fused form: 8 nnz (window CSR "
                                f"nonzeros, mean over {nnz_note}, counted on the device) + "
                                "4 (N+1) row pointers + N*M/8 infectious-mask read + state "
                                "read and written per 64-replica word (susceptible read+write, "
                                "next infectious write, infection word write, one ring read "
                                "per timer offset) + per-replica counts",
            "nnz_per_iteration": nnz,
            "design_bytes_per_launch": design,
            "design_bytes_definition": "the same state terms plus the slice epoch's packed "
                                       f"contact entries at {info['entry_bytes']} bytes each, "
                                       "which the exposure kernel streams every iteration",
            "design_achieved": design / per_launch_s / 1e9,
            # the second roof: the SpMM's nnz * M masked adds against the
            # vector ALUs' non-packed fp32 lane-add rate (256 CUs x 4 SIMD x
            # 16 lanes x 2.4 GHz); this design counts them in bit planes
            "valu_roof": {"ops_per_launch": nnz * info["M"],
                          "achieved": nnz * info["M"] / per_launch_s / 1e12,
                          "peak": VALU_LANE_ADDS_T, "unit": "T lane-adds/s",
                          "frac": nnz * info["M"] / per_launch_s / 1e12 / VALU_LANE_ADDS_T,
                          "note": "adds of the reference's fp32 SpMM; this design counts 64 "
                                  "replicas per bit-plane operation, so the fraction can "
                                  "exceed 1"}}


def time_runs(sim, T: int, steps: int, warmup: int):
    for _ in range(warmup):
        sim.reset()
        sim.run()
    sim.sync()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.reset()
        kernel_ms.append(sim.run_timed(T))
    sim.sync()
    return time.perf_counter() - t0, kernel_ms


def bench_single(args, w) -> dict:
    from replay_amd import Simulation
    t0 = time.time()
    sim = Simulation.from_contacts(w.n_nodes, w.src, w.tgt, w.start, w.end, w.params)
    sim.autotune(min(w.params.iterations, 48))
    info = sim.info()
    N, M, T = info["n_nodes"], info["M"], w.params.iterations
    log(f"{w.name}: N={N} M={M} T={T} contacts={info['n_contacts']} windows={info['n_windows']} "
        f"epochs={info['n_epochs']} entries={info['n_entries']} open {time.time() - t0:.1f}s")
    nnz, nnz_note = mean_nnz(sim, T)
    wall, kernel_ms = time_runs(sim, T, args.steps, args.warmup)
    counts = sim.counts()
    sim.close()
    per_launch = (sum(kernel_ms) / len(kernel_ms)) / 1e3 / T
    out = {
        "metric": "replica*node*steps/s (plain step loop, one network)",
        "value": N * M * T * args.steps / wall,
        "ms_per_step": wall * 1e3 / args.steps,
        "config": {"workload": w.name, "nodes": N, "replicas": M, "iterations": T,
                   "contacts": info["n_contacts"], "step_s": w.params.step_size,
                   "window_s": w.params.static_network_duration,
                   "windows_per_epoch": info["windows_per_epoch"], "variant": info["variant"],
                   "final_recovered_mean": float(counts[-1, :, 2].mean())},
        "roofline": roofline(info, per_launch, w.name, nnz, nnz_note),
    }
    if not args.no_cpu_baseline:
        # bounded samples of about 10-30 s of oracle work (replicas, iterations)
        M_s, T_s = {"config0": (100, 168), "config1": (64, 240), "config2": (32, 6)}.get(
            args.workload, (32, 16))
        out["cpu_baseline"] = cpu_baseline(w.name, w.n_nodes, w.src, w.tgt, w.start, w.end,
                                           w.params, min(16, os.cpu_count() or 1), M_s, T_s)
    return out


def dense_record(args) -> dict:
    """The high-density workload (50k nodes, 40M contacts, 500 power-law
    hubs, 4096 replicas, 4032 iterations of 5 min) timed like the headline
    one: reset + all iterations replayed from one hipGraph, inputs resident,
    after an untimed warm-up; a sub-record of the default line."""
    import replay_amd
    from replay_amd import Params, Simulation, workloads
    t0 = time.time()
    w = workloads.config2_dense()
    p = Params(**{**w.params.__dict__, "flags": replay_amd.F_GRAPH})
    sim = Simulation.from_contacts(w.n_nodes, w.src, w.tgt, w.start, w.end, p)
    variant = sim.autotune(48)
    info = sim.info()
    N, M, T = info["n_nodes"], info["M"], w.params.iterations
    nnz, nnz_note = mean_nnz(sim, T, samples=16)
    log(f"{w.name}: N={N} M={M} T={T} contacts={info['n_contacts']} variant={variant} "
        f"ring_form={info['ring_form']} state {info['state_bytes'] / 1e9:.2f} GB, "
        f"ready in {time.time() - t0:.1f}s")
    steps = max(1, min(args.steps, 2))
    wall, kernel_ms = time_runs(sim, T, steps, 1)
    counts = sim.counts()
    sim.close()
    per_launch = (sum(kernel_ms) / len(kernel_ms)) / 1e3 / T
    rf = roofline(info, per_launch, w.name, nnz, nnz_note)
    log(f"dense: {wall * 1e3 / steps:.1f} ms per run, {N * M * T * steps / wall:.3e} r*n*s/s")
    return {"workload": w.name, "value": N * M * T * steps / wall,
            "unit": "replica*node*steps/s", "steps": steps, "warmup": 1,
            "ms_per_step": wall * 1e3 / steps, "nodes": N, "replicas": M, "iterations": T,
            "contacts": info["n_contacts"], "variant": variant, "hipgraph": True,
            "ring_form": ["offsets", "compact", "history"][info["ring_form"]],
            "state_bytes": info["state_bytes"],
            "final_recovered_mean": float(counts[-1, :, 2].mean()),
            "roofline": {k: rf[k] for k in ("bound", "achieved", "peak", "unit", "frac",
                                            "bytes_per_launch", "us_per_launch",
                                            "nnz_per_iteration")},
            "roofline_note": f"bytes per iteration by the fused SpMM definition of the main "
                             f"line; window nonzeros from {nnz_note}"}


def bench_batched(args, w) -> dict:
    import numpy as np
    import replay_amd
    from replay_amd import Params, Simulation
    sp = w.params
    T, M = sp.iterations, sp.M
    t0 = time.time()
    bp = Params(**{**sp.__dict__, "flags": 0 if args.eager else replay_amd.F_GRAPH})
    batch = Simulation.from_networks(w.networks, bp)
    variant = batch.autotune(min(T, 24))
    info = batch.info()
    N = info["n_nodes"]
    nnz, nnz_note = mean_nnz(batch, T)
    log(f"{w.name}: {len(w.networks)} networks, N={N} M={M} T={T} contacts={info['n_contacts']} "
        f"epochs={info['n_epochs']} entries={info['n_entries']} variant={variant} "
        f"open {time.time() - t0:.1f}s")
    wall, kernel_ms = time_runs(batch, T, args.steps, args.warmup)
    bcounts = batch.counts()
    per_launch = (sum(kernel_ms) / len(kernel_ms)) / 1e3 / T
    batched_value = N * M * T * args.steps / wall
    log(f"batched: {wall * 1e3 / args.steps:.2f} ms per run, {batched_value:.3e} r*n*s/s")
    batch.close()
    out = {
        "metric": "replica*node*steps/s (batched persistent form; plain per-network loop beside)",
        "value": batched_value,
        "ms_per_step": wall * 1e3 / args.steps,
        "config": {"workload": w.name, "networks": len(w.networks), "nodes": N, "replicas": M,
                   "iterations": T, "contacts": info["n_contacts"], "step_s": sp.step_size,
                   "windows_per_epoch": info["windows_per_epoch"], "variant": variant,
                   "hipgraph": not args.eager,
                   "final_recovered_mean": float(bcounts[-1, :, :, 2].mean())},
        "roofline": roofline(info, per_launch, w.name, nnz, nnz_note),
    }
    if args.no_plain:
        return out

    # plain per-network launch loop: one context per network, eager launches
    t0 = time.time()

    def open_plain(nw):
        pp = Params(M=M, step_size=sp.step_size, iterations=T,
                    initial_infected=nw.initial_infected, infect_prob=nw.infect_prob,
                    exposed_duration=nw.exposed_duration,
                    infectious_duration=nw.infectious_duration,
                    resistant_duration=nw.resistant_duration, start_time=sp.start_time,
                    static_network_duration=sp.static_network_duration, seed=sp.seed,
                    stream=nw.stream, windows_per_epoch=info["windows_per_epoch"])
        return Simulation.from_contacts(nw.n_nodes, nw.src, nw.tgt, nw.start, nw.end, pp)

    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=8) as ex:
        plains = list(ex.map(open_plain, w.networks))
    log(f"plain contexts open {time.time() - t0:.1f}s")
    for s in plains:  # warmup
        s.reset()
        s.run()
    for s in plains:
        s.sync()
    plain_steps = max(1, min(args.steps, 2))
    t0 = time.perf_counter()
    for _ in range(plain_steps):
        for s in plains:  # one network at a time, as the reference runs them
            s.reset()
            s.run()
            s.sync()
    plain_wall = time.perf_counter() - t0
    plain_value = N * M * T * plain_steps / plain_wall
    identical = all(np.array_equal(s.counts(), bcounts[:, k]) for k, s in enumerate(plains))
    for s in plains:
        s.close()
    log(f"plain: {plain_wall * 1e3 / plain_steps:.2f} ms per run, {plain_value:.3e} r*n*s/s; "
        f"identical counts: {identical}")
    out["config"].update({"plain_value": plain_value,
                          "plain_ms_per_step": plain_wall * 1e3 / plain_steps,
                          "batched_over_plain": batched_value / plain_value,
                          "batched_equals_plain": identical})
    if not args.no_cpu_baseline:
        nw = w.networks[0]
        p0 = Params(M=M, step_size=sp.step_size, iterations=T,
                    initial_infected=nw.initial_infected, infect_prob=nw.infect_prob,
                    exposed_duration=nw.exposed_duration,
                    infectious_duration=nw.infectious_duration,
                    resistant_duration=nw.resistant_duration, start_time=sp.start_time,
                    static_network_duration=sp.static_network_duration, seed=sp.seed)
        out["cpu_baseline"] = cpu_baseline(w.name, nw.n_nodes, nw.src, nw.tgt, nw.start, nw.end,
                                           p0, min(16, os.cpu_count() or 1), 32, 48,
                                           note="network 0 of 64, ")
    return out


def bench_marginal(args) -> dict:
    """Differentiable marginal form: forward + adjoint of an MSE loss against
    a target curve, full workload size, timed per calibration step."""
    import torch
    from replay_amd import Simulation, workloads
    from replay_amd.marginal import MarginalEngine, marginal_seir
    w = workloads.config4_marginal()
    g = w.graph
    t0 = time.time()
    sim = Simulation.from_contacts(g.n_nodes, g.src, g.tgt, g.start, g.end, g.params)
    info = sim.info()
    N, T, B = info["n_nodes"], info["iterations"], int(w.beta.size)
    dev = "cuda:0"
    eng = MarginalEngine(sim, B, checkpoint_every=32)
    one = MarginalEngine(sim, 1, checkpoint_every=32)
    phi = torch.ones(N, device=dev)
    truth = [torch.tensor([v], dtype=torch.float32, device=dev) for v in w.truth]
    target = one.forward(*truth, phi)[:, 0, 2].clone()
    log(f"{w.name}: N={N} T={T} sets={B} contacts={info['n_contacts']} open {time.time() - t0:.1f}s")
    xs = [torch.tensor(v, device=dev, requires_grad=True) for v in (w.beta, w.sigma, w.gamma)]
    phi_p = torch.ones(N, device=dev, requires_grad=True)

    def step():
        for x in xs + [phi_p]:
            x.grad = None
        C = marginal_seir(eng, *xs, phi_p)
        loss = ((C[:, :, 2] - target[:, None]) ** 2).mean()
        loss.backward()
        return loss

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    t0 = time.perf_counter()
    fwd_ms, bwd_ms = [], []
    for _ in range(args.steps):
        for x in xs + [phi_p]:
            x.grad = None
        ev[0].record()
        C = marginal_seir(eng, *xs, phi_p)
        loss = ((C[:, :, 2] - target[:, None]) ** 2).mean()
        ev[1].record()
        loss.backward()
        ev[2].record()
        torch.cuda.synchronize()
        fwd_ms.append(ev[0].elapsed_time(ev[1]))
        bwd_ms.append(ev[1].elapsed_time(ev[2]))
    wall = time.perf_counter() - t0
    fwd = sum(fwd_ms) / len(fwd_ms)
    bwd = sum(bwd_ms) / len(bwd_ms)
    log(f"marginal: forward {fwd:.2f} ms, backward {bwd:.2f} ms per calibration step; "
        f"loss {loss.item():.4g}")
    # forward kernel: per step, the three state planes s, e, iota read and
    # written (fp32, 64 sets padded per node; r = 1 - s - e - iota), the epoch's
    # entries (8 bytes where the context holds the 8-byte copy, which the
    # forward kernel then streams) and row pointers
    Bp = (B + 63) // 64 * 64
    ent = info["n_entries"] / max(info["n_epochs"], 1)
    eb = 8 if info.get("has_entries8") else 16
    step_bytes = 24.0 * N * Bp + eb * ent + 4.0 * (N + 1)
    achieved = step_bytes / (fwd / 1e3 / T) / 1e9
    eng.close()
    one.close()
    sim.close()
    return {
        "metric": "node*step*set/s (marginal SEIR forward + adjoint per calibration step)",
        "value": N * T * B * args.steps / wall,
        "ms_per_step": wall * 1e3 / args.steps,
        "config": {"workload": w.name, "nodes": N, "iterations": T, "parameter_sets": B,
                   "contacts": info["n_contacts"], "step_s": g.params.step_size,
                   "checkpoint_every": 32, "forward_ms": fwd, "backward_ms": bwd,
                   "final_loss": float(loss.item())},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "mf_forward_kernel (+ per-step reduction)",
                     "bytes_per_launch": step_bytes,
                     "us_per_launch": fwd * 1e3 / T,
                     "bytes_definition": "per iteration: s, e, iota read and written as fp32 "
                                         "for every node and padded parameter set (r = 1 - s "
                                         "- e - iota is not stored); the slice "
                                         f"epoch's {eb}-byte contact entries and row pointers"},
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="config3")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true",
                    help="batched workload only: skip the plain per-network loop")
    ap.add_argument("--no-dense", action="store_true",
                    help="batched workload only: skip the high-density sub-record")
    ap.add_argument("--eager", action="store_true",
                    help="launch every step eagerly (no hipGraph), e.g. under rocprofv3 --pmc")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device three times (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 4, "copies": 3}))
        return 0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): replicas and networks "
              "are independent, so the path shards with no per-step exchange; multi-GPU runs "
              "are not part of this build.", file=sys.stderr)
        return 2
    from replay_amd import workloads
    t0 = time.time()
    if args.workload == "config4":
        res = bench_marginal(args)
        out = {"metric": res["metric"], "value": res["value"], "unit": "node*step*set/s",
               "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
               "ms_per_step": res["ms_per_step"], "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "f32", "data": "synthetic", "config": res["config"],
               "roofline": res["roofline"], "cpu_baseline": None}
        print(json.dumps(out))
        return 0
    if args.workload == "config3":
        w = workloads.config3_batched()
        log(f"generated {w.name} ({w.n_contacts} contacts) in {time.time() - t0:.1f}s")
        res = bench_batched(args, w)
    else:
        w = workloads.CONFIGS[args.workload]()
        log(f"generated {w.name} ({w.n_contacts} contacts) in {time.time() - t0:.1f}s")
        res = bench_single(args, w)
    out = {"metric": res["metric"], "value": res["value"], "unit": "replica*node*steps/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": res["ms_per_step"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic", "config": res["config"],
           "roofline": res["roofline"], "cpu_baseline": res.get("cpu_baseline")}
    if args.workload == "config3" and not args.no_dense:
        out["dense_workload"] = dense_record(args)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
