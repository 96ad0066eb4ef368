"""Seeded synthetic contact networks for the benchmark configurations.

Every generator returns a Workload: node count, contact arrays (ids, Unix
start/end seconds) and the simulation Params the configuration names.  The
reference generates its benchmark graphs with a closed binary
(generate_networks/, source graph_generator.c); these generators are this
project's own, seeded with numpy's PCG64, so every input is reproducible.

Mapping of rate parameters onto the reference's countdown model
(tool/gpu_cpu_temporal_sim.cu:727-729): only S->E is random, with probability
1-(1-beta)^(exposure_hours); E->I after 1/sigma, I->R after 1/gamma, both as
deterministic durations in seconds; "permanent" R is a resistant duration
longer than the horizon (kept below 2^24 s so countdowns stay exact in fp32).
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace

import numpy as np

from . import Params

T0 = 946713600  # 2000-01-01T00:00:00Z, the reference's default start (cu:55)
HOUR = 3600
DAY = 86400


@dataclass
class Workload:
    name: str
    n_nodes: int
    src: np.ndarray
    tgt: np.ndarray
    start: np.ndarray
    end: np.ndarray
    params: Params
    notes: dict = field(default_factory=dict)

    @property
    def n_contacts(self) -> int:
        return int(self.src.size)

    def write_graph(self, path: str) -> None:
        """Writes the reference text format (README.md graph format)."""
        import time as _time
        with open(path, "w") as f:
            f.write("NODE_LIST " + " ".join(str(i) for i in range(self.n_nodes)) + "\n")
            fmt = "%Y-%m-%dT%H:%M:%SZ"
            for a, b, s, e in zip(self.src.tolist(), self.tgt.tolist(), self.start.tolist(),
                                  self.end.tolist()):
                f.write(f"{a}, {b}, {_time.strftime(fmt, _time.gmtime(s))}, "
                        f"{_time.strftime(fmt, _time.gmtime(e))};\n")


def _pairs_uniform(rng, n, count):
    """Uniform unordered pairs u < v without self-loops."""
    a = rng.integers(0, n, size=count)
    b = rng.integers(0, n - 1, size=count)
    b = b + (b >= a)
    return np.minimum(a, b), np.maximum(a, b)


def config0_sir(seed: int = 0, n_nodes: int = 500, n_contacts: int = 20000, M: int = 100,
                iterations: int = 168) -> Workload:
    """Plain SIR, 500 nodes, 20k contacts over 7 days, 1 h steps, 100 runs.

    Durations Exp(15 min) capped at 8 h; beta 0.05 per contact hour;
    1/gamma = 72 h; 10 expected seeds as initial_infected = 10/N (the
    reference seeds per node with a probability, cu:84)."""
    rng = np.random.default_rng(seed)
    a, b = _pairs_uniform(rng, n_nodes, n_contacts)
    start = T0 + rng.integers(0, 7 * DAY, size=n_contacts)
    dur = np.minimum(np.rint(rng.exponential(15 * 60, size=n_contacts)), 8 * HOUR).astype(np.int64)
    p = Params(M=M, step_size=float(HOUR), iterations=iterations, initial_infected=10.0 / n_nodes,
               infect_prob=0.05, exposed_duration=0, infectious_duration=72 * HOUR,
               resistant_duration=14 * DAY, start_time=T0, static_network_duration=HOUR, seed=0)
    return Workload("config0_sir_500", n_nodes, a.astype(np.int32), b.astype(np.int32),
                    start.astype(np.int64), (start + dur).astype(np.int64), p)


def _diurnal_times(rng, count, days):
    """Start times with density proportional to 1 + sin(2 pi t / day)."""
    out = np.empty(0, np.int64)
    while out.size < count:
        t = rng.uniform(0, days * DAY, size=2 * (count - out.size) + 16)
        keep = rng.uniform(0, 2, size=t.size) < 1 + np.sin(2 * np.pi * t / DAY)
        out = np.concatenate([out, t[keep].astype(np.int64)])
    return out[:count]


def config1_village(seed: int = 0, n_nodes: int = 10000, block: int = 200,
                    n_contacts: int = 2_000_000, M: int = 1024, iterations: int = 4320,
                    days: int = 30) -> Workload:
    """Plain SEIR on a village block network: 50 blocks of 200, 2M contacts,
    90% within a block, diurnal starts over 30 days, LogNormal(ln 20 min, 1)
    durations, 10 min steps, beta 0.03, 1/sigma 48 h, 1/gamma 120 h."""
    rng = np.random.default_rng(seed)
    n_blocks = n_nodes // block
    intra = rng.uniform(size=n_contacts) < 0.9
    a = rng.integers(0, n_nodes, size=n_contacts)
    blk = a // block
    off = rng.integers(0, block - 1, size=n_contacts)
    own = a % block
    b_intra = blk * block + off + (off >= own)
    ob = rng.integers(0, n_blocks - 1, size=n_contacts)
    ob = ob + (ob >= blk)
    b_cross = ob * block + rng.integers(0, block, size=n_contacts)
    b = np.where(intra, b_intra, b_cross)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    start = T0 + _diurnal_times(rng, n_contacts, days)
    dur = np.maximum(1, np.rint(rng.lognormal(np.log(20 * 60), 1.0, size=n_contacts))).astype(np.int64)
    dur = np.minimum(dur, 2 * DAY)
    step = 600
    p = Params(M=M, step_size=float(step), iterations=iterations, initial_infected=10.0 / n_nodes,
               infect_prob=0.03, exposed_duration=48 * HOUR, infectious_duration=120 * HOUR,
               resistant_duration=days * DAY, start_time=T0, static_network_duration=step, seed=0)
    return Workload("config1_village_10k", n_nodes, lo.astype(np.int32), hi.astype(np.int32),
                    start.astype(np.int64), (start + dur).astype(np.int64), p)


def config2_dense(seed: int = 0, n_nodes: int = 50000, n_contacts: int = 40_000_000,
                  n_hubs: int = 500, M: int = 4096, iterations: int = 4032,
                  days: int = 14) -> Workload:
    """High-density stress: 50k nodes, 40M contacts, ~200 simultaneous
    neighbours per node, 500 hubs with Pareto(2.1) weights, 5 min steps over
    14 days, SEIR with the village parameters, 4096 replicas.

    Starts are uniform over [-2 d, 14 d) and durations Exp(48 h) capped at
    10 d, so about 40M * 2 d / 16 d = 5M contacts are active at a time."""
    rng = np.random.default_rng(seed)
    w = np.ones(n_nodes)
    hubs = rng.choice(n_nodes, size=n_hubs, replace=False)
    w[hubs] = (rng.pareto(1.1, size=n_hubs) + 1.0) * 4.0
    prob = w / w.sum()
    a = rng.choice(n_nodes, size=n_contacts, p=prob)
    b = rng.choice(n_nodes, size=n_contacts, p=prob)
    same = a == b
    while same.any():
        b[same] = rng.choice(n_nodes, size=int(same.sum()), p=prob)
        same = a == b
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    start = T0 + rng.integers(-2 * DAY, days * DAY, size=n_contacts)
    dur = np.minimum(np.rint(rng.exponential(48 * HOUR, size=n_contacts)), 10 * DAY).astype(np.int64)
    dur = np.maximum(dur, 1)
    step = 300
    p = Params(M=M, step_size=float(step), iterations=iterations, initial_infected=10.0 / n_nodes,
               infect_prob=0.03, exposed_duration=48 * HOUR, infectious_duration=120 * HOUR,
               resistant_duration=days * DAY + DAY, start_time=T0, static_network_duration=step,
               seed=0)
    return Workload("config2_dense_50k", n_nodes, lo.astype(np.int32), hi.astype(np.int32),
                    start.astype(np.int64), (start + dur).astype(np.int64), p)


@dataclass
class BatchWorkload:
    name: str
    networks: list          # list of replay_amd.Network
    params: Params          # shared: M, step, iterations, windows, start, seed
    notes: dict = field(default_factory=dict)

    @property
    def n_nodes(self) -> int:
        return sum(n.n_nodes for n in self.networks)

    @property
    def n_contacts(self) -> int:
        return sum(int(n.src.size) for n in self.networks)


def config3_batched(seed: int = 0, n_networks: int = 64, M: int = 256, iterations: int = 168,
                    n_min: int = 1000, n_max: int = 40000, deg_min: float = 20.0,
                    deg_max: float = 200.0, mean_days: float = 7.0) -> BatchWorkload:
    """Batched uneven workload: 64 networks with node counts log-uniform in
    1k-40k and 20-200 simultaneous neighbours, 256 replicas each, one-hour
    steps over a week, each with (beta, 1/sigma, 1/gamma) log-uniform within
    +-50% of the village configuration's (0.03, 48 h, 120 h).

    Contacts are uniform random pairs; starts are uniform over
    [t0 - mean, t0 + horizon) and durations Exp(mean) capped at 4x the mean,
    so about N * degree / 2 contacts are active at any time."""
    from . import Network
    rng = np.random.default_rng(seed)
    H = iterations * HOUR
    D = mean_days * DAY
    nets = []
    logu = lambda lo, hi: float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
    for k in range(n_networks):
        n = int(round(logu(n_min, n_max)))
        deg = logu(deg_min, deg_max)
        active = n * deg / 2.0
        n_c = int(round(active * (H + D) / D))
        a, b = _pairs_uniform(rng, n, n_c)
        start = T0 + rng.integers(-int(D), H, size=n_c)
        dur = np.minimum(np.rint(rng.exponential(D, size=n_c)), 4 * D).astype(np.int64)
        dur = np.maximum(dur, 1)
        start = np.maximum(start, 0)
        nets.append(Network(n_nodes=n, src=a.astype(np.int32), tgt=b.astype(np.int32),
                            start=start.astype(np.int64), end=(start + dur).astype(np.int64),
                            initial_infected=0.001,
                            infect_prob=0.03 * logu(0.5, 1.5),
                            exposed_duration=int(48 * HOUR * logu(0.5, 1.5)),
                            infectious_duration=int(120 * HOUR * logu(0.5, 1.5)),
                            resistant_duration=int(H + DAY), stream=k))
    p = Params(M=M, step_size=float(HOUR), iterations=iterations, start_time=T0,
               static_network_duration=HOUR, seed=0)
    return BatchWorkload("config3_batched_64nets", nets, p)


@dataclass
class MarginalWorkload:
    """A contact network and iteration plan for the marginal-probability form,
    with the parameter sets of one batch and the parameters of its target."""
    name: str
    graph: Workload         # contacts and the iteration / window plan
    beta: np.ndarray        # (B,) per hour of contact
    sigma: np.ndarray       # (B,) per hour
    gamma: np.ndarray       # (B,) per hour
    truth: tuple            # (beta, sigma, gamma) that generate the target curve


def config4_marginal(seed: int = 0, n_nodes: int = 20000, block: int = 200, days: int = 10,
                     iterations: int = 720, n_sets: int = 64) -> MarginalWorkload:
    """Differentiable marginal SEIR: the village block recipe at 20k nodes
    (100 blocks of 200, 90% of contacts within a block, diurnal starts,
    LogNormal(ln 20 min, 1) durations) at the village contact rate over 10
    days, 720 steps of 20 min, and 64 parameter sets log-uniform within +-50%
    of (beta, sigma, gamma) = (0.03 per contact hour, 1/48 h, 1/120 h), the
    parameters of the target curve."""
    per_node_day = 2_000_000 / 10000 / 30  # contacts per node and day of the village recipe
    g = config1_village(seed=seed, n_nodes=n_nodes, block=block,
                        n_contacts=int(round(per_node_day * n_nodes * days)), M=64,
                        iterations=iterations, days=days)
    step = 1200
    g = replace(g, name="config4_marginal_20k",
                params=replace(g.params, step_size=float(step), static_network_duration=step))
    rng = np.random.default_rng(seed + 1)
    logu = lambda c: (c * np.exp(rng.uniform(np.log(0.5), np.log(1.5), size=n_sets))).astype(np.float32)
    truth = (0.03, 1.0 / 48.0, 1.0 / 120.0)
    return MarginalWorkload("config4_marginal_20k", g, logu(truth[0]), logu(truth[1]),
                            logu(truth[2]), truth)


def scaled(w: Workload, **kw) -> Workload:
    return replace(w, params=replace(w.params, **kw))


CONFIGS = {"config0": config0_sir, "config1": config1_village, "config2": config2_dense}
