"""HIP engine vs the CPU oracle, through the C ABI (needs an MI355X).

Bar: bit-exact for slices (CSR of every window), countdown states and counts;
exposure bit-exact against the sequential fp32 order and within 1e-6 relative
of FP64.  Inputs are seeded; the same Philox key drives both sides."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, graph_path

import replay_amd
from replay_amd import Params, Simulation, workloads

pytestmark = pytest.mark.gpu

FIXTURES = ["readme_example.txt", "edge_cases.txt", "contactgraphgen_graph.txt",
            "random_flat.txt", "random_nested.txt", "random_dense.txt"]


def _to_oracle_params(oracle, p: Params, uniforms=None, threads=1):
    return oracle.make_params(M=p.M, step_size=p.step_size, iterations=p.iterations,
                              initial_infected=p.initial_infected, infect_prob=p.infect_prob,
                              exposed_duration=p.exposed_duration,
                              infectious_duration=p.infectious_duration,
                              resistant_duration=p.resistant_duration, start_time=p.start_time,
                              static_network_duration=p.static_network_duration, seed=p.seed,
                              threads=threads, uniforms=uniforms)


def _graph_start(path):
    import calendar
    import time
    ts = []
    for line in open(path).read().splitlines()[1:]:
        parts = [x.strip() for x in line.split(",")]
        if len(parts) >= 3:
            try:
                ts.append(calendar.timegm(time.strptime(parts[2][:19], "%Y-%m-%dT%H:%M:%S")))
            except ValueError:
                pass
    t = min(ts)
    return t - t % 600


def _random_case(seed, n=40, e=600, span=2 * 86400, long_frac=0.2):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, n, e)
    b = (a + 1 + rng.integers(0, n - 1, e)) % n
    s = 946713600 + rng.integers(0, span, e)
    d = np.where(rng.uniform(size=e) < long_frac, rng.integers(3600, span, e),
                 rng.integers(0, 5400, e))
    dup = rng.uniform(size=e) < 0.05
    s[1:][dup[1:]] = s[:-1][dup[1:]]
    d[1:][dup[1:]] = d[:-1][dup[1:]]
    return n, a.astype(np.int32), b.astype(np.int32), s.astype(np.int64), (s + d).astype(np.int64)


STEP_WINDOWS = [(600.0, 600), (600.0, 1800), (3600.0, 3600), (1000.0, 3600), (450.0, 3600)]


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("step,D", STEP_WINDOWS[:3])
def test_slices_fixture_graphs(oracle, name, step, D):
    path = graph_path(name)
    p = Params(M=2, step_size=step, static_network_duration=D, iterations=48,
               start_time=_graph_start(path))
    g = oracle.OracleGraph.from_file(path)
    with Simulation.from_file(path, p) as sim:
        assert sim.info()["nlists"] == g.nlists
        op = _to_oracle_params(oracle, p)
        for k in range(p.iterations):
            rp, col, val = sim.slice(k)
            orp, ocol, oval, _ = g.step_csr(op, k)
            np.testing.assert_array_equal(rp, orp)
            np.testing.assert_array_equal(col, ocol)
            np.testing.assert_array_equal(val.view(np.int32), oval.view(np.int32))


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("step,D", STEP_WINDOWS)
@pytest.mark.parametrize("wpe", [1, 4, 0])
def test_slices_random(oracle, seed, step, D, wpe):
    n, a, b, s, e = _random_case(seed)
    p = Params(M=3, step_size=step, static_network_duration=D, iterations=120,
               windows_per_epoch=wpe)
    g = oracle.OracleGraph.from_contacts(n, a, b, s, e)
    op = _to_oracle_params(oracle, p)
    with Simulation.from_contacts(n, a, b, s, e, p) as sim:
        for k in range(0, p.iterations, 3):
            rp, col, val = sim.slice(k)
            orp, ocol, oval, _ = g.step_csr(op, k)
            np.testing.assert_array_equal(rp, orp)
            np.testing.assert_array_equal(col, ocol)
            np.testing.assert_array_equal(val.view(np.int32), oval.view(np.int32))


@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("step,D", [(600.0, 600), (1000.0, 3600)])
def test_exposure_parity(oracle, seed, step, D):
    import ctypes as C
    n, a, b, s, e = _random_case(seed, n=64, e=1500)
    p = Params(M=130, step_size=step, static_network_duration=D, iterations=60,
               initial_infected=0.3, infect_prob=0.2, exposed_duration=1200,
               infectious_duration=7200, resistant_duration=3600, seed=seed)
    g = oracle.OracleGraph.from_contacts(n, a, b, s, e)
    op = _to_oracle_params(oracle, p)
    P = lambda arr, t: arr.ctypes.data_as(C.POINTER(t))
    with Simulation.from_contacts(n, a, b, s, e, p) as sim:
        for k in range(0, p.iterations, 7):
            state = sim.state()
            got = sim.exposure(k)
            lower, medium = p.resistant_duration, p.infectious_duration + p.resistant_duration
            inf = ((state > lower) & (state <= medium)).astype(np.float32)
            rp, col, val, _ = g.step_csr(op, k)
            ref = np.zeros(n * p.M, np.float32)
            ref64 = np.zeros(n * p.M, np.float64)
            oracle.lib().or_spmm(P(rp, C.c_int32), P(col, C.c_int32), P(val, C.c_float), len(col),
                                 n, P(inf, C.c_float), P(ref, C.c_float), p.M, 1)
            oracle.lib().or_spmm_f64(P(rp, C.c_int32), P(col, C.c_int32), P(val, C.c_float), n,
                                     P(inf, C.c_float), P(ref64, C.c_double), p.M)
            np.testing.assert_array_equal(got.view(np.int32), ref.view(np.int32))
            np.testing.assert_allclose(got, ref64, rtol=1e-6, atol=0)
            sim.run(7 if k + 7 <= p.iterations else p.iterations - k)


CASES = [
    # (M, step, D, E, I, R, p0, beta)
    (1, 600.0, 600, 0, 7200, 0, 0.2, 0.5),            # SIS-like (R = 0)
    (63, 600.0, 600, 1800, 7200, 10 ** 6, 0.05, 0.3),  # SEIR, permanent R
    (64, 1000.0, 3600, 1200, 3600, 7200, 0.1, 0.4),    # SEIRS, fractional weights
    (65, 3600.0, 3600, 0, 86400, 86400, 0.1, 0.05),    # SIRS
    (130, 450.0, 3600, 900, 5400, 0, 0.1, 1.0),        # p = 1 (NaN for zero exposure)
    (200, 1234.5, 1800, 3000, 9000, 20000, 0.05, 0.2), # non-integer step
]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("seed", [0, 7])
def test_trajectory_parity(oracle, case, seed):
    M, step, D, Ed, Id, Rd, p0, beta = case
    n, a, b, s, e = _random_case(seed + 10, n=80, e=2500)
    p = Params(M=M, step_size=step, static_network_duration=D, iterations=150,
               initial_infected=p0, infect_prob=beta, exposed_duration=Ed,
               infectious_duration=Id, resistant_duration=Rd, seed=1000 + seed)
    g = oracle.OracleGraph.from_contacts(n, a, b, s, e)
    counts, state, _ = g.simulate(_to_oracle_params(oracle, p), want_state=True)
    with Simulation.from_contacts(n, a, b, s, e, p) as sim:
        sim.run(60)
        sim.run(p.iterations - 60)
        np.testing.assert_array_equal(sim.counts(), counts)
        np.testing.assert_array_equal(sim.state(), state)
        assert counts[:, :, 1].max() > 0


@pytest.mark.parametrize("flags", [0, replay_amd.F_WIDE_ENTRIES])
@pytest.mark.parametrize("M", [1024, 1100])
@pytest.mark.parametrize("step,D", [(600.0, 600), (1000.0, 3600)])
def test_every_variant_matches_oracle_wide(oracle, M, step, D, flags):
    """Every exposure variant against the oracle at 16 and 18 words per node,
    in integer (window = step) and float windows: W = 16 takes the
    whole-group kernels of NODE4 and NODE16 (16-byte neighbour-word loads),
    reading the 8-byte copy of the entries or, with F_WIDE_ENTRIES, the
    16-byte entries; W = 18 their partial last group."""
    n, a, b, s, e = _random_case(21, n=70, e=2000)
    p = Params(M=M, step_size=step, static_network_duration=D, iterations=50,
               initial_infected=0.1, infect_prob=0.3, exposed_duration=1200,
               infectious_duration=5400, resistant_duration=3600, seed=17, flags=flags)
    g = oracle.OracleGraph.from_contacts(n, a, b, s, e)
    counts, state, _ = g.simulate(_to_oracle_params(oracle, p, threads=8), want_state=True)
    assert counts[:, :, 1].max() > 0
    with Simulation.from_contacts(n, a, b, s, e, p) as sim:
        for v in replay_amd.VARIANTS_ALL:
            sim.set_variant(v)
            sim.reset()
            sim.run()
            np.testing.assert_array_equal(sim.counts(), counts, err_msg=f"variant {v}")
            np.testing.assert_array_equal(sim.state(), state, err_msg=f"variant {v}")


SWEEP_WINDOWS = [(600.0, 600), (600.0, 1800), (3600.0, 3600), (1000.0, 3600), (450.0, 3600),
                 (300.0, 600)]


@pytest.mark.parametrize("case", range(48))
def test_random_sweep_every_variant(oracle, case):
    """Random graphs of 12-160 nodes and 100-6000 contacts, 1-1024 replicas
    (one word to 16 whole words, ragged last words), integer and float
    windows, both entry widths, 0-60% long contacts and low to high infection
    probabilities: every exposure variant's counts and states equal the
    oracle's.  Each case draws its parameters from its own seed.  Beside
    test_gpu_fuzz (one forced variant per case, at most 5 words), this runs
    all four variants on each graph and reaches the whole 16-word groups."""
    rng = np.random.default_rng(1000 + case)
    seed = int(rng.integers(0, 1 << 30))
    n = int(rng.choice([12, 40, 90, 160]))
    e = int(rng.choice([100, 600, 2500, 6000]))
    M = int(rng.choice([1, 33, 64, 130, 256, 1024]))
    step, D = SWEEP_WINDOWS[int(rng.integers(0, len(SWEEP_WINDOWS)))]
    flags = int(rng.choice([0, replay_amd.F_WIDE_ENTRIES])) | \
        int(rng.choice([0, replay_amd.F_NO_HISTORY]))
    long_frac = float(rng.choice([0.0, 0.2, 0.6]))
    print(f"case {case}: seed={seed} n={n} e={e} M={M} step={step} D={D} flags={flags}", flush=True)
    nn, a, b, s, en = _random_case(seed, n=n, e=e, long_frac=long_frac)
    p = Params(M=M, step_size=step, static_network_duration=D, iterations=40,
               initial_infected=0.1, infect_prob=float(rng.choice([0.05, 0.3, 0.9])),
               exposed_duration=1200, infectious_duration=5400, resistant_duration=3600,
               seed=seed % 1000, flags=flags)
    g = oracle.OracleGraph.from_contacts(nn, a, b, s, en)
    counts, state, _ = g.simulate(_to_oracle_params(oracle, p, threads=8), want_state=True)
    with Simulation.from_contacts(nn, a, b, s, en, p) as sim:
        for v in replay_amd.VARIANTS_ALL:
            sim.set_variant(v)
            sim.reset()
            sim.run()
            np.testing.assert_array_equal(sim.counts(), counts, err_msg=f"variant {v}")
            np.testing.assert_array_equal(sim.state(), state, err_msg=f"variant {v}")


@pytest.mark.parametrize("M", [70, 256])
def test_injected_uniform_stream(oracle, M):
    n, a, b, s, e = _random_case(3, n=50, e=1200)
    p = Params(M=M, step_size=600.0, static_network_duration=1800, iterations=80,
               initial_infected=0.1, infect_prob=0.3, exposed_duration=1200,
               infectious_duration=7200, resistant_duration=3600)
    rng = np.random.default_rng(5)
    # a stream in [0, 1] with 15-bit resolution, like rand()/RAND_MAX on Windows
    u = (rng.integers(0, 32768, size=(p.iterations + 1) * p.M * n) / 32767.0).astype(np.float32)
    g = oracle.OracleGraph.from_contacts(n, a, b, s, e)
    counts, state, _ = g.simulate(_to_oracle_params(oracle, p, uniforms=u), want_state=True)
    with Simulation.from_contacts(n, a, b, s, e, p) as sim:
        sim.set_uniforms(u)
        sim.run()
        np.testing.assert_array_equal(sim.counts(), counts)
        np.testing.assert_array_equal(sim.state(), state)
        sim.set_uniforms(None)
        sim.run()
        c2 = sim.counts()
    counts2, _, _ = g.simulate(_to_oracle_params(oracle, p), want_state=False)
    np.testing.assert_array_equal(c2, counts2)


@pytest.mark.parametrize("step,D", [(600.0, 600), (1000.0, 3600)])
def test_draws_at_probability_ties(oracle, step, D):
    """Uniforms equal to the reference probability of the exposures that
    occur, and one float below it: every such draw sits on the decision
    boundary P > u, where the kernel's fast estimate of P cannot decide and
    the reference expression (cu:113-135) must.  Contacts cover the whole run,
    so an exposure is a sum of c equal window weights and its probability is
    one of a few values P_c, computed here in the reference's fp32 order.
    Both streams must give the oracle's trajectory bit for bit, and the two
    oracle trajectories must differ (the ties are hit)."""
    import ctypes as C
    libm = C.CDLL("libm.so.6")
    libm.logf.argtypes = [C.c_float]
    libm.logf.restype = C.c_float
    rng = np.random.default_rng(11)
    n, n_c = 64, 300
    a = rng.integers(0, n, n_c)
    b = (a + 1 + rng.integers(0, n - 1, n_c)) % n
    t0 = 946713600
    j = np.arange(n_c)
    s = (t0 - 5000 - j).astype(np.int64)  # distinct, none nested in another
    e = (t0 + 10 ** 6 - j).astype(np.int64)
    beta = 0.3
    p = Params(M=130, step_size=step, static_network_duration=D, iterations=40,
               initial_infected=0.08, infect_prob=beta, exposed_duration=0,
               infectious_duration=int(6 * step), resistant_duration=int(2 * step),
               start_time=t0)
    f32 = np.float32
    log_base = f32(libm.logf(f32(1.0) - f32(beta)))
    # window weights: D / steps_in_slice for every slice length that occurs
    weights = {f32(D)} | {f32(f32(D) * (f32(1.0) / f32(k))) for k in range(1, 6)}
    P = set()
    for w in weights:
        E = f32(0.0)
        for c in range(1, 40):
            E = f32(E + w)  # ascending-column fp32 sum of c equal weights
            x = f32(log_base * f32(E / f32(3600.0)))
            P.add(f32(f32(1.0) - f32(oracle.expf(float(x)))))
    P = np.array(sorted(P), np.float32)
    size = (p.iterations + 1) * p.M * n
    pick = P[rng.integers(0, P.size, size)]
    seed_u = rng.uniform(size=p.M * n).astype(np.float32)  # slot 0: u < initial_infected
    g = oracle.OracleGraph.from_contacts(n, a.astype(np.int32), b.astype(np.int32), s, e)
    results = []
    for below in (False, True):
        u = np.nextafter(pick, f32(0.0)) if below else pick.copy()
        u[:p.M * n] = seed_u
        counts, state, _ = g.simulate(_to_oracle_params(oracle, p, uniforms=u), want_state=True)
        with Simulation.from_contacts(n, a.astype(np.int32), b.astype(np.int32), s, e, p) as sim:
            sim.set_uniforms(u)
            sim.run()
            np.testing.assert_array_equal(sim.counts(), counts)
            np.testing.assert_array_equal(sim.state(), state)
        results.append(counts)
    assert not np.array_equal(results[0], results[1])


def test_entry_forms(oracle):
    """The step kernels stream the 8-byte copy of the entries where every
    entry fits it (step_bytes counts 8 bytes an entry), and the 16-byte
    entries where an epoch spans 2^20 s or more (epoch-relative times do not
    fit 20 bits) or when F_WIDE_ENTRIES asks for them; every form gives the
    oracle's trajectory."""
    n, a, b, s, e = _random_case(5, n=60, e=1500, span=40 * 86400, long_frac=0.5)
    def params(wpe, flags):
        return Params(M=256, step_size=3600.0, static_network_duration=3600, iterations=400,
                      initial_infected=0.05, infect_prob=0.2, exposed_duration=3600,
                      infectious_duration=3 * 86400, resistant_duration=86400, seed=3,
                      windows_per_epoch=wpe, flags=flags)
    g = oracle.OracleGraph.from_contacts(n, a, b, s, e)
    counts, state, _ = g.simulate(_to_oracle_params(oracle, params(8, 0)), want_state=True)
    assert counts[:, :, 1].max() > 0
    assert 400 * 3600 >= 1 << 20  # the one epoch's times need more than 20 bits
    rest = set()
    for wpe, flags, width in ((8, 0, 8), (8, replay_amd.F_WIDE_ENTRIES, 16), (400, 0, 16)):
        with Simulation.from_contacts(n, a, b, s, e, params(wpe, flags)) as sim:
            sim.set_variant(replay_amd.V_NODE4)
            info = sim.info()
            assert info["entry_bytes"] == width, (wpe, flags)
            sim.run()
            np.testing.assert_array_equal(sim.counts(), counts, err_msg=f"wpe {wpe} {flags}")
            np.testing.assert_array_equal(sim.state(), state, err_msg=f"wpe {wpe} {flags}")
            # step_bytes prices the entries at the streamed width; the rest
            # (state words, row pointers, counts) does not depend on it
            rest.add(info["step_bytes"] - int(width * (info["n_entries"] / info["n_epochs"])))
    assert len(rest) == 1


def test_slice_nnz_counts_the_window_csr():
    n, a, b, s, e = _random_case(8, n=60, e=1500)
    p = Params(M=64, step_size=1000.0, static_network_duration=3600, iterations=40)
    with Simulation.from_contacts(n, a, b, s, e, p) as sim:
        for k in range(0, p.iterations, 3):
            rp, col, _ = sim.slice(k)
            assert sim.slice_nnz(k) == len(col) == rp[-1]


def test_truncated_windows_trajectory(oracle):
    """Windows whose query exceeds the hit cap (every contact overlapping)."""
    n = 6
    a = np.array([0, 1, 2, 3, 4, 0], np.int32)
    b = np.array([1, 2, 3, 4, 5, 5], np.int32)
    s = np.full(6, 946713600 - 100, np.int64)
    e = s + np.array([10 ** 6, 10 ** 6, 10 ** 6 - 1, 10 ** 6, 5000, 10 ** 6])
    p = Params(M=64, step_size=3600.0, static_network_duration=3600, iterations=40,
               initial_infected=0.2, infect_prob=0.9, infectious_duration=36000,
               resistant_duration=36000)
    g = oracle.OracleGraph.from_contacts(n, a, b, s, e)
    counts, state, _ = g.simulate(_to_oracle_params(oracle, p), want_state=True)
    with Simulation.from_contacts(n, a, b, s, e, p) as sim:
        assert sim.info()["truncated_windows"] > 0
        sim.run()
        np.testing.assert_array_equal(sim.counts(), counts)
        np.testing.assert_array_equal(sim.state(), state)


def test_config0_full_size(oracle):
    w = workloads.config0_sir()
    g = oracle.OracleGraph.from_contacts(w.n_nodes, w.src, w.tgt, w.start, w.end)
    counts, state, _ = g.simulate(_to_oracle_params(oracle, w.params, threads=8), want_state=True)
    with Simulation.from_contacts(w.n_nodes, w.src, w.tgt, w.start, w.end, w.params) as sim:
        sim.run()
        np.testing.assert_array_equal(sim.counts(), counts)
        np.testing.assert_array_equal(sim.state(), state)
    assert counts[-1, :, 2].mean() > 10  # the epidemic took off


def test_config1_reduced(oracle):
    w = workloads.config1_village(n_nodes=2000, n_contacts=200_000, M=96, iterations=700, days=6)
    g = oracle.OracleGraph.from_contacts(w.n_nodes, w.src, w.tgt, w.start, w.end)
    counts, state, _ = g.simulate(_to_oracle_params(oracle, w.params, threads=8), want_state=True)
    with Simulation.from_contacts(w.n_nodes, w.src, w.tgt, w.start, w.end, w.params) as sim:
        sim.run()
        np.testing.assert_array_equal(sim.counts(), counts)
        np.testing.assert_array_equal(sim.state(), state)


def test_config1_full_size_invariants():
    """Full village configuration: size-independent properties."""
    w = workloads.config1_village()
    with Simulation.from_contacts(w.n_nodes, w.src, w.tgt, w.start, w.end, w.params) as sim:
        sim.run()
        c = sim.counts()
        st = sim.state()
    assert c.shape == (w.params.iterations + 1, w.params.M, 4)
    assert (c.sum(axis=2) == w.n_nodes).all()
    # permanent R within the horizon: recovered counts never decrease
    assert (np.diff(c[:, :, 2], axis=0) >= 0).all()
    # ever-infected (E + I + R) never decreases
    assert (np.diff(c[:, :, :3].sum(axis=2), axis=0) >= 0).all()
    assert ((st == 0) | (st > 0)).all()
    assert c[-1, :, 2].mean() > 1


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return r


@pytest.mark.parametrize("M,T", [(7, 40), (64, 400)])
def test_cli_csv_byte_identical(oracle, tmp_path, M, T):
    """Both CLI output paths against the oracle CLI, byte for byte: per
    iteration (node states and summaries written by a writer thread from
    double-buffered states) and all iterations at once (summaries only)."""
    path = graph_path("random_nested.txt")
    args = [path, "--M", str(M), "--step-size", "1200", "--iterations", str(T), "--time-step", "3600",
            "--initial-infected", "0.2", "--infect-prob", "0.4", "--exposed-duration", "3600",
            "--infectious-duration", "7200", "--resistant-duration", "3600",
            "--start-time", str(_graph_start(path)), "--seed", "11"]
    ours = [str(tmp_path / "s1.csv"), str(tmp_path / "n1.csv")]
    ref = [str(tmp_path / "s2.csv"), str(tmp_path / "n2.csv")]
    r = _run([replay_amd.CLI_PATH, *args, "--summary-out", ours[0], "--node-state-out", ours[1]])
    assert f"Starting temporal Monte Carlo simulation for {T} iterations..." in r.stdout
    _run([os.path.join(ROOT, "oracle", "replay_cpu"), *args, "--summary-out", ref[0],
          "--node-state-out", ref[1]])
    for x, y in zip(ours, ref):
        assert open(x).read() == open(y).read()
    # summary only (the all-iterations-at-once path)
    _run([replay_amd.CLI_PATH, *args, "-q", "--summary-out", ours[0]])
    assert open(ours[0]).read() == open(ref[0]).read()


def test_summary_frame_matches_csv_writer(tmp_path):
    """summary_frame over a simulation's counts and block times holds the
    rows write_summary_csv writes (the reference CSV layout)."""
    from replay_amd import summary, write_summary_csv
    n, a, b, s, e = _random_case(5, n=40, e=800)
    p = Params(M=5, step_size=900.0, static_network_duration=1800, iterations=25,
               initial_infected=0.1, infect_prob=0.3, exposed_duration=1800,
               infectious_duration=5400, resistant_duration=3600, start_time=946713600)
    with Simulation.from_contacts(n, a, b, s, e, p) as sim:
        sim.run()
        f = tmp_path / "s.csv"
        write_summary_csv(str(f), sim)
        times, counts = summary.block_times(sim), sim.counts()
    t2, c2 = summary.read_summary_csv(str(f))
    np.testing.assert_array_equal(t2, times)
    np.testing.assert_array_equal(c2, counts)
    df = summary.summary_frame(counts, times)
    assert df.shape == ((p.iterations + 1) * p.M, 6)
