"""High-density contact stress (the config2 recipe: hub nodes with Pareto
weights, long overlapping contacts, 5-minute steps, thousands of replicas).

Reduced size: bit-exact against the oracle.  Reduced size also: replica r's
trajectory does not depend on how many replicas run beside it (its uniforms
are a function of (node, r, slot, stream) only).  Full size (50k nodes, 40M
contacts, 4096 replicas, a 96-iteration prefix): every kernel variant gives
identical counts and states, and each iteration's class counts sum to the
node count.  Needs an MI355X."""
import numpy as np
import pytest

import replay_amd
from replay_amd import Params, Simulation, workloads

from test_gpu_parity import _to_oracle_params

pytestmark = pytest.mark.gpu


def _reduced(M=192, T=72):
    """The recipe at 1500 nodes (~170 simultaneous neighbours) with sojourns
    shortened so that every transition happens within the horizon."""
    w = workloads.config2_dense(n_nodes=1500, n_contacts=300_000, n_hubs=15, M=M, iterations=T,
                                days=1)
    w.params = Params(**{**w.params.__dict__, "initial_infected": 0.02, "infect_prob": 0.05,
                         "exposed_duration": 3600, "infectious_duration": 3 * 3600,
                         "resistant_duration": 2 * 3600})
    return w


def test_dense_reduced_matches_oracle(oracle):
    w = _reduced()
    p = w.params
    with Simulation.from_contacts(w.n_nodes, w.src, w.tgt, w.start, w.end, p) as sim:
        info = sim.info()
        assert info["n_entries"] > 0
        sim.run()
        counts, state = sim.counts(), sim.state()
    g = oracle.OracleGraph.from_contacts(w.n_nodes, w.src, w.tgt, w.start, w.end)
    oc, ostate, _ = g.simulate(_to_oracle_params(oracle, p, threads=8), want_state=True)
    np.testing.assert_array_equal(counts, oc)
    np.testing.assert_array_equal(state.reshape(-1), ostate)


def test_dense_replicas_independent_of_M():
    w = _reduced(M=256, T=60)
    out = {}
    for M in (64, 256):
        p = Params(**{**w.params.__dict__, "M": M})
        with Simulation.from_contacts(w.n_nodes, w.src, w.tgt, w.start, w.end, p) as sim:
            sim.run()
            out[M] = (sim.counts(), sim.state())
    np.testing.assert_array_equal(out[256][0][:, :64], out[64][0])
    # states are column-major (idx = node + sim * N): replicas 0..63 come first
    np.testing.assert_array_equal(out[256][1][: 64 * w.n_nodes], out[64][1])


def test_dense_full_size_invariants():
    w = workloads.config2_dense()
    p = w.params
    T = 96  # a prefix of the 4032 iterations keeps the test short
    p = Params(**{**p.__dict__, "iterations": T})
    with Simulation.from_contacts(w.n_nodes, w.src, w.tgt, w.start, w.end, p) as sim:
        info = sim.info()
        assert info["n_nodes"] == 50000 and info["M"] == 4096
        results = []
        for v in (replay_amd.V_NODE16, replay_amd.V_NODE16 | replay_amd.V_BLOCK1,
                  replay_amd.V_NODE16 | replay_amd.V_XCD, replay_amd.V_SLICED):
            sim.set_variant(v)
            sim.reset()
            sim.run()
            results.append((sim.counts(), sim.state()))
    counts = results[0][0]
    assert np.all(counts.sum(axis=2) == 50000)
    assert counts[-1, :, :3].sum() > counts[0, :, :3].sum()  # the epidemic grew
    for c, s in results[1:]:
        np.testing.assert_array_equal(c, results[0][0])
        np.testing.assert_array_equal(s, results[0][1])


def _two_hubs(seed=0, leaves=(3000, 5000), T=24):
    """Two star hubs with more entries than the long-row threshold: with
    one-hour windows the 3000-leaf hub's weights stay below 2^24 seconds
    (integer sums), the 5000-leaf hub's exceed it (float order); leaves also
    meet each other at random."""
    rng = np.random.default_rng(seed)
    n = 2 + sum(leaves)
    src, tgt, st, en = [], [], [], []
    t0 = 946713600
    first = 2
    for h, L in enumerate(leaves):
        leaf = np.arange(first, first + L)
        first += L
        s = t0 + rng.integers(-3600, 3 * 3600, L)  # nearly all leaves in touch all along,
        d = rng.integers((T - 4) * 3600, (T + 2) * 3600, L)  # some ending inside a window
        src.append(np.full(L, h)); tgt.append(leaf); st.append(s); en.append(s + d)
    k = 20000
    a = rng.integers(2, n, k)
    b = rng.integers(2, n, k)
    keep = a != b
    s = t0 + rng.integers(0, T * 3600, k)
    src.append(a[keep]); tgt.append(b[keep]); st.append(s[keep])
    en.append(s[keep] + rng.integers(60, 4 * 3600, k)[keep])
    cat = lambda xs, dt: np.concatenate(xs).astype(dt)
    st = np.maximum(cat(st, np.int64), t0 - 7200)
    return n, cat(src, np.int32), cat(tgt, np.int32), st, cat(en, np.int64)


@pytest.mark.parametrize("variant", [None, replay_amd.V_SLICED, replay_amd.V_WORD])
def test_long_rows_match_oracle(oracle, variant):
    n, a, b, s, e = _two_hubs()
    p = Params(M=128, step_size=3600.0, static_network_duration=3600, iterations=24,
               initial_infected=0.05, infect_prob=0.2, exposed_duration=3600,
               infectious_duration=6 * 3600, resistant_duration=10 ** 6, seed=3)
    with Simulation.from_contacts(n, a, b, s, e, p) as sim:
        if variant is not None:
            sim.set_variant(variant)
            sim.reset()
        sim.run()
        counts, state = sim.counts(), sim.state()
    g = oracle.OracleGraph.from_contacts(n, a, b, s, e)
    oc, ostate, _ = g.simulate(_to_oracle_params(oracle, p, threads=8), want_state=True)
    np.testing.assert_array_equal(counts, oc)
    np.testing.assert_array_equal(state.reshape(-1), ostate)


def test_long_rows_fractional_weights_match_oracle(oracle):
    """Three 20-minute iterations per one-hour window: weights overlap / 3
    are not integers, and the hub rows are walked in the reference's fp32
    order by one wavefront per (row, word)."""
    n, a, b, s, e = _two_hubs(seed=7, T=12)
    p = Params(M=96, step_size=1200.0, static_network_duration=3600, iterations=36,
               initial_infected=0.05, infect_prob=0.2, exposed_duration=1200,
               infectious_duration=6 * 3600, resistant_duration=10 ** 6, seed=4)
    with Simulation.from_contacts(n, a, b, s, e, p) as sim:
        sim.run()
        counts, state = sim.counts(), sim.state()
    g = oracle.OracleGraph.from_contacts(n, a, b, s, e)
    oc, ostate, _ = g.simulate(_to_oracle_params(oracle, p, threads=8), want_state=True)
    np.testing.assert_array_equal(counts, oc)
    np.testing.assert_array_equal(state.reshape(-1), ostate)
    assert oc[-1, :, 1:3].sum() > oc[0, :, 1:3].sum()


def test_long_row_sums_beyond_2_24_match_oracle(oracle):
    """The 5000-leaf hub's row is listed beyond the 2^24 bound and is summed
    in integers first.  Here nearly every leaf is seeded infected, so from
    the second iteration a susceptible hub replica's sum (about 4850 leaves x
    3600 s) passes 2^24 and the row is summed again in the reference order;
    the small infection probability keeps the hub's draws near P = 0.4, far
    from saturating."""
    n, a, b, s, e = _two_hubs(seed=4)
    p = Params(M=256, step_size=3600.0, static_network_duration=3600, iterations=10,
               initial_infected=0.97, infect_prob=1e-4, exposed_duration=0,
               infectious_duration=20 * 3600, resistant_duration=10 ** 6, seed=9)
    with Simulation.from_contacts(n, a, b, s, e, p) as sim:
        sim.run()
        counts, state = sim.counts(), sim.state()
    g = oracle.OracleGraph.from_contacts(n, a, b, s, e)
    oc, ostate, _ = g.simulate(_to_oracle_params(oracle, p, threads=8), want_state=True)
    np.testing.assert_array_equal(counts, oc)
    np.testing.assert_array_equal(state.reshape(-1), ostate)


def _hub_with_runs(seed=0, leaves=6000, T=10):
    """One star hub whose leaves each meet it up to three times: one contact
    spanning the horizon and up to two short ones (partial overlaps), so the
    hub's row holds runs of entries of one column, spread over several
    2048-entry segments; leaves also meet each other at random."""
    rng = np.random.default_rng(seed)
    n = 1 + leaves
    t0 = 946713600
    leaf = np.arange(1, n)
    extra = rng.integers(0, 3, leaves)
    short = np.repeat(leaf, extra)
    ss = t0 + rng.integers(0, T * 3600, short.size)
    k = 8000
    a = rng.integers(1, n, k)
    b = rng.integers(1, n, k)
    keep = a != b
    ls = t0 + rng.integers(0, T * 3600, k)
    src = np.concatenate([np.zeros(leaves + short.size, np.int64), a[keep]]).astype(np.int32)
    tgt = np.concatenate([leaf, short, b[keep]]).astype(np.int32)
    st = np.concatenate([np.full(leaves, t0 - 3600), ss, ls[keep]]).astype(np.int64)
    en = np.concatenate([np.full(leaves, t0 + (T + 2) * 3600),
                         ss + rng.integers(60, 2 * 3600, short.size),
                         ls[keep] + rng.integers(60, 4 * 3600, k)[keep]]).astype(np.int64)
    return n, src, tgt, st, en


@pytest.mark.parametrize("M", [128, 200])
@pytest.mark.parametrize("beta", [1e-4, 0.2])
def test_hub_runs_beyond_2_24_match_oracle(oracle, M, beta):
    """A hub row of 6000 leaves, most of them seeded infectious: a
    susceptible hub replica's sum passes 2^24 s about 80% of the way along
    the row, in its third segment, and in two of three cases inside a run of
    entries of one column.  beta = 1e-4 keeps P near 0.4, so the exposure is
    summed again in the reference's fp32 order from that column on; with
    beta = 0.2 every exposure of 2^24 s or more gives P = 1 and the draw is
    settled without it."""
    n, a, b, s, e = _hub_with_runs(seed=int(M * 10 + beta * 1e4))
    p = Params(M=M, step_size=3600.0, static_network_duration=3600, iterations=10,
               initial_infected=0.97, infect_prob=beta, exposed_duration=0,
               infectious_duration=20 * 3600, resistant_duration=10 ** 6, seed=11)
    with Simulation.from_contacts(n, a, b, s, e, p) as sim:
        # the premise: susceptible hub replicas whose exposure passes 2^24 once
        # the seeds are infectious (iteration 1)
        sim.run(1)
        sus = sim.state().reshape(M, n)[:, 0] == 0
        E1 = sim.exposure(1).reshape(M, n)[:, 0]
        assert sus.any() and (E1[sus] >= 2 ** 24).any()
        sim.run()
        counts, state = sim.counts(), sim.state()
    g = oracle.OracleGraph.from_contacts(n, a, b, s, e)
    oc, ostate, _ = g.simulate(_to_oracle_params(oracle, p, threads=8), want_state=True)
    np.testing.assert_array_equal(counts, oc)
    np.testing.assert_array_equal(state.reshape(-1), ostate)


@pytest.mark.parametrize("M", [130, 256])
@pytest.mark.parametrize("seeded", [0.05, 0.9])
def test_daily_windows_beyond_2_24_match_oracle(oracle, M, seeded):
    """Daily windows (step = window = 86400 s) over nodes with about 600
    multi-day contacts each: every short row's weight bound passes 2^24, so
    the windows are not integer windows although their weights are
    integers, and each group is summed in integers first.  With 5% seeded
    the replicas' sums stay below 2^24 (the integer sums decide); with 90%
    seeded a susceptible replica's sum passes it (the reference order
    decides).  A small infection probability keeps P near 0.4."""
    import ctypes as C
    from test_host_slices import HostPlan
    rng = np.random.default_rng(13)
    n, k = 400, 120000
    a = rng.integers(0, n, k)
    b = (a + 1 + rng.integers(0, n - 1, k)) % n
    t0 = 946713600
    s = t0 + rng.integers(-2 * 86400, 3 * 86400, k)
    e = s + rng.integers(4 * 86400, 14 * 86400, k)
    p = Params(M=M, step_size=86400.0, static_network_duration=86400, iterations=10,
               initial_infected=seeded, infect_prob=1e-4, exposed_duration=0,
               infectious_duration=20 * 86400, resistant_duration=10 ** 7, seed=7,
               start_time=t0)
    # every window has integer weights but is not an integer window
    import hostlib
    L = hostlib.load()
    L.rp_host_plan_open.argtypes = [C.c_int32, C.c_int64] + [C.POINTER(C.c_int32)] * 2 + \
        [C.POINTER(C.c_int64)] * 2 + [C.POINTER(replay_amd._Params), C.POINTER(C.c_void_p)]
    L.rp_host_plan_int_mode.argtypes = [C.c_void_p, C.c_int32]
    L.rp_host_plan_int_mode.restype = C.c_int32
    L.rp_host_plan_close.argtypes = [C.c_void_p]
    hp = HostPlan(L, n, a, b, s, e, p)
    assert [L.rp_host_plan_int_mode(hp.h, it) for it in range(p.iterations)] == [0] * 10
    hp.close()
    with Simulation.from_contacts(n, a.astype(np.int32), b.astype(np.int32), s, e, p) as sim:
        sim.run()
        counts, state = sim.counts(), sim.state()
    g = oracle.OracleGraph.from_contacts(n, a.astype(np.int32), b.astype(np.int32), s, e)
    oc, ostate, _ = g.simulate(_to_oracle_params(oracle, p, threads=8), want_state=True)
    np.testing.assert_array_equal(counts, oc)
    np.testing.assert_array_equal(state.reshape(-1), ostate)
    assert (counts[-1, :, 1] + counts[-1, :, 2]).max() > seeded * n  # infections were drawn


def test_long_row_exposure_matches_oracle_order(oracle):
    """Exposure of hub rows (integer-capable and beyond 2^24 seconds) equals
    the sequential fp32 sum of the reference CSR bit for bit."""
    import ctypes as C
    n, a, b, s, e = _two_hubs(seed=2)
    p = Params(M=96, step_size=3600.0, static_network_duration=3600, iterations=12,
               initial_infected=0.3, infect_prob=0.2, exposed_duration=0,
               infectious_duration=4 * 3600, resistant_duration=2 * 3600, seed=5)
    g = oracle.OracleGraph.from_contacts(n, a, b, s, e)
    op = _to_oracle_params(oracle, p)
    P = lambda arr, t: arr.ctypes.data_as(C.POINTER(t))
    with Simulation.from_contacts(n, a, b, s, e, p) as sim:
        for k in (0, 4, 8):
            state = sim.state()
            got = sim.exposure(k)
            lower, medium = p.resistant_duration, p.infectious_duration + p.resistant_duration
            inf = ((state > lower) & (state <= medium)).astype(np.float32)
            rp, col, val, _ = g.step_csr(op, k)
            ref = np.zeros(n * p.M, np.float32)
            oracle.lib().or_spmm(P(rp, C.c_int32), P(col, C.c_int32), P(val, C.c_float), len(col),
                                 n, P(inf, C.c_float), P(ref, C.c_float), p.M, 1)
            np.testing.assert_array_equal(got.view(np.int32), ref.view(np.int32))
            sim.run(4)


def test_batch_with_hub_network_equals_plain():
    from replay_amd import Network
    from test_gpu_parity import _random_case
    nets = []
    n, a, b, s, e = _two_hubs(seed=4, leaves=(2600, 4800), T=16)
    nets.append(Network(n_nodes=n, src=a, tgt=b, start=s, end=e, initial_infected=0.03,
                        infect_prob=0.15, exposed_duration=3600, infectious_duration=5 * 3600,
                        resistant_duration=10 ** 6, stream=1))
    n2, a2, b2, s2, e2 = _random_case(11, n=300, e=4000, span=16 * 3600, long_frac=0.3)
    nets.append(Network(n_nodes=n2, src=a2, tgt=b2, start=s2, end=e2, initial_infected=0.1,
                        infect_prob=0.3, exposed_duration=0, infectious_duration=3 * 3600,
                        resistant_duration=0, stream=2))
    shared = Params(M=160, step_size=3600.0, static_network_duration=3600, iterations=16, seed=9)
    with Simulation.from_networks(nets, shared) as sim:
        sim.run()
        bc = sim.counts()
    for k, nw in enumerate(nets):
        pk = Params(M=160, step_size=3600.0, static_network_duration=3600, iterations=16, seed=9,
                    initial_infected=nw.initial_infected, infect_prob=nw.infect_prob,
                    exposed_duration=nw.exposed_duration,
                    infectious_duration=nw.infectious_duration,
                    resistant_duration=nw.resistant_duration, stream=nw.stream)
        with Simulation.from_contacts(nw.n_nodes, nw.src, nw.tgt, nw.start, nw.end, pk) as s1:
            s1.run()
            np.testing.assert_array_equal(s1.counts(), bc[:, k])
