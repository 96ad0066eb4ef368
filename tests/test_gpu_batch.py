"""Batched form (several networks in one context, one step launch for all)
against the plain form (one context per network) and the oracle, plus
hipGraph replays and kernel-variant agreement.  Needs an MI355X."""
import numpy as np
import pytest

import replay_amd
from replay_amd import Network, Params, Simulation, workloads

from test_gpu_parity import _random_case, _to_oracle_params

pytestmark = pytest.mark.gpu


def _nets(seed, sizes, long_frac=0.3):
    rng = np.random.default_rng(seed)
    out = []
    for k, n in enumerate(sizes):
        _, a, b, s, e = _random_case(seed * 31 + k, n=n, e=n * 12, long_frac=long_frac)
        out.append(Network(n_nodes=n, src=a, tgt=b, start=s, end=e,
                           initial_infected=float(rng.uniform(0.02, 0.2)),
                           infect_prob=float(rng.uniform(0.1, 0.6)),
                           exposed_duration=int(rng.integers(0, 4)) * 1800,
                           infectious_duration=int(rng.integers(2, 8)) * 1800,
                           resistant_duration=int(rng.choice([0, 3600, 10 ** 6])),
                           stream=100 + k))
    return out


def _plain_params(shared: Params, nw: Network) -> Params:
    return Params(M=shared.M, step_size=shared.step_size, iterations=shared.iterations,
                  initial_infected=nw.initial_infected, infect_prob=nw.infect_prob,
                  exposed_duration=nw.exposed_duration,
                  infectious_duration=nw.infectious_duration,
                  resistant_duration=nw.resistant_duration, start_time=shared.start_time,
                  static_network_duration=shared.static_network_duration, seed=shared.seed,
                  stream=nw.stream, windows_per_epoch=shared.windows_per_epoch)


@pytest.mark.parametrize("M", [64, 200, 1024])
@pytest.mark.parametrize("step,D", [(600.0, 600), (1000.0, 3600)])
def test_batch_equals_plain_and_oracle(oracle, M, step, D):
    nets = _nets(M + int(step), [37, 80, 5, 150])
    shared = Params(M=M, step_size=step, static_network_duration=D, iterations=90, seed=7)
    with Simulation.from_networks(nets, shared) as b:
        assert b.n_networks == 4
        b.run()
        bc = b.counts()
        bs = b.state()
        off = b.node_offsets
    N = int(off[-1])
    for k, nw in enumerate(nets):
        pp = _plain_params(shared, nw)
        with Simulation.from_contacts(nw.n_nodes, nw.src, nw.tgt, nw.start, nw.end, pp) as s:
            s.run()
            np.testing.assert_array_equal(s.counts(), bc[:, k])
            st = s.state().reshape(M, nw.n_nodes)
            np.testing.assert_array_equal(st, bs.reshape(M, N)[:, off[k]:off[k + 1]])
        g = oracle.OracleGraph.from_contacts(nw.n_nodes, nw.src, nw.tgt, nw.start, nw.end)
        op = _to_oracle_params(oracle, pp)
        op.stream = nw.stream
        counts, _, _ = g.simulate(op)
        np.testing.assert_array_equal(counts, bc[:, k])


def test_graph_replays_match_eager():
    nets = _nets(3, [60, 90, 33])
    shared = Params(M=128, step_size=900.0, static_network_duration=1800, iterations=80, seed=1)
    with Simulation.from_networks(nets, shared) as eager:
        eager.run()
        ref = eager.counts()
        ref_state = eager.state()
    gp = Params(**{**shared.__dict__, "flags": replay_amd.F_GRAPH})
    with Simulation.from_networks(nets, gp) as g:
        for rep in range(3):
            g.reset()
            g.run()
            assert g.info()["graph_ready"] == 1
            np.testing.assert_array_equal(g.counts(), ref)
            np.testing.assert_array_equal(g.state(), ref_state)
            # eager work between replays
            g.reset()
            g.run(10)
            g.sync()


@pytest.mark.parametrize("M", [256, 1000])
def test_variants_agree(M):
    nets = _nets(5, [120, 70])
    shared = Params(M=M, step_size=600.0, static_network_duration=600, iterations=60, seed=2)
    results = []
    with Simulation.from_networks(nets, shared) as sim:
        for v in replay_amd.VARIANTS_ALL:
            sim.set_variant(v)
            sim.reset()
            sim.run()
            results.append((sim.counts(), sim.state()))
        best = sim.autotune(20)
        assert best in replay_amd.VARIANTS_TUNED
        sim.run()
        results.append((sim.counts(), sim.state()))
    for c, s in results[1:]:
        np.testing.assert_array_equal(c, results[0][0])
        np.testing.assert_array_equal(s, results[0][1])


def test_autotune_times_every_variant_under_graph_flag():
    """Open-time autotune with REPLAY_F_GRAPH and T <= 32 (the autotune
    covers the whole run): every applicable variant is timed by eager
    launches, no graph of a candidate survives, and the first full run after
    it replays a graph of the chosen variant with the eager results."""
    nets = _nets(11, [60, 90, 33])
    shared = Params(M=256, step_size=900.0, static_network_duration=1800, iterations=24, seed=3)
    with Simulation.from_networks(nets, shared) as eager:
        eager.run()
        ref = eager.counts()
    gp = Params(**{**shared.__dict__, "flags": replay_amd.F_GRAPH | replay_amd.F_AUTOTUNE})
    with Simulation.from_networks(nets, gp) as g:
        info = g.info()
        assert info["graph_ready"] == 0
        assert all(t > 0 for t in info["tune_ms"]), info["tune_ms"]  # W=4: all four apply
        g.run()
        assert g.info()["graph_ready"] == 1
        np.testing.assert_array_equal(g.counts(), ref)
        first = g.info()["variant"]
        g.autotune(24)  # drops the graph; the next full run captures again
        info = g.info()
        assert info["graph_ready"] == 0 and all(t > 0 for t in info["tune_ms"])
        g.reset()
        g.run()
        assert g.info()["graph_ready"] == 1
        np.testing.assert_array_equal(g.counts(), ref)
        assert first in replay_amd.VARIANTS_TUNED


def test_config3_reduced_batch_vs_plain():
    w = workloads.config3_batched(n_networks=6, n_min=300, n_max=3000, iterations=48,
                                  mean_days=2.0)
    with Simulation.from_networks(w.networks, w.params) as b:
        b.run()
        bc = b.counts()
    for k, nw in enumerate(w.networks):
        with Simulation.from_contacts(nw.n_nodes, nw.src, nw.tgt, nw.start, nw.end,
                                      _plain_params(w.params, nw)) as s:
            s.run()
            np.testing.assert_array_equal(s.counts(), bc[:, k])
    assert bc[-1, :, :, 1:3].sum() > 0


@pytest.mark.parametrize("case", range(16))
def test_random_batches_every_variant(oracle, case):
    """Random batches of 2-6 uneven networks (5-150 nodes, their own
    infection probability, durations and seeding), 33-1024 replicas, integer
    and float windows: under every exposure variant, each network's counts
    equal the oracle's for that network alone with its stream id, and the
    batch's states are the same under every variant."""
    rng = np.random.default_rng(7000 + case)
    sizes = [int(x) for x in rng.integers(5, 151, int(rng.integers(2, 7)))]
    nets = _nets(int(rng.integers(0, 1 << 20)), sizes, long_frac=float(rng.choice([0.0, 0.3, 0.6])))
    step, D = [(600.0, 600), (600.0, 1800), (1000.0, 3600), (3600.0, 3600)][int(rng.integers(0, 4))]
    M = int(rng.choice([33, 64, 256, 1024]))
    shared = Params(M=M, step_size=step, static_network_duration=D, iterations=50,
                    seed=int(rng.integers(0, 1 << 30)))
    print(f"case {case}: sizes={sizes} M={M} step={step} D={D}", flush=True)
    expected = []
    for nw in nets:
        g = oracle.OracleGraph.from_contacts(nw.n_nodes, nw.src, nw.tgt, nw.start, nw.end)
        op = _to_oracle_params(oracle, _plain_params(shared, nw))
        op.stream = nw.stream
        expected.append(g.simulate(op)[0])
    states = []
    with Simulation.from_networks(nets, shared) as sim:
        for v in replay_amd.VARIANTS_ALL:
            sim.set_variant(v)
            sim.reset()
            sim.run()
            bc = sim.counts()
            for k, counts in enumerate(expected):
                np.testing.assert_array_equal(bc[:, k], counts, err_msg=f"variant {v} network {k}")
            states.append(sim.state())
    for s in states[1:]:
        np.testing.assert_array_equal(s, states[0])


def _uneven_params(shared: Params, nw: Network) -> Params:
    return Params(**{**_plain_params(shared, nw).__dict__, "M": nw.M or shared.M})


@pytest.mark.parametrize("case", range(10))
def test_uneven_replica_counts_every_variant(oracle, case):
    """Batches whose networks have their own replica counts (33, 256 and
    1000 replicas, i.e. 1, 4 and 16 words per node, in one launch), integer
    and float windows: under every variant that applies, each network's
    counts and countdown states equal the oracle's for that network alone
    with its replica count and stream id.  The bit-sliced variant needs one
    replica count and is refused.  In nine of the ten cases a network's word
    shift (first word less first node times words per node) is negative, so
    a masked load of global row 0 would fall before the infectious words:
    the exposure reduction's masked lanes read a row of their own network."""
    rng = np.random.default_rng(9100 + case)
    sizes = [int(x) for x in rng.integers(5, 151, int(rng.integers(2, 6)))]
    nets = _nets(int(rng.integers(0, 1 << 20)), sizes, long_frac=float(rng.choice([0.0, 0.3, 0.6])))
    Ms = [33, 256, 1000]
    rng.shuffle(Ms)
    for k, nw in enumerate(nets):
        nw.M = int(Ms[k % 3]) if k < 3 else int(rng.choice([1, 64, 65, 130, 1000]))
    step, D = [(600.0, 600), (600.0, 1800), (1000.0, 3600), (3600.0, 3600)][int(rng.integers(0, 4))]
    flags = replay_amd.F_GRAPH if case % 2 else 0
    shared = Params(M=64, step_size=step, static_network_duration=D, iterations=40,
                    seed=int(rng.integers(0, 1 << 30)), flags=flags)
    print(f"case {case}: sizes={sizes} M={[nw.M for nw in nets]} step={step} D={D}", flush=True)
    expected = []
    for nw in nets:
        g = oracle.OracleGraph.from_contacts(nw.n_nodes, nw.src, nw.tgt, nw.start, nw.end)
        op = _to_oracle_params(oracle, _uneven_params(shared, nw))
        op.stream = nw.stream
        expected.append(g.simulate(op, want_state=True)[:2])
    with Simulation.from_networks(nets, shared) as sim:
        assert not sim.uniform
        assert sim.info()["replicas_total"] == sum(nw.M for nw in nets)
        assert list(np.diff(sim.replica_offsets)) == [nw.M for nw in nets]
        with pytest.raises(replay_amd.ReplayError):
            sim.set_variant(replay_amd.V_SLICED)
        for v in [v for v in replay_amd.VARIANTS_ALL if v != replay_amd.V_SLICED]:
            sim.set_variant(v)
            for rep in range(2 if flags else 1):  # a graph is captured, then replayed
                sim.reset()
                sim.run()
                for k, (counts, state) in enumerate(expected):
                    np.testing.assert_array_equal(sim.network_counts(k), counts,
                                                  err_msg=f"variant {v} network {k}")
                    np.testing.assert_array_equal(sim.network_state(k), state,
                                                  err_msg=f"variant {v} network {k}")


def test_uneven_batch_equals_plain_with_hub_rows():
    """A hub network (rows beyond the long-row threshold, 130 replicas) with
    a random network of 1000 replicas in one batch: each equals its own
    single-network context, counts and states; the batch-wide calls whose
    layout needs one replica count refuse."""
    from test_gpu_dense import _two_hubs
    n, a, b, s, e = _two_hubs(seed=4, leaves=(2600, 4800), T=16)
    nets = [Network(n_nodes=n, src=a, tgt=b, start=s, end=e, initial_infected=0.03,
                    infect_prob=0.15, exposed_duration=3600, infectious_duration=5 * 3600,
                    resistant_duration=10 ** 6, stream=1, M=130)]
    n2, a2, b2, s2, e2 = _random_case(11, n=300, e=4000, span=16 * 3600, long_frac=0.3)
    nets.append(Network(n_nodes=n2, src=a2, tgt=b2, start=s2, end=e2, initial_infected=0.1,
                        infect_prob=0.3, exposed_duration=0, infectious_duration=3 * 3600,
                        resistant_duration=0, stream=2, M=1000))
    shared = Params(M=64, step_size=3600.0, static_network_duration=3600, iterations=16, seed=9)
    with Simulation.from_networks(nets, shared) as sim:
        sim.run()
        got = [(sim.network_counts(k), sim.network_state(k)) for k in range(2)]
        with pytest.raises(ValueError):
            sim.counts()
        for call in (sim.state, lambda: sim.exposure(0),
                     lambda: sim.set_uniforms(np.zeros(17 * 1130 * sim.n_nodes, np.float32))):
            with pytest.raises(replay_amd.ReplayError) as ei:
                call()
            assert ei.value.code == replay_amd.E_UNSUPPORTED
    for k, nw in enumerate(nets):
        with Simulation.from_contacts(nw.n_nodes, nw.src, nw.tgt, nw.start, nw.end,
                                      _uneven_params(shared, nw)) as s1:
            s1.run()
            np.testing.assert_array_equal(s1.counts(), got[k][0])
            np.testing.assert_array_equal(s1.state(), got[k][1])
    assert got[0][0][-1, :, 1:3].sum() > 0
