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
        for v in (replay_amd.V_WORD, replay_amd.V_NODE4, replay_amd.V_NODE16, replay_amd.V_SLICED):
            sim.set_variant(v)
            sim.reset()
            sim.run()
            results.append((sim.counts(), sim.state()))
        best = sim.autotune(20)
        assert best in (replay_amd.V_WORD, replay_amd.V_NODE4, replay_amd.V_NODE16,
                        replay_amd.V_SLICED)
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
        assert first in (replay_amd.V_WORD, replay_amd.V_NODE4, replay_amd.V_NODE16,
                         replay_amd.V_SLICED)


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
        for v in (replay_amd.V_WORD, replay_amd.V_NODE4, replay_amd.V_NODE16, replay_amd.V_SLICED):
            sim.set_variant(v)
            sim.reset()
            sim.run()
            bc = sim.counts()
            for k, counts in enumerate(expected):
                np.testing.assert_array_equal(bc[:, k], counts, err_msg=f"variant {v} network {k}")
            states.append(sim.state())
    for s in states[1:]:
        np.testing.assert_array_equal(s, states[0])
