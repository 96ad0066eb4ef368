"""The 64-network batched workload (the headline benchmark) on the GPU.

Reduced recipe: 8 networks of 500-4000 nodes drawn by the same generator
(uniform random pairs, 20-200 simultaneous neighbours, per-network beta,
durations and stream id), 256 replicas, 48 one-hour iterations, against the
oracle network by network under every exposure variant.

Full size (64 networks, about 807k nodes, 66M contacts, 256 replicas, 168
iterations): properties that hold at any size -- the per-replica SEIR counts
from the device's prefix sums equal a classification of the countdown state
(cu:505-568) done here in numpy; exposed + infectious + resistant never falls
(the resistant duration outlasts the horizon); every variant gives the same
state; the batch equals the per-network contexts for three networks."""
import numpy as np
import pytest

import replay_amd
from replay_amd import Params, Simulation, workloads

from test_gpu_batch import _plain_params
from test_gpu_parity import _to_oracle_params

pytestmark = pytest.mark.gpu

VARIANTS = replay_amd.VARIANTS_ALL


def test_config3_reduced_recipe_matches_oracle(oracle):
    w = workloads.config3_batched(n_networks=8, n_min=500, n_max=4000, iterations=48)
    expected = []
    for nw in w.networks:
        g = oracle.OracleGraph.from_contacts(nw.n_nodes, nw.src, nw.tgt, nw.start, nw.end)
        op = _to_oracle_params(oracle, _plain_params(w.params, nw), threads=16)
        op.stream = nw.stream
        expected.append(g.simulate(op)[0])
    print(f"oracle done: sizes {[nw.n_nodes for nw in w.networks]}", flush=True)
    with Simulation.from_networks(w.networks, w.params) as sim:
        for v in VARIANTS:
            sim.set_variant(v)
            sim.reset()
            sim.run()
            bc = sim.counts()
            for k, counts in enumerate(expected):
                np.testing.assert_array_equal(bc[:, k], counts, err_msg=f"variant {v} network {k}")
    assert sum(int(c[-1, :, 1:3].sum()) for c in expected) > 0


def _classify(state_k: np.ndarray, Ed: int, Id: int, Rd: int) -> np.ndarray:
    """Per-replica (E, I, R, S) counts of one network's countdowns, shape
    (M, n), by cu:530-550."""
    up = Ed + Id + Rd
    c = state_k.astype(np.int64)
    el = np.where((c <= 0) | (c > up), 0, up - c)
    E = (c != 0) & (Ed > 0) & (el < Ed)
    I = (c != 0) & ~E & (Id > 0) & (el < Ed + Id)
    R = (c != 0) & ~E & ~I & (Rd > 0) & (el < Ed + Id + Rd)
    S = ~(E | I | R)
    return np.stack([E.sum(1), I.sum(1), R.sum(1), S.sum(1)], axis=1).astype(np.int32)


def test_config3_full_size_properties():
    w = workloads.config3_batched()
    T, M = w.params.iterations, w.params.M
    gp = Params(**{**w.params.__dict__, "flags": replay_amd.F_GRAPH})
    with Simulation.from_networks(w.networks, gp) as sim:
        off = sim.node_offsets
        N = int(off[-1])
        print(f"opened: N={N} networks={len(w.networks)}", flush=True)
        ref_state = None
        for v in VARIANTS:
            sim.set_variant(v)
            sim.reset()
            sim.run()
            counts = sim.counts()
            state = sim.state().reshape(M, N)
            print(f"variant {v} ran", flush=True)
            if ref_state is None:
                ref_state, ref_counts = state, counts
                continue
            np.testing.assert_array_equal(counts, ref_counts, err_msg=f"variant {v}")
            assert np.array_equal(state, ref_state), f"variant {v} state differs"
            del state
    counts = ref_counts
    assert counts.shape == (T + 1, len(w.networks), M, 4)
    assert (counts >= 0).all()
    for k, nw in enumerate(w.networks):
        ck = counts[:, k]
        assert (ck.sum(axis=2) == nw.n_nodes).all()
        # resistant outlasts the horizon: nobody returns to S
        ever = ck[:, :, :3].sum(axis=2)
        assert (np.diff(ever, axis=0) >= 0).all(), f"network {k}: E+I+R fell"
        # the last block of counts equals the state classified here
        got = _classify(ref_state[:, off[k]:off[k + 1]], nw.exposed_duration,
                        nw.infectious_duration, nw.resistant_duration)
        np.testing.assert_array_equal(ck[-1], got, err_msg=f"network {k}")
    assert counts[-1, :, :, 1:3].sum() > 0
    # batch equals the per-network contexts
    for k in (0, len(w.networks) // 2, len(w.networks) - 1):
        nw = w.networks[k]
        with Simulation.from_contacts(nw.n_nodes, nw.src, nw.tgt, nw.start, nw.end,
                                      _plain_params(w.params, nw)) as s:
            s.run()
            np.testing.assert_array_equal(s.counts(), counts[:, k], err_msg=f"network {k}")
            st = s.state().reshape(M, nw.n_nodes)
            assert np.array_equal(st, ref_state[:, off[k]:off[k + 1]]), f"network {k} state"
