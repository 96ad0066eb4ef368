"""Contact graphs at the edges of the input domain, HIP engine vs the oracle
(needs an MI355X).

The reference keeps a contact whose two ends are the same node: the pair
packs to (u, u) (temporal_contact_matrix.c:80-90) and the window CSR gets
the entry twice, merged (gpu_cpu_temporal_sim.cu:238-350).  Such an entry
adds w * I[u] to E[u], which is zero whenever u is susceptible, so it never
changes a draw, but it is a row entry the exposure kernels walk.  Nodes of
the node list without a contact have empty rows in every window; a window
with no contact at all gives every row an empty slice.  Every case is
compared with the oracle bit for bit, counts and countdown states, under
every exposure variant."""
import numpy as np
import pytest

import replay_amd
from replay_amd import Params, Simulation
from test_gpu_parity import _to_oracle_params

pytestmark = pytest.mark.gpu

T0 = 946713600
VARIANTS = [None, *replay_amd.VARIANTS_ALL]


def _self_loops_one_node():
    s = np.array([T0, T0 + 1800, T0 + 7000], np.int64)
    return 1, np.zeros(3, np.int32), np.zeros(3, np.int32), s, s + 3600


def _mostly_isolated(seed=7):
    # 60 nodes, contacts among the first six only, a third of them self-loops
    rng = np.random.default_rng(seed)
    e = 90
    a = rng.integers(0, 6, e)
    b = np.where(rng.uniform(size=e) < 0.33, a, rng.integers(0, 6, e))
    s = T0 + rng.integers(0, 6 * 3600, e)
    d = rng.integers(60, 4 * 3600, e)
    return 60, a.astype(np.int32), b.astype(np.int32), s.astype(np.int64), (s + d).astype(np.int64)


def _outside_horizon(seed=11):
    # contacts ending before the start or beginning after the last window,
    # plus a few inside, so that most windows hold no contact
    rng = np.random.default_rng(seed)
    n, e = 25, 120
    a = rng.integers(0, n, e)
    b = (a + 1 + rng.integers(0, n - 1, e)) % n
    where = rng.integers(0, 3, e)
    s = np.where(where == 0, T0 - 86400 + rng.integers(0, 3600, e),
                 np.where(where == 1, T0 + 30 * 86400 + rng.integers(0, 3600, e),
                          T0 + rng.integers(0, 8 * 3600, e)))
    d = np.where(where == 2, rng.integers(60, 1800, e), rng.integers(60, 3600, e))
    return n, a.astype(np.int32), b.astype(np.int32), s.astype(np.int64), (s + d).astype(np.int64)


CASES = {"self_loops_one_node": _self_loops_one_node, "mostly_isolated": _mostly_isolated,
         "outside_horizon": _outside_horizon}


@pytest.mark.parametrize("M", [1, 65, 256])
@pytest.mark.parametrize("case", sorted(CASES))
def test_edge_graph_every_variant(oracle, case, M):
    n, a, b, s, e = CASES[case]()
    p = Params(M=M, step_size=600.0, static_network_duration=1200, iterations=60,
               initial_infected=0.3, infect_prob=0.6, exposed_duration=1200,
               infectious_duration=6000, resistant_duration=1800, start_time=T0, seed=3)
    g = oracle.OracleGraph.from_contacts(n, a, b, s, e)
    counts, state, _ = g.simulate(_to_oracle_params(oracle, p), want_state=True)
    with Simulation.from_contacts(n, a, b, s, e, p) as sim:
        for v in VARIANTS:
            if v is not None:
                sim.set_variant(v)
            sim.reset()
            sim.run()
            np.testing.assert_array_equal(sim.counts(), counts, err_msg=f"{case} variant {v}")
            np.testing.assert_array_equal(sim.state(), state, err_msg=f"{case} variant {v}")
