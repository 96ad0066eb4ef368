"""Seeded random configurations against the oracle (needs an MI355X).

Each case draws a small contact graph and a configuration -- replicas M
(one to five words, whole or partial), step and window (integer and float
windows), windows per epoch, durations (zero included), seeding and
infection probability, a forced exposure variant, Philox or an injected
uniform stream -- and requires the oracle's counts and countdown states
bit for bit.  The forced variant includes the block-width and XCD-mapping
schedules.  300 cases take a few seconds on one GPU."""
import numpy as np
import pytest

import replay_amd
from replay_amd import Params, Simulation
from test_gpu_parity import _random_case, _to_oracle_params

pytestmark = pytest.mark.gpu

VARIANTS = [None, *replay_amd.VARIANTS_ALL]


def _case(i):
    rng = np.random.default_rng(1000 + i)
    step = float(rng.choice([300.0, 600.0, 900.0, 1234.5, 3600.0]))
    D = int(rng.choice([300, 600, 1800, 3600]))
    M = int(rng.choice([1, 7, 63, 64, 65, 128, 200, 256, 300]))
    dur = lambda: int(rng.choice([0, int(step), int(rng.integers(1, 4) * step), int(rng.integers(1, 200000))]))
    p = Params(M=M, step_size=step, static_network_duration=D,
               iterations=int(rng.integers(20, 60)),
               initial_infected=float(rng.choice([0.02, 0.1, 0.3])),
               infect_prob=float(rng.choice([0.01, 0.2, 0.6, 1.0])),
               exposed_duration=dur(), infectious_duration=max(1, dur()), resistant_duration=dur(),
               windows_per_epoch=int(rng.choice([0, 1, 4])), seed=int(rng.integers(0, 1 << 30)))
    graph = _random_case(int(rng.integers(0, 1 << 20)), n=int(rng.integers(20, 90)),
                         e=int(rng.integers(200, 2500)), long_frac=float(rng.uniform(0, 0.6)))
    variant = VARIANTS[int(rng.integers(0, len(VARIANTS)))]
    inject = bool(rng.uniform() < 0.3)
    return p, graph, variant, inject, rng


@pytest.mark.parametrize("i", range(300))
def test_random_configuration(oracle, i):
    p, (n, a, b, s, e), variant, inject, rng = _case(i)
    u = None
    if inject:
        u = rng.uniform(size=(p.iterations + 1) * p.M * n).astype(np.float32)
    g = oracle.OracleGraph.from_contacts(n, a, b, s, e)
    counts, state, _ = g.simulate(_to_oracle_params(oracle, p, uniforms=u), want_state=True)
    with Simulation.from_contacts(n, a, b, s, e, p) as sim:
        if variant is not None:
            sim.set_variant(variant)
        if u is not None:
            sim.set_uniforms(u)
        sim.reset()
        sim.run()
        np.testing.assert_array_equal(sim.counts(), counts, err_msg=f"case {i}")
        np.testing.assert_array_equal(sim.state(), state, err_msg=f"case {i}")
