"""Ensemble agreement under independent random streams (two-sample KS).

The reference draws its uniforms from clock-seeded generators (curand
XORWOW on the GPU, rand() on the CPU path; tool/gpu_cpu_temporal_sim.cu:
81-82, 149-150, 470-480), so its trajectories cannot be replayed; what must
hold is that ensembles driven by this project's Philox stream and by an
independent generator are draws of the same process.  The independent
stream here is numpy's PCG64, injected through the oracle's uniform-stream
input, playing the part of the reference's generator.

CPU test: oracle (Philox) vs oracle (PCG64).  GPU test: the HIP engine
(Philox) vs the oracle (PCG64).  Statistics: final attack size, peak
prevalence and the day of the peak per replica; bar: KS p > 0.01 for each.
"""
import numpy as np
import pytest
from scipy import stats

from replay_amd import Params, workloads

M = 400


def _workload():
    w = workloads.config0_sir(n_nodes=150, n_contacts=5000, M=M, iterations=96)
    p = w.params
    p = Params(**{**p.__dict__, "initial_infected": 0.03, "infect_prob": 0.3,
                  "infectious_duration": 48 * 3600})
    return w, p


def _summaries(counts):
    # counts (T + 1, M, 4): exposed, infectious, resistant, susceptible
    inf = counts[:, :, 1]
    return {"attack": counts[-1, :, 2] + counts[-1, :, 1] + counts[-1, :, 0],
            "peak": inf.max(axis=0), "peak_time": inf.argmax(axis=0)}


def _oracle_counts(oracle, w, p, uniforms=None):
    g = oracle.OracleGraph.from_contacts(w.n_nodes, w.src, w.tgt, w.start, w.end)
    op = oracle.make_params(M=p.M, step_size=p.step_size, iterations=p.iterations,
                            initial_infected=p.initial_infected, infect_prob=p.infect_prob,
                            exposed_duration=p.exposed_duration,
                            infectious_duration=p.infectious_duration,
                            resistant_duration=p.resistant_duration, start_time=p.start_time,
                            static_network_duration=p.static_network_duration, seed=p.seed,
                            threads=4, uniforms=uniforms)
    counts, _, _ = g.simulate(op)
    return counts


def _pcg_uniforms(w, p, seed):
    rng = np.random.default_rng(seed)
    return rng.random(((p.iterations + 1) * p.M * w.n_nodes), dtype=np.float32)


def _assert_same_process(a, b):
    sa, sb = _summaries(a), _summaries(b)
    assert sa["attack"].std() > 0  # the statistics vary, so the test has power
    for k in sa:
        p = stats.ks_2samp(sa[k], sb[k]).pvalue
        assert p > 0.01, (k, p)


def test_oracle_philox_vs_independent_stream(oracle):
    w, p = _workload()
    a = _oracle_counts(oracle, w, p)
    b = _oracle_counts(oracle, w, p, _pcg_uniforms(w, p, 12345))
    _assert_same_process(a, b)


def test_ks_detects_a_changed_process(oracle):
    """Power check: a 30% lower transmission probability is rejected."""
    w, p = _workload()
    a = _oracle_counts(oracle, w, p)
    q = Params(**{**p.__dict__, "infect_prob": p.infect_prob * 0.7})
    b = _oracle_counts(oracle, w, q, _pcg_uniforms(w, q, 7))
    sa, sb = _summaries(a), _summaries(b)
    assert stats.ks_2samp(sa["attack"], sb["attack"]).pvalue < 0.01


@pytest.mark.gpu
def test_gpu_philox_vs_independent_stream(oracle):
    from replay_amd import Simulation
    w, p = _workload()
    with Simulation.from_contacts(w.n_nodes, w.src, w.tgt, w.start, w.end,
                                  Params(**{**p.__dict__, "seed": 99})) as sim:
        sim.run()
        a = sim.counts()
    b = _oracle_counts(oracle, w, p, _pcg_uniforms(w, p, 2024))
    _assert_same_process(a, b)
