"""Countdowns above 2^24 s against the oracle (needs an MI355X).

The reference keeps each replica's countdown as an int and steps it with
c = (int)((float)c - step) (tool/gpu_cpu_temporal_sim.cu:161-176, 489-500),
after c0 = (int)((float)upper + step) (cu:153-158).  Above 2^24 the float
conversions round, so a countdown falls by more or less than `step` per
iteration and the E/I/R boundaries move; the reference's own documented
settings reach this range (resistant duration 99,999,999 s, step 30 s, in
screenshot/replay_params.png).  The engine tabulates the same float sequence
on the host; these tests run the cases where it is inexact, assert on the
host (numpy float32, independent of the engine) that the sequence really has
inexact steps, and require the oracle's counts and countdown states bit for
bit."""
import numpy as np
import pytest

import replay_amd
from replay_amd import Params, Simulation

from test_gpu_parity import _random_case, _to_oracle_params

pytestmark = pytest.mark.gpu


def countdown_sequence(upper: int, step: float, n: int) -> np.ndarray:
    """c(0..n) of a new infection with the reference's float operations."""
    f = np.float32
    out = np.empty(n + 1, np.int64)
    c = int(f(f(upper) + f(step)))  # cu:155
    out[0] = c
    for j in range(1, n + 1):
        if c > 0:
            c = int(f(f(c) - f(step)))  # cu:169-170, truncation toward zero
        if c < 0:
            c = 0
        out[j] = c
    return out


def inexact_steps(seq: np.ndarray, step: float, upper: int) -> int:
    """Offsets whose countdown differs from exact arithmetic on its
    predecessor: trunc(upper + step) for the first, trunc(c - step) after
    (integers and halves, exact in float64)."""
    prev = seq[:-1]
    exact = np.where(prev > 0, np.trunc(prev.astype(np.float64) - step), 0.0)
    exact = np.maximum(exact, 0.0).astype(np.int64)
    first = int(np.trunc(float(upper) + step)) != seq[0]
    return int(first) + int(np.count_nonzero(exact != seq[1:]))


def first_at_least(seq: np.ndarray, upper: int, elapsed: int) -> int:
    """First offset whose elapsed time (upper - c) reaches `elapsed`."""
    el = upper - seq
    return int(np.argmax((seq > 0) & (el >= elapsed)))


def _run_case(oracle, n, a, b, s, e, p, want_states=True):
    """The oracle's counts and states, bit for bit, with every form of the
    infection ring: every iteration kept (the default for runs of up to
    HISTORY_MAX_ITERATIONS), slots back to the longest timer offset
    (REPLAY_F_NO_HISTORY, and the default for longer runs), and the compact
    form (forced by a one-byte ring budget)."""
    from dataclasses import replace
    g = oracle.OracleGraph.from_contacts(n, a, b, s, e)
    oc, ostate, _ = g.simulate(_to_oracle_params(oracle, p, threads=8), want_state=want_states)
    default = (replay_amd.RING_HISTORY if p.iterations <= replay_amd.HISTORY_MAX_ITERATIONS
               else replay_amd.RING_OFFSETS)
    forms = ((dict(), default),
             (dict(flags=replay_amd.F_NO_HISTORY), replay_amd.RING_OFFSETS),
             (dict(max_ring_bytes=1), replay_amd.RING_COMPACT))
    for kw, form in forms:
        with Simulation.from_contacts(n, a, b, s, e, replace(p, **kw)) as sim:
            info = sim.info()
            assert info["ring_form"] == form
            assert info["ring_depth"] >= 1
            sim.run()
            counts, state = sim.counts(), sim.state()
        np.testing.assert_array_equal(counts, oc, err_msg=f"ring form {form}")
        if want_states:
            np.testing.assert_array_equal(state.reshape(-1), ostate, err_msg=f"ring form {form}")
    return counts


def test_resistant_99999999_at_30s_steps(oracle):
    """R = 99,999,999 s and I = 7 days at 30 s steps (the reference's
    RollerNet settings): upper is about 1.006e8, where floats are 8 apart,
    so each step takes 32 s off the countdown instead of 30 and the I -> R
    boundary comes about 1260 iterations early.  The horizon runs past it."""
    Ed, Id, Rd, step = 0, 604800, 99_999_999, 30.0
    upper = Ed + Id + Rd
    seq = countdown_sequence(upper, step, 25000)
    assert inexact_steps(seq, step, upper) > 1000
    n_off = first_at_least(seq, upper, Ed + Id)
    exact_off = int(np.ceil((Ed + Id) / step))
    assert 0 < n_off < exact_off - 500, (n_off, exact_off)
    T = n_off + 150
    n, a, b, s, e = _random_case(41, n=60, e=900, span=int(T * step) + 3600, long_frac=0.5)
    p = Params(M=64, step_size=step, static_network_duration=30, iterations=T,
               initial_infected=0.1, infect_prob=0.2, exposed_duration=Ed, infectious_duration=Id,
               resistant_duration=Rd, seed=3, start_time=946713600)
    print(f"T={T} n_off={n_off} exact={exact_off}", flush=True)
    counts = _run_case(oracle, n, a, b, s, e, p)
    # replicas infected early have left I for R by the end
    assert counts[-1, :, 2].sum() > 0 and counts[:, :, 1].max() > 0


def test_resistant_99999999_fractional_step(oracle):
    """R = 99,999,999 s with a 1234.5 s step and a 1800 s window: fractional
    steps and weights together with rounded countdowns."""
    Ed, Id, Rd, step = 3000, 90000, 99_999_999, 1234.5
    upper = Ed + Id + Rd
    seq = countdown_sequence(upper, step, 200)
    assert inexact_steps(seq, step, upper) > 50
    n, a, b, s, e = _random_case(42, n=80, e=2500, span=200 * 1300, long_frac=0.3)
    p = Params(M=130, step_size=step, static_network_duration=1800, iterations=200,
               initial_infected=0.05, infect_prob=0.3, exposed_duration=Ed,
               infectious_duration=Id, resistant_duration=Rd, seed=4)
    counts = _run_case(oracle, n, a, b, s, e, p)
    assert counts[-1, :, 2].sum() > 0


def test_sirs_beyond_2_24(oracle):
    """SIRS with E + I + R = 20,000,000 s > 2^24 at 99,999 s steps: floats
    are 2 apart there, so every step of an odd length rounds while the
    countdown exceeds 2^24 (the first 33); the replicas return to S within
    the horizon and are infected again."""
    Ed, Id, Rd, step = 0, 18_000_000, 2_000_000, 99999.0
    upper = Ed + Id + Rd
    assert upper > 1 << 24
    seq = countdown_sequence(upper, step, 400)
    assert inexact_steps(seq, step, upper) > 30
    n_s = int(np.argmax(seq == 0))
    assert 0 < n_s < 320
    n, a, b, s, e = _random_case(43, n=70, e=3000, span=400 * 100000, long_frac=0.6)
    p = Params(M=100, step_size=step, static_network_duration=100000, iterations=400,
               initial_infected=0.1, infect_prob=0.01, exposed_duration=Ed,
               infectious_duration=Id, resistant_duration=Rd, seed=5)
    counts = _run_case(oracle, n, a, b, s, e, p)
    sus = counts[:, :, 3].sum(axis=1)
    assert (np.diff(sus) > 0).any(), "no replica returned to S"
    assert (np.diff(sus[n_s + 2:]) < 0).any(), "no infection after a return to S"


def test_hub_rows_with_big_timers(oracle):
    """Hub rows beyond the long-row threshold (integer and float-order sums)
    with R = 999,999,999 s: floats are 64 apart there and 3600 is not a
    multiple of 64, so every step rounds; long_row_kernel's draws feed the
    rounded timer sequence."""
    from test_gpu_dense import _two_hubs
    Ed, Id, Rd, step = 3600, 5 * 3600, 999_999_999, 3600.0
    seq = countdown_sequence(Ed + Id + Rd, step, 16)
    assert inexact_steps(seq, step, Ed + Id + Rd) >= 16
    n, a, b, s, e = _two_hubs(seed=6, leaves=(2600, 4800), T=16)
    p = Params(M=160, step_size=step, static_network_duration=3600, iterations=16,
               initial_infected=0.03, infect_prob=0.05, exposed_duration=Ed,
               infectious_duration=Id, resistant_duration=Rd, seed=8)
    with Simulation.from_contacts(n, a, b, s, e, p) as sim:
        sim.run(1)
        assert sim.info()["n_entries"] > 0
    for v in (None, replay_amd.V_WORD):
        g = oracle.OracleGraph.from_contacts(n, a, b, s, e)
        oc, ostate, _ = g.simulate(_to_oracle_params(oracle, p, threads=8), want_state=True)
        with Simulation.from_contacts(n, a, b, s, e, p) as sim:
            if v is not None:
                sim.set_variant(v)
            sim.run()
            np.testing.assert_array_equal(sim.counts(), oc, err_msg=f"variant {v}")
            np.testing.assert_array_equal(sim.state(), ostate, err_msg=f"variant {v}")
        assert oc[-1, :, 1:3].sum() > 0


@pytest.mark.parametrize("case", range(12))
def test_compact_ring_random_matches_oracle(oracle, case):
    """Random graphs and timer settings (SEIRS with all three flip offsets
    inside the horizon, SIS, permanent R) with the compact ring forced:
    counts and states equal the oracle's under every variant."""
    rng = np.random.default_rng(4400 + case)
    n = int(rng.choice([20, 70, 150]))
    M = int(rng.choice([1, 64, 130, 256, 1024]))
    step, D = [(600.0, 600), (600.0, 1800), (1000.0, 3600)][int(rng.integers(0, 3))]
    Ed, Id, Rd = [(1200, 5400, 3600), (0, 7200, 0), (1800, 3000, 10 ** 7),
                  (600, 1800, 1200)][int(rng.integers(0, 4))]
    nn, a, b, s, e = _random_case(int(rng.integers(0, 1 << 30)), n=n, e=n * 20, long_frac=0.3)
    p = Params(M=M, step_size=step, static_network_duration=D, iterations=60,
               initial_infected=0.1, infect_prob=float(rng.choice([0.1, 0.4])),
               exposed_duration=Ed, infectious_duration=Id, resistant_duration=Rd,
               seed=int(rng.integers(0, 1000)), max_ring_bytes=1)
    print(f"case {case}: n={n} M={M} step={step} D={D} E,I,R={Ed},{Id},{Rd}", flush=True)
    g = oracle.OracleGraph.from_contacts(nn, a, b, s, e)
    oc, ostate, _ = g.simulate(_to_oracle_params(oracle, p, threads=8), want_state=True)
    with Simulation.from_contacts(nn, a, b, s, e, p) as sim:
        assert sim.info()["ring_form"] in (replay_amd.RING_COMPACT, replay_amd.RING_OFFSETS)
        for v in replay_amd.VARIANTS_ALL:
            sim.set_variant(v)
            sim.reset()
            sim.run()
            np.testing.assert_array_equal(sim.counts(), oc, err_msg=f"variant {v}")
            np.testing.assert_array_equal(sim.state(), ostate, err_msg=f"variant {v}")


def test_ring_beyond_device_memory_runs_compact():
    """50,000 nodes x 4096 replicas at the reference's RollerNet settings
    (30 s steps, 7-day infectious period, R = 99,999,999 s) over 20,400
    iterations: the infectious period ends about 18,900 iterations after an
    infection, so a full ring would hold that many slots of 25.6 MB, about
    484 GB, more than the 288 GB of the device.  The context opens in the
    compact form by default and runs; replicas 0-63 equal a 64-replica run
    of the same network with the full ring (a replica's draws do not depend
    on M)."""
    Ed, Id, Rd, step = 0, 604800, 99_999_999, 30.0
    seq = countdown_sequence(Ed + Id + Rd, step, 25000)
    n_off = first_at_least(seq, Ed + Id + Rd, Ed + Id)
    T = n_off + 1500
    n, M = 50_000, 4096
    depth_bytes = n * (M // 64) * 8 * n_off
    assert depth_bytes > 288e9, depth_bytes
    rng = np.random.default_rng(77)
    k = 400_000
    a = rng.integers(0, n, k)
    b = (a + 1 + rng.integers(0, n - 1, k)) % n
    s = 946713600 + rng.integers(0, int(T * step), k)
    e = s + np.minimum(rng.exponential(3600, k).astype(np.int64) + 1, 8 * 3600)
    p = Params(M=M, step_size=step, static_network_duration=30, iterations=T,
               initial_infected=2e-4, infect_prob=0.3, exposed_duration=Ed,
               infectious_duration=Id, resistant_duration=Rd, seed=11)
    args = (n, a.astype(np.int32), b.astype(np.int32), s.astype(np.int64), e.astype(np.int64))
    with Simulation.from_contacts(*args, p) as sim:
        info = sim.info()
        print(f"T={T} ring_form={info['ring_form']} ring_bytes={info['ring_bytes']}", flush=True)
        assert info["ring_form"] == replay_amd.RING_COMPACT and info["ring_bytes"] < 16e9
        sim.run()
        counts = sim.counts()
    from dataclasses import replace
    with Simulation.from_contacts(*args, replace(p, M=64)) as small:
        assert small.info()["ring_form"] != replay_amd.RING_COMPACT
        small.run()
        c64 = small.counts()
    np.testing.assert_array_equal(counts[:, :64], c64)
    assert (counts.sum(axis=2) == n).all()
    assert counts[-1, :, 2].sum() > 0  # some replicas reached R within the horizon
