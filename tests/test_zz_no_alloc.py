"""replay_get_exposure_noalloc against replay_get_exposure, through the C ABI
(needs an MI355X).

The no-allocation entry point stages the exposure matrix in the library's
per-device static array instead of a hipMalloc'd buffer; its output must be
bit-identical to the allocating entry point's.  Two contexts in a row, the
second smaller than the first (the array keeps the first call's contents),
integer and float windows, and the size limit."""
import ctypes as C

import numpy as np
import pytest

import replay_amd
from replay_amd import Params, Simulation

from test_gpu_parity import _random_case

pytestmark = pytest.mark.gpu

LIMIT = 1 << 24  # floats of the static staging array (64 MiB)


def _noalloc(sim: Simulation, k: int) -> np.ndarray:
    """The entry point called through the C ABI into a NaN-filled buffer."""
    L = replay_amd.load_library()
    out = np.full(sim.n_nodes * sim.M, np.nan, np.float32)
    rc = L.replay_get_exposure_noalloc(sim._ctx, k, out.ctypes.data_as(C.POINTER(C.c_float)),
                                       out.size)
    sim._check(rc)
    return out


@pytest.mark.parametrize("D", [1800, 3600])
def test_noalloc_matches_allocating_entry(D):
    cases = [(300, 256, 11), (40, 64, 5)]  # the second call is smaller
    for seed, (n, M, k) in enumerate(cases):
        n, a, b, s, e = _random_case(900 + seed, n=n, e=n * 15)
        p = Params(M=M, step_size=1800.0, static_network_duration=D, iterations=24,
                   initial_infected=0.2, infect_prob=0.3, exposed_duration=1800,
                   infectious_duration=7200, resistant_duration=3600, seed=seed)
        print(f"noalloc case n={n} M={M} k={k} D={D}", flush=True)
        with Simulation.from_contacts(n, a, b, s, e, p) as sim:
            sim.run(k)
            sim.sync()
            ref = sim.exposure(k)
            got = _noalloc(sim, k)
            np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
            assert np.count_nonzero(ref) > 0
            # a second call on the same context, another window
            sim.run(3)
            sim.sync()
            np.testing.assert_array_equal(_noalloc(sim, k + 3).view(np.uint32),
                                          sim.exposure(k + 3).view(np.uint32))


def test_noalloc_rejects_beyond_limit():
    n, M = 65600, 256  # N * M = 16,793,600 > 2^24
    assert n * M > LIMIT
    rng = np.random.default_rng(5)
    a = rng.integers(0, n, 100).astype(np.int32)
    b = ((a + 1) % n).astype(np.int32)
    s = (946713600 + rng.integers(0, 3600, 100)).astype(np.int64)
    p = Params(M=M, step_size=3600.0, static_network_duration=3600, iterations=2, seed=1)
    with Simulation.from_contacts(n, a, b, s, s + 600, p) as sim:
        with pytest.raises(replay_amd.ReplayError) as ei:
            _noalloc(sim, 0)
        assert ei.value.code == replay_amd.E_ARG
        assert "no-allocation staging limit" in str(ei.value)
