"""Ingestion, window planning and slice packing of libreplay_hip, checked on
the CPU: the packed slices evaluated on the host (same rules as the device
slice kernel) equal the oracle's per-iteration CSR bit for bit."""
import ctypes as C

import numpy as np
import pytest

import hostlib
import replay_amd
from replay_amd import Params

from test_gpu_parity import STEP_WINDOWS, _random_case


@pytest.fixture(scope="module")
def lib():
    L = hostlib.load()
    P = C.POINTER
    L.rp_host_plan_open.argtypes = [C.c_int32, C.c_int64, P(C.c_int32), P(C.c_int32),
                                    P(C.c_int64), P(C.c_int64), P(replay_amd._Params),
                                    P(C.c_void_p)]
    L.rp_host_plan_nlists.argtypes = [C.c_void_p]
    L.rp_host_plan_nlists.restype = C.c_int64
    L.rp_host_plan_epochs.argtypes = [C.c_void_p]
    L.rp_host_slice.argtypes = [C.c_void_p, C.c_int32, P(C.c_int32), P(C.c_int32), P(C.c_float),
                                C.c_int64, P(C.c_int64)]
    L.rp_host_plan_close.argtypes = [C.c_void_p]
    return L


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class HostPlan:
    def __init__(self, lib, n, a, b, s, e, params: Params):
        self.lib = lib
        self.n = n
        self.h = C.c_void_p()
        self._keep = [np.ascontiguousarray(x, dt) for x, dt in
                      ((a, np.int32), (b, np.int32), (s, np.int64), (e, np.int64))]
        aa, bb, ss, ee = self._keep
        pc = params._c()
        rc = lib.rp_host_plan_open(n, len(aa), _p(aa, C.c_int32), _p(bb, C.c_int32),
                                   _p(ss, C.c_int64), _p(ee, C.c_int64), C.byref(pc),
                                   C.byref(self.h))
        assert rc == 0, rc

    def slice(self, k, cap):
        rp = np.zeros(self.n + 1, np.int32)
        col = np.zeros(cap, np.int32)
        val = np.zeros(cap, np.float32)
        nnz = C.c_int64()
        assert self.lib.rp_host_slice(self.h, k, _p(rp, C.c_int32), _p(col, C.c_int32),
                                      _p(val, C.c_float), cap, C.byref(nnz)) == 0
        return rp, col[:nnz.value], val[:nnz.value]

    def close(self):
        self.lib.rp_host_plan_close(self.h)


def _oracle_params(oracle, p):
    return oracle.make_params(M=p.M, step_size=p.step_size, iterations=p.iterations,
                              start_time=p.start_time,
                              static_network_duration=p.static_network_duration)


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("step,D", STEP_WINDOWS)
@pytest.mark.parametrize("wpe", [1, 3, 16, 0])
def test_host_slices_match_oracle(lib, oracle, seed, step, D, wpe):
    n, a, b, s, e = _random_case(seed)
    p = Params(M=1, step_size=step, static_network_duration=D, iterations=120,
               windows_per_epoch=wpe)
    g = oracle.OracleGraph.from_contacts(n, a, b, s, e)
    hp = HostPlan(lib, n, a, b, s, e, p)
    assert lib.rp_host_plan_nlists(hp.h) == g.nlists
    op = _oracle_params(oracle, p)
    cap = 2 * len(a)
    for k in range(p.iterations):
        rp, col, val = hp.slice(k, cap)
        orp, ocol, oval, _ = g.step_csr(op, k)
        np.testing.assert_array_equal(rp, orp)
        np.testing.assert_array_equal(col, ocol)
        np.testing.assert_array_equal(val.view(np.int32), oval.view(np.int32))
    hp.close()


def test_host_slices_truncation(lib, oracle):
    """Every contact spans every window: the cap cuts each window's last hits."""
    rng = np.random.default_rng(4)
    n, E = 12, 40
    a = rng.integers(0, n, E)
    b = (a + 1 + rng.integers(0, n - 1, E)) % n
    s = 946713600 - rng.integers(1, 5000, E)
    e = 946713600 + 10 ** 5 + rng.integers(0, 5000, E)
    p = Params(M=1, step_size=1800.0, static_network_duration=3600, iterations=40)
    g = oracle.OracleGraph.from_contacts(n, a, b, s, e)
    op = _oracle_params(oracle, p)
    hp = HostPlan(lib, n, a, b, s, e, p)
    for k in range(p.iterations):
        rp, col, val = hp.slice(k, 4 * E)
        orp, ocol, oval, _ = g.step_csr(op, k)
        np.testing.assert_array_equal(rp, orp)
        np.testing.assert_array_equal(col, ocol)
        np.testing.assert_array_equal(val.view(np.int32), oval.view(np.int32))
    hp.close()


def _large_case(seed=9, n=3000, E=160000, hub_deg=3000):
    """Enough contacts that sorting, span, fill and row passes split across
    worker threads: long contacts over two days, duplicates (equal start and
    end, so rank order among equal columns matters), and a hub whose row is
    long in every epoch."""
    rng = np.random.default_rng(seed)
    a = rng.integers(1, n, E)
    b = (a + 1 + rng.integers(0, n - 2, E)) % n
    b[b == 0] = 1
    s = 946713600 + rng.integers(-6 * 3600, 2 * 86400, E)
    d = rng.integers(60, 86400, E)
    dup = rng.uniform(size=E) < 0.1
    s[1:][dup[1:]] = s[:-1][dup[1:]]
    d[1:][dup[1:]] = d[:-1][dup[1:]]
    ha = np.zeros(hub_deg, np.int64)
    hb = rng.integers(1, n, hub_deg)
    hs = 946713600 - rng.integers(3600, 7200, hub_deg)
    he = 946713600 + 3 * 86400 + rng.integers(0, 3600, hub_deg)
    a = np.concatenate([a, ha]); b = np.concatenate([b, hb])
    s = np.concatenate([s, hs]); e = np.concatenate([s[:E] + d, he])
    return n, a.astype(np.int32), b.astype(np.int32), s.astype(np.int64), e.astype(np.int64)


@pytest.mark.parametrize("step,D,wpe", [(3600.0, 3600, 8), (1800.0, 3600, 4)])
def test_host_slices_parallel_build(lib, oracle, step, D, wpe):
    """The multi-threaded ingest (stable parallel sort, spans, partitioned
    fill, row sorts, long-row lists) on a graph large enough to split:
    every window's CSR equals the oracle's bit for bit, and each epoch's long
    rows are exactly its rows above the threshold, in row order, with the
    integer flag of their weight bound."""
    P = C.POINTER
    lib.rp_host_plan_epoch.argtypes = [C.c_void_p, C.c_int32, P(C.c_int32), P(C.c_int64),
                                       C.c_int64]
    lib.rp_host_plan_epoch.restype = C.c_int64
    lib.rp_host_plan_long_rows.argtypes = [C.c_void_p, C.c_int32, P(C.c_int32), P(C.c_uint8),
                                           C.c_int64]
    lib.rp_host_plan_long_rows.restype = C.c_int64
    lib.rp_host_long_row_threshold.restype = C.c_int32
    n, a, b, s, e = _large_case()
    p = Params(M=1, step_size=step, static_network_duration=D, iterations=int(48 * 3600 / step),
               windows_per_epoch=wpe)
    g = oracle.OracleGraph.from_contacts(n, a, b, s, e)
    hp = HostPlan(lib, n, a, b, s, e, p)
    assert lib.rp_host_plan_nlists(hp.h) == g.nlists
    op = _oracle_params(oracle, p)
    cap = 2 * len(a)
    for k in range(0, p.iterations, 3):
        rp, col, val = hp.slice(k, cap)
        orp, ocol, oval, _ = g.step_csr(op, k)
        np.testing.assert_array_equal(rp, orp)
        np.testing.assert_array_equal(col, ocol)
        np.testing.assert_array_equal(val.view(np.int32), oval.view(np.int32))
    thr = lib.rp_host_long_row_threshold()
    total = 0
    for ep in range(lib.rp_host_plan_epochs(hp.h)):
        rowptr = np.zeros(n + 1, np.int32)
        ent = np.zeros((4 * len(a), 4), np.int64)
        m = lib.rp_host_plan_epoch(hp.h, ep, _p(rowptr, C.c_int32), _p(ent, C.c_int64), len(ent))
        assert m >= 0
        total += m
        lens = np.diff(rowptr)
        # every row: columns ascending, ranks ascending among equal columns
        x = ent[:m]
        row_of = np.repeat(np.arange(n), lens)
        inner = row_of[1:] == row_of[:-1]
        dcol = np.diff(x[:, 0])[inner]
        drank = np.diff(x[:, 3])[inner]
        assert np.all(dcol >= 0)
        assert np.all(drank[dcol == 0] > 0)
        want = np.flatnonzero(lens > thr)
        rows = np.zeros(n, np.int32)
        flags = np.zeros(n, np.uint8)
        got = lib.rp_host_plan_long_rows(hp.h, ep, _p(rows, C.c_int32), _p(flags, C.c_uint8), n)
        np.testing.assert_array_equal(rows[:got], want)
        for j, r in enumerate(want):
            x = ent[rowptr[r]:rowptr[r + 1]]
            bound = np.minimum(x[:, 2] - x[:, 1], D).sum()
            assert flags[j] == (1 if bound < (1 << 24) else 0)
        assert want.size > 0  # the hub is long in every epoch
    assert total > 8 * 65536  # the fill split into several parts
    hp.close()
