"""Ingestion, window planning and slice packing of libreplay_hip, checked on
the CPU: the packed slices evaluated on the host (same rules as the device
slice kernel) equal the oracle's per-iteration CSR bit for bit."""
import ctypes as C
import os

import numpy as np
import pytest

import replay_amd
from replay_amd import Params

from test_gpu_parity import STEP_WINDOWS, _random_case


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(replay_amd.LIB_PATH):
        from replay_amd import build
        build.build()
    L = replay_amd.load_library()
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
