"""Pins of the oracle's arithmetic: Philox, exp, windows, CSR, exposure,
classification and whole trajectories, each against something the
mathematics or an independent implementation fixes."""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

# Random123 known-answer vectors for philox4x32-10 (counter, key, output).
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


def test_philox_known_answers(oracle):
    for ctr, key, out in KAT:
        assert tuple(oracle.philox4x32_10(ctr, key)) == out


def test_philox_matches_torch_engine(oracle, tmp_path):
    """torch's at::philox_engine (ATen/core/PhiloxRNGEngine.h, header only)
    as an independent implementation: key = seed, counter = {offset, subseq}."""
    import torch
    inc = os.path.join(os.path.dirname(torch.__file__), "include")
    if not os.path.exists(os.path.join(inc, "ATen", "core", "PhiloxRNGEngine.h")):
        pytest.skip("torch headers not present")
    src = tmp_path / "p.cpp"
    src.write_text(
        "#include <ATen/core/PhiloxRNGEngine.h>\n#include <cstdio>\n#include <cstdlib>\n"
        "int main(int c, char** v){ for(int i=1;i+2<c;i+=3){"
        " at::philox_engine e(strtoull(v[i],0,10), strtoull(v[i+1],0,10), strtoull(v[i+2],0,10));"
        " for(int k=0;k<4;k++) printf(\"%u \", e()); printf(\"\\n\");} }\n")
    exe = tmp_path / "p"
    r = subprocess.run(["g++", "-std=c++17", f"-I{inc}", str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("cannot compile torch philox header: " + r.stderr[:200])
    rng = np.random.default_rng(1)
    cases = [(int(rng.integers(0, 2**63)), int(rng.integers(0, 2**63)), int(rng.integers(0, 2**63)))
             for _ in range(64)]
    args = [str(x) for c in cases for x in c]
    out = subprocess.run([str(exe), *args], capture_output=True, text=True, check=True).stdout.split("\n")
    for (seed, subseq, offset), line in zip(cases, out):
        ctr = (offset & 0xffffffff, offset >> 32, subseq & 0xffffffff, subseq >> 32)
        key = (seed & 0xffffffff, seed >> 32)
        assert oracle.philox4x32_10(ctr, key) == [int(x) for x in line.split()]


def test_uniform_definition(oracle):
    """Replica r takes word (r >> 6) & 3 of Philox(counter {node, (r >> 8) * 64
    + (r & 63), slot, stream}): four replicas 64 apart share one block."""
    seed = 0x1234567890
    for stream in (0, 9):
        p = oracle.make_params(M=4096, seed=seed, stream=stream)
        for node, rep, slot in [(0, 0, 0), (5, 2, 17), (123456, 1, 4032), (7, 200, 3),
                                (7, 4095, 3), (11, 1000, 8)]:
            ctr = (node, ((rep >> 8) << 6) | (rep & 63), slot, stream)
            words = oracle.philox4x32_10(ctr, (seed & 0xffffffff, seed >> 32))
            u = oracle.uniform(p, slot, node, rep, 10)
            assert u == (words[(rep >> 6) & 3] >> 8) * 2.0 ** -24
            assert 0.0 <= u < 1.0
        # the four replicas sharing a block take its four words, in order
        us = [oracle.uniform(p, 5, 3, 64 * k + 9, 10) for k in range(4)]
        words = oracle.philox4x32_10((3, 9, 5, stream), (seed & 0xffffffff, seed >> 32))
        assert us == [(x >> 8) * 2.0 ** -24 for x in words]


def _ulp_dist(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    ia = a.view(np.int32).astype(np.int64)
    ib = b.view(np.int32).astype(np.int64)
    return np.abs(ia - ib)


def test_expf_accuracy(oracle):
    """or_expf against exp in float64 rounded to float32 (correctly rounded
    except within 1e-16 of a midpoint) and against glibc expf."""
    import ctypes as C
    lib = C.CDLL("libm.so.6")
    lib.expf.argtypes = [C.c_float]
    lib.expf.restype = C.c_float
    rng = np.random.default_rng(7)
    xs = np.concatenate([rng.uniform(-104, 0, 200000), rng.uniform(-20, 0, 200000),
                         rng.uniform(-1e-3, 0, 20000), rng.uniform(0, 88.7, 20000),
                         np.array([0.0, -0.0, -1e-30, -87.3, -103.9, -104.0, 88.72])]).astype(np.float32)
    got = np.array([oracle.expf(float(x)) for x in xs], np.float32)
    ref = np.exp(xs.astype(np.float64)).astype(np.float32)
    d = _ulp_dist(got, ref)
    assert d.max() <= 1
    assert (d != 0).mean() < 1e-4
    libm = np.array([lib.expf(float(x)) for x in xs[:50000]], np.float32)
    d2 = _ulp_dist(got[:50000], libm)
    assert d2.max() <= 1
    assert (d2 != 0).mean() < 2e-3
    assert math.isnan(oracle.expf(float("nan")))
    assert oracle.expf(float("-inf")) == 0.0
    assert oracle.expf(float("inf")) == float("inf")
    assert oracle.expf(0.0) == 1.0


def test_window_arithmetic_closed_forms(oracle):
    import ctypes as C

    class Cache(C.Structure):
        _fields_ = [("valid", C.c_int), ("start", C.c_int64), ("steps", C.c_int)]

    def run(step, D, T, start=946713600):
        p = oracle.make_params(step_size=step, static_network_duration=D, iterations=T,
                               start_time=start)
        c = Cache(0, 0, 1)
        out = []
        for it in range(T):
            t, ws, we, sc = C.c_int64(), C.c_int64(), C.c_int64(), C.c_float()
            assert oracle.lib().or_window(C.byref(p), C.byref(c), it, C.byref(t), C.byref(ws),
                                          C.byref(we), C.byref(sc)) == 0
            out.append((t.value, ws.value, we.value, sc.value))
        return out

    # window == step: one window per iteration, divisor 1
    for k, (t, ws, we, sc) in enumerate(run(600.0, 600, 20)):
        assert (t, ws, we, sc) == (946713600 + 600 * k, 946713600 + 600 * k, 946713600 + 600 * (k + 1), 1.0)
    # three iterations per window, the last window holds the remaining two
    res = run(1200.0, 3600, 8)
    scales = [r[3] for r in res]
    assert scales[:6] == [np.float32(1 / 3)] * 6 and scales[6:] == [0.5, 0.5]
    assert [r[1] for r in res] == [946713600 + 3600 * (k // 3) for k in range(8)]
    # non-integer steps: offsets are llround(iter * step), halves away from zero
    res = run(1.5, 2, 6)
    assert [r[0] - 946713600 for r in res] == [0, 2, 3, 5, 6, 8]


def _small_graph(oracle, seed, n=9, e=40, span=20000):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, n, e)
    b = (a + 1 + rng.integers(0, n - 1, e)) % n
    s = 946713600 + rng.integers(0, span, e)
    d = rng.integers(0, 4000, e)
    return oracle.OracleGraph.from_contacts(n, a, b, s, s + d), (a, b, s, s + d)


@pytest.mark.parametrize("seed", range(5))
@pytest.mark.parametrize("step,D", [(600.0, 600), (600.0, 1800), (1000.0, 3600)])
def test_csr_against_dense_bruteforce(oracle, seed, step, D):
    g, (a, b, s, e) = _small_graph(oracle, seed)
    T = 30
    p = oracle.make_params(step_size=step, static_network_duration=D, iterations=T, M=1)
    iv = g.intervals()
    cap = max(g.nlists, 1)
    order = np.lexsort((-iv[:, 1], iv[:, 0]))
    win = [int(k * step) // D for k in range(T)]  # integer steps: offsets are exact
    for it in range(T):
        rowptr, col, val, (ws, we, sc) = g.step_csr(p, it)
        # closed form: the divisor is the number of iterations sharing the window
        assert sc == np.float32(1.0) / np.float32(win.count(win[it]))
        assert ws == 946713600 + win[it] * D and we == ws + D
        dense = np.zeros((g.n_nodes, g.n_nodes), np.float64)
        got = np.zeros_like(dense)
        for r in range(g.n_nodes):
            for j in range(rowptr[r], rowptr[r + 1]):
                got[r, col[j]] = val[j]
            assert (np.diff(col[rowptr[r]:rowptr[r + 1]]) > 0).all()
        ranked = iv[order]
        hits = ranked[(ranked[:, 0] < we) & (ws < ranked[:, 1])][:cap]
        for st, en, tid, _ in hits:
            ov = min(en, we) - max(st, ws)
            if ov > 0:
                u, v = tid >> 32, tid & 0xffffffff
                w = float(np.float32(np.float32(ov) * np.float32(sc)))
                dense[u, v] += w
                dense[v, u] += w
        np.testing.assert_allclose(got, dense, rtol=1e-6, atol=0)
        if sc == 1.0:
            np.testing.assert_array_equal(got, dense)


def test_spmm_order_and_fp64(oracle):
    import ctypes as C
    rng = np.random.default_rng(3)
    n, m = 200, 16
    dens = rng.uniform(size=(n, n)) < 0.05
    w = np.where(dens, rng.uniform(1, 300, (n, n)), 0).astype(np.float32)
    w = np.triu(w, 1)
    w = w + w.T
    rowptr = np.zeros(n + 1, np.int32)
    cols, vals = [], []
    for i in range(n):
        nz = np.nonzero(w[i])[0]
        cols += nz.tolist()
        vals += w[i, nz].tolist()
        rowptr[i + 1] = len(cols)
    col = np.array(cols, np.int32)
    val = np.array(vals, np.float32)
    B = (rng.uniform(size=n * m) < 0.3).astype(np.float32)
    out = np.zeros(n * m, np.float32)
    out64 = np.zeros(n * m, np.float64)
    P = lambda a, t: a.ctypes.data_as(C.POINTER(t))
    oracle.lib().or_spmm(P(rowptr, C.c_int32), P(col, C.c_int32), P(val, C.c_float), len(col), n,
                         P(B, C.c_float), P(out, C.c_float), m, 1)
    oracle.lib().or_spmm_f64(P(rowptr, C.c_int32), P(col, C.c_int32), P(val, C.c_float), n,
                             P(B, C.c_float), P(out64, C.c_double), m)
    np.testing.assert_allclose(out, out64, rtol=1e-6)
    # sequential ascending-column fp32 sum, written out independently
    Bm = B.reshape(m, n)
    for c in range(0, m, 5):
        for i in range(0, n, 7):
            acc = np.float32(0)
            for j in range(rowptr[i], rowptr[i + 1]):
                acc = np.float32(acc + val[j] * Bm[c, col[j]])
            assert out[c * n + i] == acc
    dense64 = (w.astype(np.float64) @ Bm.T.astype(np.float64)).T.ravel()
    np.testing.assert_allclose(out64, dense64, rtol=1e-12)


def test_classification_boundaries(oracle):
    """cu:530-550 on hand-derived cases (E=100, I=200, R=300; upper=600)."""
    p = oracle.make_params(exposed_duration=100, infectious_duration=200, resistant_duration=300)
    cases = {0: 0, 600: 1, 501: 1, 500: 2, 301: 2, 300: 3, 1: 3, 700: 1, -5: 1}
    for c, cls in cases.items():
        assert oracle.state_class(p, c) == cls, c
    p0 = oracle.make_params(exposed_duration=0, infectious_duration=200, resistant_duration=0)
    assert [oracle.state_class(p0, c) for c in (0, 200, 1, 201)] == [0, 2, 2, 1 * 0 + 2]


def test_trajectory_pure_timers(oracle):
    """No transmission (p=0), everyone seeded: counts follow the timers in
    closed form.  Seed c0 = (int)(medium + step); each step subtracts step."""
    E, I, R, step = 7200, 10800, 3600, 3600.0
    g, _ = _small_graph(oracle, 0, n=7)
    T = 8
    p = oracle.make_params(M=3, iterations=T, initial_infected=1.0, infect_prob=0.0,
                           exposed_duration=E, infectious_duration=I, resistant_duration=R,
                           step_size=step, static_network_duration=3600)
    counts, state, _ = g.simulate(p, want_state=True)
    up = E + I + R
    c = I + R + int(step)
    for k in range(T + 1):
        elapsed = 0 if (c <= 0 or c > up) else up - c
        if c == 0:
            cls = 3
        elif elapsed < E:
            cls = 0
        elif elapsed < E + I:
            cls = 1
        elif elapsed < E + I + R:
            cls = 2
        else:
            cls = 3
        expect = [0, 0, 0, 0]
        expect[cls] = 7
        assert counts[k].tolist() == [expect] * 3, k
        c = max(0, int(c - step)) if c > 0 else 0
    assert (state == 0).all()


def _chain_sim(oracle, with_dummy):
    n = 5
    a, b = [0, 1], [1, 2]
    s = [946713600 - 10] * 2
    e = [946713600 - 10 + 10 ** 6] * 2
    if with_dummy:  # starts later: last in rank order, nested under contact (1,2)
        a.append(3); b.append(4); s.append(946713600 - 5); e.append(e[0])
    g = oracle.OracleGraph.from_contacts(n, a, b, s, e)
    u = np.full((7, 1, n), 0.999, np.float32)
    u[0, 0, :] = [0.0, 1.0, 1.0, 1.0, 1.0]  # seed node 0 only (u < p0 = 0.5)
    p = oracle.make_params(M=1, iterations=6, initial_infected=0.5, infect_prob=1.0,
                           exposed_duration=0, infectious_duration=100000, resistant_duration=0,
                           step_size=3600.0, static_network_duration=3600, uniforms=u)
    counts, _, _ = g.simulate(p, want_state=True)
    return g, counts[:, 0, 1].tolist()


def test_trajectory_certain_infection(oracle):
    """p = 1 and contacts spanning every window: a susceptible neighbour of an
    infectious node is infected in that step (P = 1 > u).  The seed is not
    infectious during the first step (c0 = medium + step, cu:85)."""
    g, infectious = _chain_sim(oracle, with_dummy=True)
    assert g.nlists == 2
    assert infectious == [1, 1, 2, 3, 3, 3, 3]


def test_trajectory_hit_cap_drops_last_contact(oracle):
    """Two identical contacts are two top-level intervals: nlists = 1, so every
    window returns only the first (cu:988-989, idb:258-261, 465-467) and
    node 2 is never exposed."""
    g, infectious = _chain_sim(oracle, with_dummy=False)
    assert g.nlists == 1
    assert infectious == [1, 1, 2, 2, 2, 2, 2]
