"""The contact-line parser's canonical-form fast path against the
reference's sscanf patterns (tcm:131-135, 272-295, 95-120), and the
product's file ingestion against the oracle on every fixture graph (CPU).

Fast path: on every line it accepts, the line yields the same ids, the same
two times and the same counted flag as the sscanf + timegm path; lines it
does not accept fall back to that path.  The fixtures are pinned to the
reference's own parser by test_oracle_golden.py."""
import ctypes as C
import os

import numpy as np
import pytest

import hostlib
from conftest import GRAPHS

FIXTURES = sorted(f for f in os.listdir(GRAPHS) if f.endswith(".txt"))


@pytest.fixture(scope="module")
def lib():
    L = hostlib.load()
    P = C.POINTER
    L.rp_host_parse_line.argtypes = [C.c_char_p, C.c_int64, C.c_int32, P(C.c_int32),
                                     P(C.c_int32), P(C.c_int64), P(C.c_int64), P(C.c_int32)]
    L.rp_host_parse_line.restype = C.c_int32
    L.rp_host_parse_file.argtypes = [C.c_char_p, P(C.c_int64), C.c_int64]
    L.rp_host_parse_file.restype = C.c_int64
    return L


def _parse(lib, line: bytes, fast: int):
    src, tgt, cnt = C.c_int32(), C.c_int32(), C.c_int32()
    s, e = C.c_int64(), C.c_int64()
    r = lib.rp_host_parse_line(line, len(line), fast, C.byref(src), C.byref(tgt), C.byref(s),
                               C.byref(e), C.byref(cnt))
    return r, (src.value, tgt.value, s.value, e.value), cnt.value


def _time(rng):
    y = int(rng.choice([1969, 1970, 1999, 2000, 2024, 2100, 1900, 9999]))
    mo = int(rng.choice([0, 1, 2, 6, 12, 13, 99]))
    d = int(rng.choice([0, 1, 15, 28, 29, 30, 31, 45, 99]))
    h, mi, se = (int(rng.choice([0, 1, 23, 24, 59, 60, 99])) for _ in range(3))
    t = f"{y:04d}-{mo:02d}-{d:02d}T{h:02d}:{mi:02d}:{se:02d}"
    return t + str(rng.choice(["", "Z", "+00:00", ".5Z", "0", "Zxyz", "+0530", "9"]))


def _canonical(rng):
    sp = lambda: str(rng.choice(["", " ", "  ", "\t"]))
    u, v = rng.integers(0, 10 ** 6, 2)
    return (sp() + f"{u},{sp()}{v},{sp()}{_time(rng)},{sp()}{_time(rng)};" +
            str(rng.choice(["\n", "\r\n", ""]))).encode()


def _line(rng):
    if rng.uniform() < 0.5:
        return _canonical(rng)
    sp = lambda: str(rng.choice(["", " ", "  ", "\t", " \t"]))
    u = str(rng.choice(["0", "7", "42", "007", "123456789", "1234567890", "-3", "+5", "x1"]))
    v = str(rng.choice(["1", "99", "00", "2147483647", "5 "]))
    t1, t2 = _time(rng), _time(rng)
    if rng.uniform() < 0.1:
        t1 = t1.replace("-", "/", 1)
    if rng.uniform() < 0.05:
        t2 = "x" * 70
    line = (sp() + u + str(rng.choice([",", " ,"])) + sp() + v + "," + sp() + t1 +
            str(rng.choice([",", " ,"])) + sp() + t2 + str(rng.choice([";", " ;", "", ";x"])) +
            str(rng.choice(["\n", "\r\n", ""])))
    if rng.uniform() < 0.03:
        line = "#" + line
    return line.encode()


def test_fast_path_equals_sscanf_path(lib):
    rng = np.random.default_rng(0)
    taken = 0
    for _ in range(20000):
        line = _line(rng)
        rf, vf, cf = _parse(lib, line, 1)
        rs, vs, cs = _parse(lib, line, 0)
        if rf == -1:
            continue
        taken += 1
        assert rf == rs == 1, line
        assert vf == vs and cf == cs == 1, line
    assert taken > 1500  # the canonical form is exercised


def test_fast_path_takes_the_writer_format(lib):
    for line in (b"3, 4, 2000-01-01T00:00:00Z, 2000-01-01T00:30:00Z;\n",
                 b"  12,  7, 2023-02-29T25:61:61, 2023-03-01T00:00:00;\r\n",
                 b"1, 2, 1969-12-31T23:59:58Z, 1970-01-01T00:00:00Z;"):
        rf, vf, cf = _parse(lib, line, 1)
        rs, vs, cs = _parse(lib, line, 0)
        assert rf == 1 and (rf, vf, cf) == (rs, vs, cs), line
    # -1 is the reference's failure value: left to the slow path, which drops it
    line = b"1, 2, 1969-12-31T23:59:59Z, 1970-01-01T00:00:00Z;"
    assert _parse(lib, line, 1)[0] == -1
    assert _parse(lib, line, 0)[0] == 0


@pytest.mark.parametrize("name", FIXTURES)
def test_parse_file_matches_oracle(lib, oracle, name):
    path = os.path.join(GRAPHS, name)
    g = oracle.OracleGraph.from_file(path, build_ncls=False)
    iv = g.intervals()  # parse order: start, end, (min << 32) | max, sublist
    order = np.lexsort((-iv[:, 1], iv[:, 0]))  # stable: start asc, end desc, file order
    ref = iv[order]
    out = np.zeros((max(len(iv), 1), 4), np.int64)
    n = lib.rp_host_parse_file(path.encode(), out.ctypes.data_as(C.POINTER(C.c_int64)), len(out))
    assert n == len(iv)
    np.testing.assert_array_equal(out[:n, 0], ref[:, 0])
    np.testing.assert_array_equal(out[:n, 1], ref[:, 1])
    np.testing.assert_array_equal((out[:n, 2] << 32) | out[:n, 3], ref[:, 2])
