"""The group-structured input generator (CPU) against the reference's own
generator (generate_networks/graph_generator.c:39-157): byte-identical files
for the same arguments -- against goldens written by the compiled reference
generator and, where the reference sources are built here (oracle/_ref),
against fresh runs -- and the rand() stream it reproduces against the
system C library's rand()."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import replay_amd
from conftest import ROOT

GOLD = os.path.join(ROOT, "tests", "golden", "generator")
REF_GEN = os.path.join(ROOT, "oracle", "_ref", "graph_gen")
T2000 = 946684800


def _ours(tmp_path, name, *args):
    out = str(tmp_path / name)
    n = replay_amd.generate_groups(out, *args)
    return out, n


@pytest.mark.parametrize("gold,args", [
    ("groups_40n_4g_5s_seed12.txt", (T2000, 5, 60, 40, 4, 0.3, 0.05, 12)),
    ("groups_25n_1g_3s_seed0.txt", (1625097540, 3, 7, 25, 1, 0.5, 0.5, 0)),
])
def test_matches_golden(tmp_path, gold, args):
    out, n = _ours(tmp_path, "g.txt", *args)
    ref = open(os.path.join(GOLD, gold), "rb").read()
    assert open(out, "rb").read() == ref
    assert n == ref.count(b";\n")


@pytest.mark.skipif(not os.path.exists(REF_GEN), reason="reference generator not built here")
@pytest.mark.parametrize("args", [
    ("2000-01-01T00:00:00", 3, 600, 60, 6, 0.4, 0.02, 7),
    ("1999-12-31T23:30:00", 2, 1, 30, 30, 1.0, 0.0, 1),
    ("2010-02-28T12:00:00", 4, 3600, 35, 3, 0.15, 0.15, 123456789),
    ("2000-13-40T25:61:61", 1, 17, 20, 2, 0.9, 0.1, 5),   # fields normalised as mktime does
])
def test_matches_reference_generator(tmp_path, args):
    start, steps, ss, pop, grp, pi, pe, seed = args
    ref = str(tmp_path / "ref.txt")
    ours = str(tmp_path / "ours.txt")
    env = dict(os.environ, TZ="UTC")
    cli = [str(x) for x in (start, steps, ss, pop, grp, pi, pe)]
    r1 = subprocess.run([REF_GEN, *cli, ref, str(seed)], env=env, capture_output=True, text=True)
    r2 = subprocess.run([replay_amd.GEN_PATH, *cli, ours, str(seed)], capture_output=True,
                        text=True)
    assert r1.returncode == r2.returncode == 0
    assert open(ours, "rb").read() == open(ref, "rb").read()
    # the progress lines are the reference's (the last line names the file)
    assert r1.stdout.splitlines()[:-1] == r2.stdout.splitlines()[:-1]


@pytest.mark.parametrize("seed", [0, 1, 2, 12345, 2147483647])
def test_rand_stream_matches_libc(tmp_path, seed):
    """A pair in one group with probability 1 and a step of 2^31 - 1 seconds
    writes two raw draws per step as times: they must be the C library's
    rand() values 2, 3, 5, 6, ... after srand(seed) (draws 1, 4, ... are the
    contact tests)."""
    libc = C.CDLL("libc.so.6")
    libc.rand.restype = C.c_int
    libc.srand(C.c_uint(seed))
    steps, ss = 3, 2147483647
    draws = [libc.rand() for _ in range(3 * steps)]
    out, n = _ours(tmp_path, "r.txt", 0, steps, ss, 2, 1, 1.0, 1.0, seed)
    assert n == steps
    import calendar
    import time
    lines = open(out).read().splitlines()[1:]
    for k, line in enumerate(lines):
        parts = [x.strip() for x in line.rstrip(";").split(",")]
        t = sorted(calendar.timegm(time.strptime(p, "%Y-%m-%dT%H:%M:%SZ")) - k * ss
                   for p in parts[2:4])
        assert t == sorted([draws[3 * k + 1] % ss, draws[3 * k + 2] % ss])


def test_generated_file_parses(tmp_path, oracle):
    out, n = _ours(tmp_path, "p.txt", T2000, 4, 900, 50, 5, 0.2, 0.02, 3)
    g = oracle.OracleGraph.from_file(out, build_ncls=False)
    assert len(g.intervals()) == n


def test_generator_errors(tmp_path):
    with pytest.raises(replay_amd.ReplayError):
        replay_amd.generate_groups(str(tmp_path / "x.txt"), T2000, 1, 0, 10, 2, 0.5, 0.5, 1)
    with pytest.raises(replay_amd.ReplayError):
        replay_amd.generate_groups(str(tmp_path / "x.txt"), T2000, 1, 60, 10, 2, 0.5, 0.5, -1)
    with pytest.raises(replay_amd.ReplayError):
        replay_amd.generate_groups(str(tmp_path / "no" / "x.txt"), T2000, 1, 60, 10, 2, 0.5,
                                   0.5, 1)
