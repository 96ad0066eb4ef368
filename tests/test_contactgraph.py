"""The command-file generator (CPU): the native `replay_contactgraph` /
`replay_cgg` against the oracle restatement of
contactgraphgen/contactgraphgen.py, and the oracle pinned against the
reference's own example (example.txt -> graph.txt) and hand-derived
expectations."""
import calendar
import os
import random
import subprocess
import time

import pytest

import replay_amd
from conftest import ROOT

import sys
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import contactgraph_ref as ref  # noqa: E402

GOLD = os.path.join(ROOT, "tests", "golden")
EXAMPLE = os.path.join(GOLD, "contactgraph", "example_commands.txt")
EXAMPLE_GRAPH = os.path.join(GOLD, "graphs", "contactgraphgen_graph.txt")


def native(tmp_path, data: bytes, name="c.txt"):
    """(stderr lines, graph text or None, exit status) of replay_cgg."""
    cmd = tmp_path / name
    cmd.write_bytes(data)
    out = tmp_path / (name + ".graph")
    if out.exists():
        out.unlink()
    r = subprocess.run([replay_amd.CGG_PATH, str(cmd), "--out", str(out)], capture_output=True)
    graph = out.read_bytes().decode() if out.exists() else None
    return r.stderr.decode().splitlines(), graph, r.returncode


def _iso(t):
    return time.strftime("%Y-%m-%dT%H:%M:%SZ", time.gmtime(t))


# ---- the oracle, pinned ----------------------------------------------------

def test_oracle_reference_example():
    e = ref.expand(open(EXAMPLE, "rb").read())
    assert e.messages == [] and not e.failed and e.contacts == 10
    assert e.graph == open(EXAMPLE_GRAPH).read()


def test_oracle_hand_derived_cases():
    # a triangle twice, a duplicate of one pair, DELTA 0 overriding the
    # horizon, UNTIL equal to START (no start), ids declared out of order
    e = ref.expand(b"ADDNODE 7\nADDNODE 3\nADDNODE 12\n"
                   b"ADDEDGE [12,7,3,7] 30 10 20 OCCURRENCES 2\n"
                   b"ADDEDGE [3,7] 50 10 0 OCCURRENCES 9\n"
                   b"ADDEDGE [3,12] 40 5 1 UNTIL 40\n")
    assert e.messages == [] and not e.failed
    assert e.graph == ("NODE_LIST 3 7 12\n"
                       "3, 7, 1970-01-01T00:00:30Z, 1970-01-01T00:00:40Z; \n"
                       "3, 12, 1970-01-01T00:00:30Z, 1970-01-01T00:00:40Z; \n"
                       "7, 12, 1970-01-01T00:00:30Z, 1970-01-01T00:00:40Z; \n"
                       "3, 7, 1970-01-01T00:00:50Z, 1970-01-01T00:01:00Z; \n"
                       "3, 12, 1970-01-01T00:00:50Z, 1970-01-01T00:01:00Z; \n"
                       "7, 12, 1970-01-01T00:00:50Z, 1970-01-01T00:01:00Z; \n")
    # an edge before its nodes exist, and one id missing of three
    e = ref.expand(b"ADDEDGE [1,2] 0 1 1\nADDNODE 1\nADDNODE 2\nADDEDGE [2,1,5] 0 1 1\n")
    assert e.messages == [
        "ERROR: ID 1 is undefined. Add it with `ADDNODE 1`.",
        "ERROR: ID 2 is undefined. Add it with `ADDNODE 2`.",
        "Skipping ADDEDGE on line 1.",
        "ERROR: ID 5 is undefined. Add it with `ADDNODE 5`.",
        "Skipping ADDEDGE on line 4."]
    assert e.graph == "NODE_LIST 1 2\n" and e.contacts == 0
    # no nodes at all: the header has no trailing space
    assert ref.expand(b"# nothing\n\n").graph == "NODE_LIST\n"


def test_oracle_counts_closed_form():
    """C(m, 2) pairs at each start; UNTIL T gives ceil((T - S) / D) starts."""
    rng = random.Random(5)
    for _ in range(40):
        m = rng.randint(2, 6)
        S, D, dur = rng.randint(0, 10**6), rng.randint(1, 5000), rng.randint(0, 9000)
        ids = rng.sample(range(1, 10**9), m)
        cmds = "".join(f"ADDNODE {i}\n" for i in ids)
        if rng.random() < 0.5:
            n = rng.randint(0, 12)
            cmds += f"ADDEDGE [{','.join(map(str, ids))}] {S} {dur} {D} OCCURRENCES {n}\n"
        else:
            T = S + rng.randint(-3 * D, 12 * D)
            cmds += f"ADDEDGE [{','.join(map(str, ids))}] {S} {dur} {D} UNTIL {_iso(T)}\n"
            n = max(0, -(-(T - S) // D))
        e = ref.expand(cmds.encode())
        assert e.contacts == m * (m - 1) // 2 * n
        rows = [ln.rstrip("; ").split(", ") for ln in e.graph.splitlines()[1:]]
        keys = [(calendar.timegm(time.strptime(a, "%Y-%m-%dT%H:%M:%SZ")), int(u), int(v),
                 calendar.timegm(time.strptime(b, "%Y-%m-%dT%H:%M:%SZ")))
                for u, v, a, b in rows]
        assert keys == sorted(keys) and all(k[3] - k[0] == dur and k[1] < k[2] for k in keys)
        assert sorted({k[0] for k in keys}) == [S + k * D for k in range(n)][:len(set(
            k[0] for k in keys))]


def test_oracle_stops_past_year_9999():
    last = 253402300799
    e = ref.expand(f"ADDNODE 1\nADDNODE 2\nADDEDGE [1,2] {last - 2} 1 1 OCCURRENCES 9\n".encode())
    assert e.failed and e.contacts == 2
    assert e.graph.splitlines()[-1] == "1, 2, 9999-12-31T23:59:58Z, 9999-12-31T23:59:59Z; "


# ---- the native generator against the oracle -------------------------------

def test_reference_example(tmp_path):
    msgs, graph, rc = native(tmp_path, open(EXAMPLE, "rb").read())
    assert (msgs, rc) == ([], 0)
    assert graph == open(EXAMPLE_GRAPH).read()
    n = replay_amd.contact_graph(EXAMPLE, str(tmp_path / "g.txt"), quiet=True)
    assert n == 10 and (tmp_path / "g.txt").read_text() == graph


EDGE_CASES = {
    "lexical": (b"  ADDNODE 1\t\r\nADDNODE\x0b2\rADDNODE 03\n\n# c\n ADDNODE\xc2\xa04 \n"
                b"ADDNODE\xe3\x80\x805\nADDNODE 6 7\nADDNODE\nADDNODE x\nADDNODE 000\n"
                b"ADDNODE5\nADDNODEX 8\naddnode 9\n\xef\xbb\xbfADDNODE 10\n"
                b"ADDEDGE[1,2]0 1 1\nADDEDGE [1,2]\t5 1 1 OCCURRENCES\x1c2\n"),
    "id lists": (b"ADDNODE 1\nADDNODE 2\nADDNODE 3\n"
                 b"ADDEDGE\nADDEDGE 1,2\nADDEDGE [1,2\nADDEDGE [1,[2]] 0 1 1\n"
                 b"ADDEDGE [1,,2] 0 1 1\nADDEDGE [ ] 0 1 1\nADDEDGE [1,1] 0 1 1\n"
                 b"ADDEDGE [1, 0] 0 1 1\nADDEDGE [1,00] 0 1 1\nADDEDGE [1,-2] 0 1 1\n"
                 b"ADDEDGE [1,2]] 0 1 1\nADDEDGE [ 3 , 1 ,2,2 ] 0 1 1\nADDEDGE [1,2]\n"
                 b"ADDEDGE [1,2] 0 1\nADDEDGE [1,2] a b c\nADDEDGE [,] 0 1 1\n"),
    "horizons": (b"ADDNODE 1\nADDNODE 2\n"
                 b"ADDEDGE [1,2] 0 1 1 OCCURRENCES 2 UNTIL 5\n"
                 b"ADDEDGE [1,2] 0 1 1 UNTIL 5 OCCURRENCES\n"
                 b"ADDEDGE [1,2] 0 1 1 OCCURRENCES\nADDEDGE [1,2] 0 1 1 OCCURRENCES 2 3\n"
                 b"ADDEDGE [1,2] 0 1 1 OCCURRENCES -2\nADDEDGE [1,2] 0 1 1 UNTIL\n"
                 b"ADDEDGE [1,2] 0 1 1 UNTIL -5\nADDEDGE [1,2] 0 1 1 FOREVER 3\n"
                 b"ADDEDGE [1,2] 10 1 3 OCCURRENCES 0\nADDEDGE [1,2] 10 1 0 OCCURRENCES 0\n"
                 b"ADDEDGE [1,2] 20 2 3 UNTIL 29\nADDEDGE [1,2] 20 2 3 UNTIL 30\n"
                 b"ADDEDGE [1,2] 40 2 0 UNTIL 1\nADDEDGE [1,2] 50 2 7\n"
                 b"ADDEDGE [1,2] 60 0 5 OCCURRENCES 3\nADDEDGE [1,2] 60 0 5 OCCURRENCES 3\n"),
    "iso": (b"ADDNODE 1\nADDNODE 2\n"
            b"ADDEDGE [1,2] 0 1 60 UNTIL 1970-1-1T0:3:0Z\n"
            b"ADDEDGE [1,2] 1000 1 60 UNTIL 1970-01-01t00:20:00z\n"
            b"ADDEDGE [1,2] 0 1 1 UNTIL 1970-01-01T00:00:60Z\n"
            b"ADDEDGE [1,2] 0 1 1 UNTIL 1970-02-29T00:00:00Z\n"
            b"ADDEDGE [1,2] 0 1 1 UNTIL 1972-02-30T00:00:00Z\n"
            b"ADDEDGE [1,2] 0 1 86400 UNTIL 1972-02-29T00:00:01Z\n"
            b"ADDEDGE [1,2] 0 1 1 UNTIL 0000-01-01T00:00:00Z\n"
            b"ADDEDGE [1,2] 0 1 1 UNTIL 0001-01-01T00:00:00Z\n"
            b"ADDEDGE [1,2] 0 1 1 UNTIL 1970-00-01T00:00:00Z\n"
            b"ADDEDGE [1,2] 0 1 1 UNTIL 1970-01-01T24:00:00Z\n"
            b"ADDEDGE [1,2] 0 1 1 UNTIL 1970-01-01T00:00:00\n"
            b"ADDEDGE [1,2] 0 1 1 UNTIL 1970-001-01T00:00:00Z\n"
            b"ADDEDGE [1,2] 0 1 1 UNTIL 11970-01-01T00:00:00Z\n"
            b"ADDEDGE [1,2] 0 1 1 UNTIL 1970-01-01T00:00:00Zx\n"
            b"ADDEDGE [1,2] 0 1 1 UNTIL 1970-01-01 00:00:00Z\n"
            b"ADDEDGE [1,2] 7200 1 1800 UNTIL 1970-01-01T03:00:00Z\n"),
    # (a horizon of many starts past year 9999 is left out: the reference
    # expands every start before it writes, which does not finish)
    "big numbers": (b"ADDNODE 123456789012345678901234567890\nADDNODE 0000000000000000000000000000000000000000001\n"
                    b"ADDEDGE [1, 123456789012345678901234567890] 99999999999999999999 1 1\n"
                    b"ADDEDGE [1,123456789012345678901234567890] 0 1 99999999999999999999 "
                    b"OCCURRENCES 3\n"
                    b"ADDEDGE [1,123456789012345678901234567890] 5 0 99999999999999999999 "
                    b"UNTIL 99999999999999999999\n"),
    "past 9999": (b"ADDNODE 1\nADDNODE 2\nADDNODE 3\n"
                  b"ADDEDGE [1,2,3] 253402300790 5 4 OCCURRENCES 5\n"
                  b"ADDEDGE [1,2] 0 10 0\n"),
    "end past 9999": (b"ADDNODE 1\nADDNODE 2\nADDNODE 3\n"
                      b"ADDEDGE [2,3] 253402300700 10 0\n"
                      b"ADDEDGE [1,3] 253402300700 100 0\n"
                      b"ADDEDGE [1,2] 253402300700 5 0\n"),
    "empty": b"",
    "no newline at end": b"ADDNODE 1\nADDNODE 2\nADDEDGE [1,2] 5 5 5",
}


@pytest.mark.parametrize("name", sorted(EDGE_CASES))
def test_edge_cases_match_oracle(tmp_path, name):
    data = EDGE_CASES[name]
    e = ref.expand(data)
    msgs, graph, rc = native(tmp_path, data)
    assert rc == (1 if e.failed else 0)
    assert graph == e.graph
    # a write that stops past year 9999 ends with a traceback there and a
    # message of the native generator's own here
    assert (msgs[:-1] if e.failed else msgs) == e.messages


_WS = [" ", "  ", "\t", "\x0b", "\x0c", "\x1f", " ", " ", "　"]


def _random_commands(rng: random.Random) -> bytes:
    ids = [rng.choice([rng.randint(1, 40), rng.randint(1, 10**12)]) for _ in range(8)]
    lines = []
    for _ in range(rng.randint(1, 30)):
        ws = lambda: rng.choice(_WS) if rng.random() < 0.2 else " "  # noqa: E731
        r = rng.random()
        if r < 0.3:
            tok = rng.choice([str(rng.choice(ids)), "0", "x", "-3", "0" + str(rng.choice(ids)),
                              str(rng.choice(ids)) + " 4", ""])
            lines.append(f"ADDNODE{ws()}{tok}")
        elif r < 0.85:
            m = rng.randint(0, 5)
            lst = [str(rng.choice(ids + [0, 41])) for _ in range(m)]
            if rng.random() < 0.05:
                lst.append(rng.choice(["", "a", "[3]"]))
            sep = rng.choice([",", ", ", " ,"])
            S = rng.choice([rng.randint(0, 10**5), 253402300700 + rng.randint(0, 200)])
            nums = [str(S), str(rng.randint(0, 600)), str(rng.choice([0, 1, 7, 60, 3600]))]
            if rng.random() < 0.05:
                nums[rng.randrange(3)] = rng.choice(["q", "-1", "1.5"])
            if rng.random() < 0.1:
                nums = nums[:rng.randint(0, 2)]
            tail = ""
            h = rng.random()
            if h < 0.3:
                tail = f" OCCURRENCES{ws()}{rng.randint(0, 6)}"
            elif h < 0.6:
                T = S + rng.randint(-100, 4000)
                tail = " UNTIL " + rng.choice([str(max(T, 0)), _iso(max(T, 0)),
                                               _iso(max(T, 0)).lower(), "1970-13-01T00:00:00Z"])
            elif h < 0.65:
                tail = rng.choice([" UNTIL 5 OCCURRENCES 2", " OCCURRENCES", " BOGUS 3",
                                   " OCCURRENCES 2 2"])
            lines.append(f"ADDEDGE{ws()}[{sep.join(lst)}]{ws()}{' '.join(nums)}{tail}")
        elif r < 0.9:
            lines.append("# " + "comment")
        elif r < 0.95:
            lines.append(rng.choice(["", "   ", "FOO 1", "addedge [1,2] 0 1 1"]))
        else:
            lines.append(f"ADDEDGE [{rng.choice(ids)},{rng.choice(ids)}")
    eols = [rng.choice(["\n", "\r\n", "\r"]) if rng.random() < 0.2 else "\n" for _ in lines]
    text = "".join(a + b for a, b in zip(lines, eols))
    if rng.random() < 0.2:
        text = text.rstrip("\r\n")
    return text.encode()


def test_random_command_files_match_oracle(tmp_path):
    rng = random.Random(2024)
    for case in range(300):
        data = _random_commands(rng)
        e = ref.expand(data)
        msgs, graph, rc = native(tmp_path, data)
        ctx = f"case {case}: {data!r}"
        assert rc == (1 if e.failed else 0), ctx
        assert graph == e.graph, ctx
        assert (msgs[:-1] if e.failed else msgs) == e.messages, ctx


def test_large_expansion_matches_oracle(tmp_path):
    """Overlapping recurring cliques: many duplicates and interleaved starts."""
    rng = random.Random(9)
    lines = [f"ADDNODE {i}" for i in rng.sample(range(1, 5000), 60)]
    ids = [int(x.split()[1]) for x in lines]
    for _ in range(40):
        grp = rng.sample(ids, rng.randint(2, 12))
        lines.append(f"ADDEDGE [{','.join(map(str, grp))}] {rng.randint(0, 3600)} "
                     f"{rng.choice([60, 300, 900])} {rng.choice([300, 600, 3600])} "
                     f"OCCURRENCES {rng.randint(1, 30)}")
    data = ("\n".join(lines) + "\n").encode()
    e = ref.expand(data)
    msgs, graph, rc = native(tmp_path, data)
    assert (msgs, rc) == ([], 0)
    assert e.contacts > 10000 and graph == e.graph


def test_io_errors(tmp_path):
    r = subprocess.run([replay_amd.CGG_PATH, str(tmp_path / "none.txt"), "--out",
                        str(tmp_path / "g.txt")], capture_output=True, text=True)
    assert r.returncode == 1
    assert r.stderr == f"ERROR: Command file not found: {tmp_path / 'none.txt'}\n"
    (tmp_path / "c.txt").write_text("ADDNODE 1\n")
    bad = tmp_path / "no" / "g.txt"
    r = subprocess.run([replay_amd.CGG_PATH, str(tmp_path / "c.txt"), "--out", str(bad)],
                       capture_output=True, text=True)
    assert r.returncode == 1
    assert r.stderr == (f"ERROR: Failed to write output '{bad}': [Errno 2] No such file or "
                        f"directory: '{bad}'\n")
    (tmp_path / "u.txt").write_bytes(b"ADDNODE 1\n\xff\n")
    r = subprocess.run([replay_amd.CGG_PATH, str(tmp_path / "u.txt"), "--out",
                        str(tmp_path / "g2.txt")], capture_output=True, text=True)
    assert r.returncode == 1 and not (tmp_path / "g2.txt").exists()
    assert ref.expand(b"ADDNODE 1\n\xff\n").failed
    with pytest.raises(replay_amd.ReplayError):
        replay_amd.contact_graph(str(tmp_path / "none.txt"), str(tmp_path / "g.txt"), quiet=True)


def test_cli_arguments(tmp_path):
    (tmp_path / "c.txt").write_text("ADDNODE 1\nADDNODE 2\nADDEDGE [1,2] 0 1 1\n")
    exe = replay_amd.CGG_PATH
    for args in (["--out", "a.txt", "c.txt"], ["c.txt", "--out=b.txt"], ["c.txt", "--ou", "d.txt"]):
        r = subprocess.run([exe, *args], cwd=tmp_path, capture_output=True)
        assert r.returncode == 0
    assert {(tmp_path / f).read_text() for f in ("a.txt", "b.txt", "d.txt")} == {
        "NODE_LIST 1 2\n1, 2, 1970-01-01T00:00:00Z, 1970-01-01T00:00:01Z; \n"}
    for args in ([], ["c.txt"], ["--out", "x.txt"], ["c.txt", "c.txt", "--out", "y.txt"],
                 ["c.txt", "--bogus", "--out", "z.txt"]):
        r = subprocess.run([exe, *args], cwd=tmp_path, capture_output=True, text=True)
        assert r.returncode == 2 and r.stderr.startswith("usage:")
    r = subprocess.run([exe, "-h"], capture_output=True, text=True)
    assert r.returncode == 0 and "--out OUT" in r.stdout
