"""The oracle's ingestion against the reference's own compiled ingestion.

tests/golden/ingest_golden.json.gz holds what tool/temporal_contact_matrix.c
and tool/ncls/ncls/src/intervaldb.c (compiled from the reference sources into
oracle/_ref by tests/golden/make_golden.py) return for the graphs under
tests/golden/graphs: node counts, intervals, NCLS order, the hit cap and the
find_intervals hits of many windows.  When oracle/_ref is built, more random
graphs are compared directly.
"""
import gzip
import json
import os

import numpy as np
import pytest

from conftest import GRAPHS, graph_path

GOLDEN = os.path.join(os.path.dirname(GRAPHS), "ingest_golden.json.gz")


@pytest.fixture(scope="module")
def golden():
    with gzip.open(GOLDEN, "rt") as f:
        return json.load(f)


def test_golden_covers_fixtures(golden):
    names = {n for n in os.listdir(GRAPHS) if n.endswith(".txt")}
    assert names == set(golden)
    # the cap truncates at least one query in most graphs: the vectors pin it
    assert sum(any(len(q["hits_capped"]) < len(q["hits_all"]) for q in g["queries"])
               for g in golden.values()) >= 5


@pytest.mark.parametrize("name", sorted(n for n in os.listdir(GRAPHS) if n.endswith(".txt")))
def test_parse_and_ncls_match_reference(oracle, golden, name):
    ref = golden[name]
    g = oracle.OracleGraph.from_file(graph_path(name), build_ncls=False)
    assert g.n_nodes == ref["n_nodes"]
    assert g.n_edges == ref["n_edges"]
    np.testing.assert_array_equal(g.intervals()[:, :3], np.array(ref["file_order"]))
    g.build_ncls()
    assert g.n_top == ref["n_top"]
    assert g.nlists == ref["nlists"]
    np.testing.assert_array_equal(g.intervals(), np.array(ref["ncls_order"]))
    for q in ref["queries"]:
        got = g.find(q["ws"], q["we"], g.nlists)
        np.testing.assert_array_equal(got[:, :3].reshape(-1, 3),
                                      np.array(q["hits_capped"]).reshape(-1, 3))
        got_all = g.find(q["ws"], q["we"], g.n_edges + 1)
        np.testing.assert_array_equal(got_all[:, :3].reshape(-1, 3),
                                      np.array(q["hits_all"]).reshape(-1, 3))


def _rank_order(iv_file: np.ndarray) -> np.ndarray:
    """Stable sort by (start asc, end desc): the claimed enumeration order."""
    idx = np.lexsort((-iv_file[:, 1], iv_file[:, 0]))
    return iv_file[idx]


@pytest.mark.parametrize("name", sorted(n for n in os.listdir(GRAPHS) if n.endswith(".txt")))
def test_hits_equal_rank_order_filter(golden, name):
    """For well-formed intervals the DFS hit order is the rank order: the
    property the HIP slices are built on, checked on the reference's vectors."""
    ref = golden[name]
    iv = np.array(ref["file_order"], dtype=np.int64)
    assert (iv[:, 1] >= iv[:, 0]).all()
    ranked = _rank_order(iv)
    cap = max(ref["nlists"], 1)
    for q in ref["queries"]:
        if q["ws"] < 0:
            continue  # mirrored-query path of idb:437-444; the HIP engine rejects it
        ov = ranked[(ranked[:, 0] < q["we"]) & (q["ws"] < ranked[:, 1])]
        np.testing.assert_array_equal(ov.reshape(-1, 3), np.array(q["hits_all"]).reshape(-1, 3))
        np.testing.assert_array_equal(ov[:cap].reshape(-1, 3),
                                      np.array(q["hits_capped"]).reshape(-1, 3))


def _random_graph_file(path, seed, n_nodes, n_contacts, span):
    rng = np.random.default_rng(seed)
    import time
    lines = ["NODE_LIST " + " ".join(str(i * 7 + 3) for i in range(n_nodes))]
    for _ in range(n_contacts):
        a, b = rng.choice(n_nodes, size=2, replace=False)
        s = 946713600 + int(rng.integers(0, span))
        kind = rng.uniform()
        d = int(rng.integers(span // 3, span)) if kind < 0.25 else (0 if kind < 0.3 else int(rng.integers(1, 7200)))
        f = "%Y-%m-%dT%H:%M:%SZ"
        lines.append(f"{a * 7 + 3}, {b * 7 + 3}, {time.strftime(f, time.gmtime(s))}, "
                     f"{time.strftime(f, time.gmtime(s + d))};")
        if rng.uniform() < 0.05:
            lines.append(lines[-1])
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


@pytest.mark.parametrize("seed", range(6))
def test_random_graphs_against_reference_build(oracle, tmp_path, seed):
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    p = str(tmp_path / f"g{seed}.txt")
    _random_graph_file(p, seed, 8 + 4 * seed, 50 + 60 * seed, 86400)
    ref = oracle.RefIngest(p)
    g = oracle.OracleGraph.from_file(p)
    assert (g.n_nodes, g.n_edges, g.n_top, g.nlists) == (ref.n_nodes, ref.n_edges, ref.n_top, ref.nlists)
    np.testing.assert_array_equal(g.intervals(), ref.ncls_order)
    for ws in range(946713600 - 3600, 946713600 + 86400 + 3600, 1800):
        for d in (600, 3600, 20000):
            np.testing.assert_array_equal(g.find(ws, ws + d, g.nlists), ref.find(ws, ws + d, ref.nlists))


def test_iso8601_against_reference(oracle):
    cases = ["2000-01-01T05:24:44Z", "2000-01-01T25:59:59Z", "1970-01-01T00:00:00Z",
             "2001-02-30T00:00:00Z", "2000-13-01T00:00:00", "1999-12-31T23:59:60Z",
             "2000-01-01T00:00", "garbage", "1969-12-31T23:59:59Z", "2038-01-19T03:14:08Z",
             "2000-00-00T-1:-1:-1Z", " 2000-1-1T1:1:1Zextra"]
    expect_fixed = {"2000-01-01T05:24:44Z": 946704284, "1970-01-01T00:00:00Z": 0,
                    "2000-01-01T00:00": -1, "garbage": -1}
    for c in cases:
        got = oracle.lib().or_parse_iso8601(c.encode())
        if c in expect_fixed:
            assert got == expect_fixed[c], c
        if oracle.ref_available():
            assert got == oracle.ref_lib().parse_iso8601_to_unix(c.encode()), c
