"""Writes tests/golden/ingest_golden.json.gz from the reference's own ingestion.

Runs tool/temporal_contact_matrix.c + tool/ncls/ncls/src/intervaldb.c from the
reference, compiled into oracle/_ref/libref_ingest.so by `make -C oracle ref`,
on every graph under tests/golden/graphs/ and records:
  - NODE_LIST count and valid contact count (tcm:148-321),
  - the intervals in file order (start, end, packed pair),
  - the NCLS order, top-level length and the modified sublist count
    (idb:227-326, the query cap used at cu:988-989),
  - the hits find_intervals returns for a set of windows, with the step
    loop's cap (nbuf = nlists) and with an unbounded buffer.

The random graphs are (re)generated here with a fixed seed first.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import gzip
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle_py  # noqa: E402

GRAPHS = os.path.join(HERE, "graphs")
T0 = 946713600


def _iso(t: int) -> str:
    import time
    return time.strftime("%Y-%m-%dT%H:%M:%SZ", time.gmtime(t))


def write_random_graph(path: str, seed: int, n_nodes: int, n_contacts: int, span: int,
                       nested: bool) -> None:
    rng = np.random.default_rng(seed)
    ids = rng.choice(10 * n_nodes, size=n_nodes, replace=False)
    lines = ["NODE_LIST " + " ".join(str(int(i)) for i in ids)]
    for _ in range(n_contacts):
        a, b = rng.choice(n_nodes, size=2, replace=False)
        s = T0 + int(rng.integers(0, span))
        if nested and rng.uniform() < 0.3:
            d = int(rng.integers(span // 4, span))       # long container
        elif rng.uniform() < 0.1:
            d = 0                                          # zero length
        else:
            d = int(rng.integers(1, 3 * 3600))
        if rng.uniform() < 0.1 and len(lines) > 1:        # exact duplicate
            lines.append(lines[-1])
            continue
        lines.append(f"{ids[a]}, {ids[b]}, {_iso(s)}, {_iso(s + d)};")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


RANDOM = [("random_flat.txt", 1, 12, 60, 2 * 86400, False),
          ("random_nested.txt", 2, 15, 120, 86400, True),
          ("random_dense.txt", 3, 30, 400, 3 * 86400, True)]


def windows_for(iv: np.ndarray) -> list[tuple[int, int]]:
    lo, hi = int(iv[:, 0].min()), int(iv[:, 1].max())
    out = [(lo - 10, hi + 10), (lo, lo + 1), (hi - 1, hi + 5)]
    for d in (600, 3600, 7200):
        t0 = lo - (lo % d) - d
        n = (hi + d - t0) // d + 1
        stride = max(1, n // 120)          # at most ~120 windows per size
        for k in range(0, n, stride):
            out.append((t0 + k * d, t0 + (k + 1) * d))
    return out


def main() -> None:
    if not oracle_py.ref_available():
        raise SystemExit("build oracle/_ref first: make -C oracle ref")
    for name, seed, n, e, span, nested in RANDOM:
        write_random_graph(os.path.join(GRAPHS, name), seed, n, e, span, nested)
    out = {}
    for name in sorted(os.listdir(GRAPHS)):
        if not name.endswith(".txt"):
            continue
        ref = oracle_py.RefIngest(os.path.join(GRAPHS, name))
        wins = windows_for(ref.file_order)
        cap = ref.nlists
        out[name] = {
            "n_nodes": ref.n_nodes,
            "n_edges": ref.n_edges,
            "file_order": ref.file_order[:, :3].tolist(),
            "ncls_order": ref.ncls_order.tolist(),
            "n_top": ref.n_top,
            "nlists": ref.nlists,
            "queries": [{"ws": ws, "we": we,
                         "hits_capped": ref.find(ws, we, cap)[:, :3].tolist(),
                         "hits_all": ref.find(ws, we, ref.n_edges + 1)[:, :3].tolist()}
                        for ws, we in wins],
        }
    with gzip.open(os.path.join(HERE, "ingest_golden.json.gz"), "wt") as f:
        json.dump(out, f, separators=(",", ":"))
    print(f"wrote {len(out)} graphs")


if __name__ == "__main__":
    main()
