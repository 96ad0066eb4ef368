"""bench.py's contract pieces that need no GPU: the roofline arithmetic
(algorithmic bytes by the fused SpMM definition, the design's streamed
bytes, the VALU roof), the committed traffic summary it quotes, and the
one-GPU refusal."""
import json
import os
import subprocess
import sys

import pytest

from conftest import ROOT

sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_roofline_arithmetic():
    # 10 epochs of 1000 entries streamed at 8 bytes; 5,000 state bytes
    info = {"step_bytes": 8 * 1000 + 5000, "n_entries": 10_000, "n_epochs": 10,
            "entry_bytes": 8, "M": 256}
    r = bench.roofline(info, 1e-6, "no_such_workload", nnz=700.0, nnz_note="test")
    assert r["bytes_per_launch"] == 8 * 700 + 5000
    assert r["achieved"] == pytest.approx((8 * 700 + 5000) / 1e-6 / 1e9)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0)
    assert r["design_bytes_per_launch"] == 13000
    assert r["traffic"] is None
    assert r["valu_roof"]["ops_per_launch"] == 700 * 256
    assert r["valu_roof"]["frac"] == pytest.approx(700 * 256 / 1e-6 / 1e12 / 39.3)
    # the 16-byte stream prices the same state terms
    wide = dict(info, step_bytes=16 * 1000 + 5000, entry_bytes=16)
    assert bench.roofline(wide, 1e-6, "x", 700.0, "t")["bytes_per_launch"] == 8 * 700 + 5000


def test_committed_traffic_summary():
    t = bench.measured_traffic("config3_batched_64nets")
    assert t is not None
    # every kernel of the step and the once-per-run tally, per step
    parts = sum(v for k, v in t.items() if k.endswith(("_fetch_bytes", "_write_bytes")))
    assert {"expo_fetch_bytes", "expo_write_bytes", "update_fetch_bytes",
            "update_write_bytes"} <= set(t)
    assert abs(parts - t["hbm_bytes_per_step"]) < 1
    assert "FETCH_SIZE" in t["source"] and "WRITE_SIZE" in t["source"]


def test_more_than_one_gpu_exits():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""
