"""Summary series helpers (CPU): the CSV the oracle executable writes reads
back to the oracle's counts and block times, in-memory frames have the
CSV's rows, and the plotted unit follows visualize_summary.py's rule."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, graph_path
from replay_amd import summary
from test_gpu_parity import _graph_start, _to_oracle_params
from replay_amd import Params


def test_read_oracle_csv(oracle, tmp_path):
    path = graph_path("random_nested.txt")
    start = _graph_start(path)
    p = Params(M=9, step_size=1200.0, static_network_duration=3600, iterations=30,
               initial_infected=0.2, infect_prob=0.4, exposed_duration=3600,
               infectious_duration=7200, resistant_duration=3600, start_time=start, seed=5)
    out = str(tmp_path / "s.csv")
    subprocess.run([os.path.join(ROOT, "oracle", "replay_cpu"), path, "--M", "9",
                    "--step-size", "1200", "--iterations", "30", "--time-step", "3600",
                    "--initial-infected", "0.2", "--infect-prob", "0.4",
                    "--exposed-duration", "3600", "--infectious-duration", "7200",
                    "--resistant-duration", "3600", "--start-time", str(start), "--seed", "5",
                    "-q", "--summary-out", out], check=True, capture_output=True)
    times, counts = summary.read_summary_csv(out)
    g = oracle.OracleGraph.from_file(path)
    ref, _, _ = g.simulate(_to_oracle_params(oracle, p), want_state=False)
    np.testing.assert_array_equal(counts, ref)
    # initial block and block 1 at the start time, then start + llround(k * step)
    want = np.array([start] + [start + int(round(k * 1200.0)) for k in range(30)], np.int64)
    np.testing.assert_array_equal(times, want)
    # the in-memory frame writes the same text
    df = summary.summary_frame(counts, times)
    text = ",".join(summary.COLUMNS) + "\n" + "".join(
        ",".join(str(int(v)) for v in row) + "\n" for row in df.itertuples(index=False))
    assert text == open(out).read()


def test_frame_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    counts = rng.integers(0, 1000, size=(6, 4, 4)).astype(np.int32)
    times = np.array([100, 100, 700, 1300, 1900, 2500], np.int64)
    df = summary.summary_frame(counts, times)
    assert list(df.columns) == summary.COLUMNS
    assert df.shape == (24, 6)
    f = tmp_path / "r.csv"
    df.to_csv(f, index=False)
    t2, c2 = summary.read_summary_csv(str(f))
    np.testing.assert_array_equal(t2, times)
    np.testing.assert_array_equal(c2, counts)


def test_read_rejects_ragged(tmp_path):
    f = tmp_path / "bad.csv"
    f.write_text(",".join(summary.COLUMNS) + "\n1,0,1,1,1,1\n1,1,1,1,1,1\n2,0,1,1,1,1\n")
    with pytest.raises(ValueError):
        summary.read_summary_csv(str(f))
    f.write_text("time,sim\n1,0\n")
    with pytest.raises(ValueError):
        summary.read_summary_csv(str(f))


@pytest.mark.parametrize("span,unit", [(0, "seconds"), (59, "seconds"), (60, "minutes"),
                                       (3599, "minutes"), (3600, "hours"), (86400, "days"),
                                       (30 * 86400 - 1, "days"), (30 * 86400, "months"),
                                       (365 * 86400, "years")])
def test_time_unit(span, unit):
    assert summary.time_unit(span)[0] == unit


def test_plotted_series_order():
    counts = np.arange(3 * 2 * 4, dtype=np.int32).reshape(3, 2, 4)
    times = np.array([50, 50, 3650], np.int64)
    t, series, unit = summary.plotted_series(counts, times)
    assert unit == "hours"
    np.testing.assert_allclose(t, [0.0, 0.0, 1.0])
    np.testing.assert_array_equal(series["susceptible"], counts[:, :, 3].T)
    np.testing.assert_array_equal(series["exposed"], counts[:, :, 0].T)
