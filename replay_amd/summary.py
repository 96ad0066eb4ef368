"""Summary series in the layout of the reference's summary CSV and of
visualize_summary.py, without a round trip through a file.

The CSV (tool/gpu_cpu_temporal_sim.cu:505-568, header cu:707-709) holds one
row per (time block, replica): ``time,simulation_id,exposed,infectious,
resistant,susceptible``.  The first block is the initial state at the start
time; block k + 1 follows iteration k and carries its time
(``start + llround(k * step)``, cu:1080-1087), so two blocks share the start
time.  visualize_summary.py:8-41 reads that file with pandas, sorts it by
(simulation_id, time), measures time from the first block and plots each
replica's four curves on the largest time unit the span reaches.
"""
from __future__ import annotations

import numpy as np

COLUMNS = ["time", "simulation_id", "exposed", "infectious", "resistant", "susceptible"]

# visualize_summary.py:23-30
TIME_UNITS = [("seconds", 1), ("minutes", 60), ("hours", 3600), ("days", 86400),
              ("months", 30 * 86400), ("years", 365 * 86400)]


def block_times(sim) -> np.ndarray:
    """Time of each count block of a Simulation: the start time, then the
    time of iteration k for block k + 1."""
    n = sim.counts().shape[0]
    return np.array([sim.iteration_time(k - 1) for k in range(n)], np.int64)


def summary_frame(counts: np.ndarray, times: np.ndarray):
    """pandas DataFrame with the CSV's columns and row order (block by
    block, replicas in order) from counts (B, M, 4: exposed, infectious,
    resistant, susceptible) and the blocks' times (B)."""
    import pandas as pd
    counts = np.asarray(counts)
    B, M, _ = counts.shape
    flat = counts.reshape(B * M, 4)
    return pd.DataFrame({
        "time": np.repeat(np.asarray(times, np.int64), M),
        "simulation_id": np.tile(np.arange(M, dtype=np.int64), B),
        "exposed": flat[:, 0], "infectious": flat[:, 1],
        "resistant": flat[:, 2], "susceptible": flat[:, 3],
    }, columns=COLUMNS)


def read_summary_csv(path: str):
    """(times (B,), counts (B, M, 4)) from a summary CSV written by the
    `replay` executable, the reference or write_summary_csv.  Raises
    ValueError when the header differs or the rows are not whole blocks of
    replicas 0 .. M-1."""
    with open(path) as f:
        header = f.readline().strip()
        if header != ",".join(COLUMNS):
            raise ValueError(f"unexpected summary header: {header!r}")
        rows = np.loadtxt(f, delimiter=",", dtype=np.int64, ndmin=2)
    if rows.size == 0:
        return np.zeros(0, np.int64), np.zeros((0, 0, 4), np.int32)
    sim = rows[:, 1]
    M = int(sim.max()) + 1
    if rows.shape[0] % M or not np.array_equal(sim, np.tile(np.arange(M), rows.shape[0] // M)):
        raise ValueError("summary rows are not whole blocks of replicas 0 .. M-1")
    B = rows.shape[0] // M
    times = rows[::M, 0].copy()
    if not np.array_equal(rows[:, 0], np.repeat(times, M)):
        raise ValueError("a block's rows carry different times")
    return times, rows[:, 2:6].reshape(B, M, 4).astype(np.int32)


def time_unit(span_seconds: float):
    """(name, seconds) of the unit visualize_summary.py plots a span of
    span_seconds in: the largest unit the span reaches, else seconds
    (visualize_summary.py:32-36)."""
    for name, val in reversed(TIME_UNITS):
        if span_seconds / val >= 1:
            return name, val
    return "seconds", 1


def plotted_series(counts: np.ndarray, times: np.ndarray):
    """What visualize_summary.py draws for each replica: the scaled time axis
    and, per state (susceptible, exposed, infectious, resistant), an
    (M, B) array of curves; plus the unit name."""
    counts = np.asarray(counts)
    t = np.asarray(times, np.int64)
    delta = t - t.min()
    unit, scale = time_unit(float(delta.max()))
    order = np.argsort(delta, kind="stable")
    series = {
        "susceptible": counts[order, :, 3].T,
        "exposed": counts[order, :, 0].T,
        "infectious": counts[order, :, 1].T,
        "resistant": counts[order, :, 2].T,
    }
    return delta[order] / scale, series, unit
