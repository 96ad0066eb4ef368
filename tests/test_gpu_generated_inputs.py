"""Generated inputs end to end on the GPU: a file from the group generator
(`replay_gen` / generate_groups) and one expanded from an ADDNODE / ADDEDGE
command file (`replay_cgg` / contact_graph) each run through the `replay`
CLI, whose summary and node-state CSVs must equal the oracle CLI's byte for
byte."""
import os
import subprocess

import pytest

import replay_amd
from conftest import ROOT

pytestmark = pytest.mark.gpu

T2000 = 946684800  # 2000-01-01T00:00:00Z


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return r


def _cli_matches_oracle(tmp_path, graph, args):
    ours = [str(tmp_path / "s1.csv"), str(tmp_path / "n1.csv")]
    ref = [str(tmp_path / "s2.csv"), str(tmp_path / "n2.csv")]
    _run([replay_amd.CLI_PATH, graph, *args, "-q", "--summary-out", ours[0],
          "--node-state-out", ours[1]])
    _run([os.path.join(ROOT, "oracle", "replay_cpu"), graph, *args, "-q", "--summary-out", ref[0],
          "--node-state-out", ref[1]])
    for x, y in zip(ours, ref):
        a, b = open(x).read(), open(y).read()
        assert a == b and a.count("\n") > 10


@pytest.mark.parametrize("M", [33, 256])
def test_group_generator_file(tmp_path, M):
    g = str(tmp_path / "groups.txt")
    n = replay_amd.generate_groups(g, T2000, 8, 3600, 80, 6, 0.25, 0.01, seed=4)
    assert n > 500
    args = ["--M", str(M), "--step-size", "1800", "--iterations", "16", "--time-step", "3600",
            "--initial-infected", "0.1", "--infect-prob", "0.5", "--exposed-duration", "1800",
            "--infectious-duration", "7200", "--resistant-duration", "3600",
            "--start-time", str(T2000), "--seed", "21"]
    _cli_matches_oracle(tmp_path, g, args)


def test_command_file_expansion(tmp_path):
    lines = [f"ADDNODE {i}" for i in range(1, 41)]
    # households of 4 meeting nightly, classrooms of 10 on weekday mornings,
    # and a few one-off gatherings that overlap them
    for h in range(10):
        ids = ",".join(str(4 * h + k) for k in range(1, 5))
        lines.append(f"ADDEDGE [{ids}] {T2000 + 18 * 3600} 36000 86400 OCCURRENCES 4")
    for c in range(4):
        ids = ",".join(str(10 * c + k) for k in range(1, 11))
        lines.append(f"ADDEDGE [{ids}] {T2000 + 8 * 3600} 14400 86400 UNTIL 2000-01-04T00:00:00Z")
    lines.append(f"ADDEDGE [3,17,25,38] {T2000 + 40000} 5400 0")
    lines.append(f"ADDEDGE [1,40] {T2000 + 100000} 1200 7200 OCCURRENCES 6")
    cmd = tmp_path / "commands.txt"
    cmd.write_text("\n".join(lines) + "\n")
    g = str(tmp_path / "graph.txt")
    n = replay_amd.contact_graph(str(cmd), g, quiet=True)
    assert n > 300
    args = ["--M", "128", "--step-size", "3600", "--iterations", "96", "--time-step", "3600",
            "--initial-infected", "0.05", "--infect-prob", "0.3", "--exposed-duration", "3600",
            "--infectious-duration", "36000", "--resistant-duration", "86400",
            "--start-time", str(T2000), "--seed", "5"]
    _cli_matches_oracle(tmp_path, g, args)
