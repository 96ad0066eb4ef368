"""C-ABI library and `replay` CLI checks that need no GPU: the library loads
and exports every function include/replay_hip.h declares, and the error paths
the reference takes before touching the device (argv, header, ids, files)
return the documented codes instead of exiting."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, graph_path

import replay_amd
from replay_amd import Params, ReplayError, Simulation


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(replay_amd.LIB_PATH):
        from replay_amd import build
        build.build()
    return replay_amd.load_library()


def test_exports_every_declared_symbol(lib):
    hdr = open(os.path.join(ROOT, "include", "replay_hip.h")).read()
    declared = set(re.findall(r"^\s*(?:const\s+)?[a-z0-9_]+\s*\*?\s*(replay_[a-z_]+)\s*\(", hdr, re.M))
    assert declared == set(replay_amd.EXPORTED_SYMBOLS)
    out = subprocess.run(["nm", "-D", "--defined-only", replay_amd.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (replay_[a-z_]+)$", out, re.M))
    assert declared <= exported
    assert lib.replay_abi_version() == replay_amd.ABI_VERSION


def test_params_defaults_match_reference(lib):
    import ctypes as C
    p = replay_amd._Params()
    lib.replay_params_init(C.byref(p))
    assert (p.M, p.step_size, p.iterations, p.start_time, p.static_network_duration) == \
           (1000, 3600.0, 42, 946713600, 3600)
    d = Params()
    assert (d.M, d.step_size, d.iterations, d.start_time, d.static_network_duration) == \
           (1000, 3600.0, 42, 946713600, 3600)


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


@pytest.mark.parametrize("text,code", [
    ("", replay_amd.E_FORMAT),
    ("NODELIST 1 2\n1, 2, 2000-01-01T00:00:00Z, 2000-01-01T01:00:00Z;\n", replay_amd.E_FORMAT),
    ("NODE_LIST\n", replay_amd.E_FORMAT),
    ("NODE_LIST 1 x2\n", replay_amd.E_NODE_ID),
    ("NODE_LIST 1 -2\n", replay_amd.E_NODE_ID),
    ("NODE_LIST 1 2 1\n", replay_amd.E_DUP_ID),
    ("NODE_LIST  \n", replay_amd.E_NODE_ID),
    ("NODE_LIST 1 2\n# only a comment\n", replay_amd.E_NO_EDGES),
    ("NODE_LIST 1 2\n1, 2, bad, 2000-01-01T01:00:00Z;\n", replay_amd.E_NO_EDGES),
    ("NODE_LIST 1 2\n1, 3, 2000-01-01T00:00:00Z, 2000-01-01T01:00:00Z;\n"
     "2, 4, 2000-01-01T00:00:00Z, 2000-01-01T01:00:00Z;\n", replay_amd.E_NODE_RANGE),
    ("NODE_LIST 1 2\n1 , 2, 2000-01-01T00:00:00Z, 2000-01-01T01:00:00Z;\n", replay_amd.E_NO_EDGES),
    ("NODE_LIST 1 2\n1, 2, 2000-01-01T00:00:00Z, 2000-01-01T01:00:00Z;\n"
     "1 , 2, 2000-01-01T00:00:00Z, 2000-01-01T01:00:00Z;\n", replay_amd.E_COUNT),
    ("NODE_LIST 1 2\n1, 2, 1969-01-01T00:00:00Z, 1969-01-01T01:00:00Z;\n", replay_amd.E_UNSUPPORTED),
    ("NODE_LIST 1 2\n1, 2, 2000-01-01T02:00:00Z, 2000-01-01T01:00:00Z;\n", replay_amd.E_UNSUPPORTED),
])
def test_open_errors_without_device(lib, tmp_path, oracle, text, code):
    """Each of these fails before any device call, with the code for the
    reference's fatal path (tcm:159-254, 308-312; cu:735-738)."""
    path = _write(tmp_path, "g.txt", text)
    with pytest.raises(ReplayError) as e:
        Simulation.from_file(path, Params(M=2, iterations=2))
    assert e.value.code == code
    # the oracle (restating the reference) rejects the same inputs, except the
    # two the engine refuses on purpose (pre-1970 / inverted intervals)
    if code != replay_amd.E_UNSUPPORTED:
        with pytest.raises(ValueError):
            oracle.OracleGraph.from_file(path)


def test_missing_file(lib):
    with pytest.raises(ReplayError) as e:
        Simulation.from_file("/nonexistent/graph.txt", Params())
    assert e.value.code == replay_amd.E_IO


@pytest.mark.parametrize("kw", [dict(M=0), dict(step_size=0.0), dict(static_network_duration=0),
                                dict(iterations=-1)])
def test_bad_params(lib, kw):
    with pytest.raises(ReplayError) as e:
        Simulation.from_file(graph_path("readme_example.txt"), Params(**kw))
    assert e.value.code == replay_amd.E_ARG


def _cli(*args):
    if not os.path.exists(replay_amd.CLI_PATH):
        pytest.skip("replay CLI not built")
    return subprocess.run([replay_amd.CLI_PATH, *args], capture_output=True, text=True)


def test_cli_argv_errors(lib, tmp_path):
    r = _cli()
    assert r.returncode == 1 and "Usage:" in r.stderr
    r = _cli("g.txt", "--N", "5")
    assert r.returncode == 1 and "Unrecognized option: --N" in r.stderr
    r = _cli("g.txt", "--M")  # value flag without a value
    assert r.returncode == 1 and "Unrecognized option: --M" in r.stderr
    r = _cli("--M", "4")
    assert r.returncode == 1 and "Missing contact network file argument." in r.stderr
    r = _cli("g.txt", "--summary-out", str(tmp_path / "no" / "dir" / "s.csv"))
    assert r.returncode == 1 and "Could not open summary output file" in r.stderr
    r = _cli("g.txt", "--node-state-out", str(tmp_path / "no" / "n.csv"))
    assert r.returncode == 1 and "Could not open node state output file" in r.stderr


def test_cli_cpu_only_is_refused(lib, tmp_path):
    """The reference's --cpu-only runs its CPU twin (cu:1109-1144); this
    build has no CPU engine, so the flag fails loudly instead of handing a
    caller that times the CPU path a GPU run."""
    path = _write(tmp_path, "g.txt", "NODE_LIST 1 2\n1 2 2000-01-01T00:00:00 2000-01-01T01:00:00;\n")
    r = _cli(path, "--cpu-only", "--M", "2", "--iterations", "1")
    assert r.returncode == 1
    assert "--cpu-only is not supported" in r.stderr
    assert "Starting temporal Monte Carlo" not in r.stdout
    r = _cli("--cpu-threads")  # a value flag without its value is still an argv error
    assert r.returncode == 1 and "Unrecognized option: --cpu-threads" in r.stderr


def test_cli_parse_error_exit_code(lib, tmp_path):
    path = _write(tmp_path, "bad.txt", "NODE_LIST 1 1\n")
    out = tmp_path / "s.csv"
    out.write_text("stale")
    r = _cli(path, "-q", "--summary-out", str(out))
    assert r.returncode == 1
    assert "No intervals read or parse error." in r.stderr
    # the old file was replaced by a fresh one holding only the header
    assert out.read_text() == "time,simulation_id,exposed,infectious,resistant,susceptible\n"
