"""Builds libreplay_hip.so (HIP kernels for gfx950 + C ABI), the `replay` CLI
and the `replay_gen` input generator.

    python -m replay_amd.build        # or replay_amd.build.build()

Outputs (in-tree, git-ignored): replay_amd/lib/libreplay_hip.so,
replay_amd/bin/replay, replay_gen and replay_cgg, and
replay_amd/lib/libreplay_hosttest.so (host evaluation of ingestion and slice
packing, for the CPU tests only).  Compiles every time it is called.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIB_DIR = os.path.join(PKG, "lib")
BIN_DIR = os.path.join(PKG, "bin")
LIB = os.path.join(LIB_DIR, "libreplay_hip.so")
# host evaluation of ingestion and slice packing for the CPU tests; not part
# of the product library or its ABI
HOSTTEST_LIB = os.path.join(LIB_DIR, "libreplay_hosttest.so")
CLI = os.path.join(BIN_DIR, "replay")
GEN_CLI = os.path.join(BIN_DIR, "replay_gen")
CGG_CLI = os.path.join(BIN_DIR, "replay_cgg")

ARCH = os.environ.get("REPLAY_HIP_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

# Exact IEEE fp32 semantics are part of the contract (bit-exact parity with
# the reference order of operations): no contraction into FMAs, no fast math.
# Packed FP32 instructions are disabled as a precaution on this hardware pool.
HIP_FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument", f"-I{INCLUDE}", f"-I{CSRC}"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-pthread", "-ffp-contract=off", f"-I{INCLUDE}",
             f"-I{CSRC}", "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__"]


def _run(cmd: list[str], quiet: bool) -> None:
    if not quiet:
        print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")


def build(quiet: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    os.makedirs(BIN_DIR, exist_ok=True)
    host_objs = []
    for src in ("ingest.cpp", "generator.cpp", "contactgraph.cpp"):
        obj = os.path.join(BUILD, src.replace(".cpp", ".o"))
        _run(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj], quiet)
        host_objs.append(obj)
    hs_obj = os.path.join(BUILD, "host_slices.o")
    _run(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, "host_slices.cpp"), "-o", hs_obj], quiet)
    tmp = HOSTTEST_LIB + ".tmp"
    _run(["g++", "-shared", "-pthread", host_objs[0], hs_obj, "-Wl,--no-undefined", "-o", tmp],
         quiet)
    os.replace(tmp, HOSTTEST_LIB)
    hip_objs = []
    for src in ("replay_hip.hip", "marginal.hip"):
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj], quiet)
        hip_objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", *host_objs, *hip_objs,
          "-o", tmp], quiet)
    os.replace(tmp, LIB)
    _run(["g++", "-O2", "-std=c++17", f"-I{INCLUDE}", os.path.join(CSRC, "replay_main.cpp"),
          f"-L{LIB_DIR}", "-lreplay_hip", "-Wl,-rpath,$ORIGIN/../lib", "-o", CLI], quiet)
    _run(["g++", "-O2", "-std=c++17", f"-I{INCLUDE}", os.path.join(CSRC, "replay_gen_main.cpp"),
          f"-L{LIB_DIR}", "-lreplay_hip", "-Wl,-rpath,$ORIGIN/../lib", "-o", GEN_CLI], quiet)
    _run(["g++", "-O2", "-std=c++17", f"-I{INCLUDE}", os.path.join(CSRC, "replay_cgg_main.cpp"),
          f"-L{LIB_DIR}", "-lreplay_hip", "-Wl,-rpath,$ORIGIN/../lib", "-o", CGG_CLI], quiet)
    return LIB


if __name__ == "__main__":
    print(build(quiet=False))
