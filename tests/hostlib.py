"""Loader of libreplay_hosttest.so: ingestion, window planning and slice
packing evaluated on the host (replay_amd/csrc/host_slices.cpp), for the CPU
tests.  It is not the product library and not part of its ABI."""
import ctypes as C
import os

import replay_amd

PATH = os.path.join(os.path.dirname(replay_amd.LIB_PATH), "libreplay_hosttest.so")


def load() -> C.CDLL:
    if not os.path.exists(PATH):
        from replay_amd import build
        build.build()
    return C.CDLL(PATH)
