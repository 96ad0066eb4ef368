"""replay_amd -- Python host for libreplay_hip (MI355X / gfx950).

ctypes bindings of include/replay_hip.h.  The reference drives its engine as a
subprocess (`replay <graph> [flags]`, see run_sim_gui.py:110-141 and
benchmark/benchmark_with_trials.py:57-63 in the reference); this module gives
the same engine in-process: open a graph (file or arrays), run iterations on
the GPU, read back SEIR counts, states, slices and exposures.

There is no CPU fallback: if libreplay_hip.so is missing or no HIP device is
present, opening a simulation raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libreplay_hip.so")
CLI_PATH = os.path.join(PKG_DIR, "bin", "replay")
GEN_PATH = os.path.join(PKG_DIR, "bin", "replay_gen")
CGG_PATH = os.path.join(PKG_DIR, "bin", "replay_cgg")

ABI_VERSION = 3

# error codes (include/replay_hip.h)
E_ARG, E_IO, E_FORMAT, E_NODE_ID, E_DUP_ID, E_NO_EDGES, E_COUNT, E_NODE_RANGE = \
    -1, -2, -3, -4, -5, -6, -7, -8
E_NOMEM, E_DEVICE, E_OVERFLOW, E_UNSUPPORTED, E_STATE = -9, -10, -11, -12, -13


class ReplayError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"replay_hip error {code}: {message}")
        self.code = code


class _Params(C.Structure):
    _fields_ = [("M", C.c_int32), ("step_size", C.c_float), ("iterations", C.c_int32),
                ("initial_infected", C.c_float), ("infect_prob", C.c_float),
                ("exposed_duration", C.c_int32), ("infectious_duration", C.c_int32),
                ("resistant_duration", C.c_int32), ("start_time", C.c_int64),
                ("static_network_duration", C.c_int64), ("seed", C.c_uint64),
                ("device", C.c_int32), ("windows_per_epoch", C.c_int32),
                ("max_slice_bytes", C.c_int64), ("stream", C.c_uint32), ("flags", C.c_int32),
                ("max_ring_bytes", C.c_int64)]


class _Network(C.Structure):
    _fields_ = [("n_node_list", C.c_int32), ("node_ids", C.POINTER(C.c_int64)),
                ("n_contacts", C.c_int64), ("src", C.POINTER(C.c_int32)),
                ("tgt", C.POINTER(C.c_int32)), ("start", C.POINTER(C.c_int64)),
                ("end", C.POINTER(C.c_int64)), ("initial_infected", C.c_float),
                ("infect_prob", C.c_float), ("exposed_duration", C.c_int32),
                ("infectious_duration", C.c_int32), ("resistant_duration", C.c_int32),
                ("stream", C.c_uint32), ("M", C.c_int32)]


class _Info(C.Structure):
    _fields_ = [("n_nodes", C.c_int32), ("M", C.c_int32), ("n_contacts", C.c_int64),
                ("nlists", C.c_int64), ("iterations", C.c_int32), ("n_windows", C.c_int32),
                ("n_epochs", C.c_int32), ("windows_per_epoch", C.c_int32),
                ("n_entries", C.c_int64), ("slice_bytes", C.c_int64),
                ("state_bytes", C.c_int64), ("steps_done", C.c_int32),
                ("truncated_windows", C.c_int32), ("step_bytes", C.c_int64),
                ("ring_depth", C.c_int32), ("n_networks", C.c_int32), ("variant", C.c_int32),
                ("graph_ready", C.c_int32), ("entry_bytes", C.c_int32),
                ("replicas_total", C.c_int32), ("ring_form", C.c_int32),
                ("has_entries8", C.c_int32),
                ("ring_bytes", C.c_int64), ("tune_ms", C.c_float * 4)]


F_GRAPH = 1
F_AUTOTUNE = 2
F_WIDE_ENTRIES = 4
F_NO_HISTORY = 8
RING_OFFSETS, RING_COMPACT, RING_HISTORY = 0, 1, 2
HISTORY_MAX_ITERATIONS = 1024  # longest run the library keeps in the history form
V_WORD, V_NODE4, V_NODE16, V_SLICED = 0, 4, 16, 1
# or'ed with V_NODE4 / V_NODE16: the exposure kernel in blocks of one wavefront
V_BLOCK1 = 0x100
# or'ed with V_WORD / V_NODE4 / V_NODE16 (and V_BLOCK1): XCD x takes the x-th
# contiguous eighth of the nodes instead of every eighth block
V_XCD = 0x200
# the variants and schedules the parity tests run (replay_set_variant also
# accepts the other combinations of V_XCD; SLICED only for a uniform M)
VARIANTS_ALL = (V_WORD, V_NODE4, V_NODE4 | V_BLOCK1, V_NODE16, V_NODE16 | V_BLOCK1, V_SLICED,
                V_WORD | V_XCD, V_NODE4 | V_XCD, V_NODE16 | V_BLOCK1 | V_XCD)
# every variant and schedule replay_autotune times
VARIANTS_TUNED = tuple(v | x for x in (0, V_XCD)
                       for v in (V_WORD, V_NODE4, V_NODE4 | V_BLOCK1, V_NODE16,
                                 V_NODE16 | V_BLOCK1)) + (V_SLICED,)


EXPORTED_SYMBOLS = (
    "replay_abi_version", "replay_params_init", "replay_open", "replay_open_contacts",
    "replay_get_info", "replay_reset", "replay_run", "replay_run_timed", "replay_sync",
    "replay_get_counts",
    "replay_get_state", "replay_iteration_time", "replay_set_uniforms", "replay_get_slice",
    "replay_get_exposure", "replay_last_error", "replay_error_string", "replay_close",
    "replay_open_batch", "replay_get_network_offsets", "replay_autotune", "replay_set_variant",
    "replay_marginal_create", "replay_marginal_forward", "replay_marginal_backward",
    "replay_marginal_destroy", "replay_generate_groups", "replay_contactgraph",
    "replay_get_replica_offsets", "replay_get_network_state", "replay_get_exposure_noalloc",
)

_lib = None


def library_path() -> str:
    return LIB_PATH


def load_library() -> C.CDLL:
    """Loads libreplay_hip.so; raises if it has not been built.

    A process must hold one HIP runtime.  PyTorch-ROCm ships its own
    libamdhip64 / libhsa-runtime64, loaded by path; if this library were
    loaded first, its dependency would resolve to the system ROCm copies and
    torch would then load a second runtime that finds no device (and device
    pointers would not be shared between the two).  So torch, when present,
    is imported first: this library's libamdhip64.so.7 dependency then binds
    to the runtime torch already holds."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m replay_amd.build`")
    try:
        import torch  # noqa: F401  (one HIP runtime per process, see above)
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    ctx_p = C.c_void_p
    L.replay_abi_version.restype = C.c_int32
    L.replay_params_init.argtypes = [P(_Params)]
    L.replay_params_init.restype = None
    L.replay_open.argtypes = [C.c_char_p, P(_Params), P(ctx_p)]
    L.replay_open_contacts.argtypes = [C.c_int32, P(C.c_int64), C.c_int64, P(C.c_int32),
                                       P(C.c_int32), P(C.c_int64), P(C.c_int64), P(_Params),
                                       P(ctx_p)]
    L.replay_open_batch.argtypes = [C.c_int32, P(_Network), P(_Params), P(ctx_p)]
    L.replay_get_network_offsets.argtypes = [ctx_p, P(C.c_int32), C.c_int32]
    L.replay_get_replica_offsets.argtypes = [ctx_p, P(C.c_int32), C.c_int32]
    L.replay_get_network_state.argtypes = [ctx_p, C.c_int32, P(C.c_int32), C.c_int64]
    L.replay_autotune.argtypes = [ctx_p, C.c_int32, P(C.c_int32)]
    L.replay_set_variant.argtypes = [ctx_p, C.c_int32]
    L.replay_get_info.argtypes = [ctx_p, P(_Info)]
    L.replay_reset.argtypes = [ctx_p]
    L.replay_run.argtypes = [ctx_p, C.c_int32]
    L.replay_run_timed.argtypes = [ctx_p, C.c_int32, P(C.c_double)]
    L.replay_sync.argtypes = [ctx_p]
    L.replay_get_counts.argtypes = [ctx_p, P(C.c_int32), C.c_int64]
    L.replay_get_state.argtypes = [ctx_p, P(C.c_int32), C.c_int64]
    L.replay_iteration_time.argtypes = [ctx_p, C.c_int32]
    L.replay_iteration_time.restype = C.c_int64
    L.replay_set_uniforms.argtypes = [ctx_p, P(C.c_float), C.c_int64]
    L.replay_get_slice.argtypes = [ctx_p, C.c_int32, P(C.c_int32), P(C.c_int32), P(C.c_float),
                                   C.c_int64, P(C.c_int64)]
    L.replay_get_exposure.argtypes = [ctx_p, C.c_int32, P(C.c_float), C.c_int64]
    L.replay_get_exposure_noalloc.argtypes = [ctx_p, C.c_int32, P(C.c_float), C.c_int64]
    L.replay_last_error.argtypes = [ctx_p]
    L.replay_last_error.restype = C.c_char_p
    L.replay_error_string.argtypes = [C.c_int32]
    L.replay_error_string.restype = C.c_char_p
    L.replay_close.argtypes = [ctx_p]
    L.replay_close.restype = None
    vp = C.c_void_p
    L.replay_marginal_create.argtypes = [ctx_p, C.c_int32, C.c_int32, P(vp)]
    L.replay_marginal_forward.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.replay_marginal_backward.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.replay_marginal_destroy.argtypes = [vp]
    L.replay_marginal_destroy.restype = None
    L.replay_generate_groups.argtypes = [C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                         C.c_double, C.c_double, C.c_int32, C.c_char_p,
                                         P(C.c_int64), C.c_int32]
    L.replay_generate_groups.restype = C.c_int32
    L.replay_contactgraph.argtypes = [C.c_char_p, C.c_char_p, P(C.c_int64), C.c_int32]
    L.replay_contactgraph.restype = C.c_int32
    for name in ("replay_marginal_create", "replay_marginal_forward", "replay_marginal_backward"):
        getattr(L, name).restype = C.c_int32
    for name in ("replay_open", "replay_open_contacts", "replay_get_info", "replay_reset",
                 "replay_run", "replay_run_timed", "replay_sync", "replay_get_counts", "replay_get_state",
                 "replay_set_uniforms", "replay_get_slice", "replay_get_exposure",
                 "replay_get_exposure_noalloc",
                 "replay_open_batch", "replay_get_network_offsets", "replay_autotune",
                 "replay_get_replica_offsets", "replay_get_network_state",
                 "replay_set_variant"):
        getattr(L, name).restype = C.c_int32
    if L.replay_abi_version() != ABI_VERSION:
        raise RuntimeError("libreplay_hip ABI version mismatch")
    _lib = L
    return L


@dataclass
class Params:
    """Simulation parameters; defaults are the reference's (cu:25-27, 48-56)."""

    M: int = 1000
    step_size: float = 3600.0
    iterations: int = 42
    initial_infected: float = 0.0
    infect_prob: float = 0.0
    exposed_duration: int = 0
    infectious_duration: int = 0
    resistant_duration: int = 0
    start_time: int = 946713600
    static_network_duration: int = 3600
    seed: int = 0
    device: int = 0
    windows_per_epoch: int = 0
    max_slice_bytes: int = 0
    stream: int = 0
    flags: int = 0
    max_ring_bytes: int = 0

    def _c(self) -> _Params:
        return _Params(self.M, self.step_size, self.iterations, self.initial_infected,
                       self.infect_prob, self.exposed_duration, self.infectious_duration,
                       self.resistant_duration, self.start_time, self.static_network_duration,
                       self.seed, self.device, self.windows_per_epoch, self.max_slice_bytes,
                       self.stream, self.flags, self.max_ring_bytes)


@dataclass
class Network:
    """One network of a batch: contacts (ids, Unix start/end) and its own
    epidemiological parameters (replay_network)."""

    n_nodes: int
    src: np.ndarray
    tgt: np.ndarray
    start: np.ndarray
    end: np.ndarray
    initial_infected: float = 0.0
    infect_prob: float = 0.0
    exposed_duration: int = 0
    infectious_duration: int = 0
    resistant_duration: int = 0
    stream: int = 0
    M: int = 0  # replicas of this network; 0 takes the batch's Params.M


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class Simulation:
    """One network with M replicas on one GPU (a replay_ctx)."""

    def __init__(self, ctx: C.c_void_p, params: Params):
        self._ctx = ctx
        self.params = params
        info = self.info()
        self.n_nodes = info["n_nodes"]
        self.M = info["M"]
        self.n_networks = info["n_networks"]
        off = np.zeros(self.n_networks + 1, np.int32)
        self._check(load_library().replay_get_network_offsets(ctx, _ptr(off, C.c_int32), off.size))
        self.node_offsets = off
        roff = np.zeros(self.n_networks + 1, np.int32)
        self._check(load_library().replay_get_replica_offsets(ctx, _ptr(roff, C.c_int32),
                                                              roff.size))
        self.replica_offsets = roff
        # every network has M replicas (the counts and state layouts below)
        self.uniform = bool((np.diff(roff) == self.M).all())

    # construction -------------------------------------------------------
    @classmethod
    def from_file(cls, path: str, params: Params) -> "Simulation":
        L = load_library()
        ctx = C.c_void_p()
        p = params._c()
        rc = L.replay_open(os.fsencode(path), C.byref(p), C.byref(ctx))
        if rc != 0:
            raise ReplayError(rc, L.replay_last_error(None).decode())
        return cls(ctx, params)

    @classmethod
    def from_contacts(cls, n_nodes: int, src, tgt, start, end, params: Params,
                      node_ids=None) -> "Simulation":
        L = load_library()
        src = np.ascontiguousarray(src, dtype=np.int32)
        tgt = np.ascontiguousarray(tgt, dtype=np.int32)
        start = np.ascontiguousarray(start, dtype=np.int64)
        end = np.ascontiguousarray(end, dtype=np.int64)
        ids = None if node_ids is None else np.ascontiguousarray(node_ids, dtype=np.int64)
        ctx = C.c_void_p()
        p = params._c()
        rc = L.replay_open_contacts(n_nodes, None if ids is None else _ptr(ids, C.c_int64),
                                    len(src), _ptr(src, C.c_int32), _ptr(tgt, C.c_int32),
                                    _ptr(start, C.c_int64), _ptr(end, C.c_int64),
                                    C.byref(p), C.byref(ctx))
        if rc != 0:
            raise ReplayError(rc, L.replay_last_error(None).decode())
        return cls(ctx, params)

    @classmethod
    def from_networks(cls, networks: list, params: Params) -> "Simulation":
        """Batched form: every network advances in the same step launches."""
        L = load_library()
        keep = []
        arr = (_Network * len(networks))()
        for k, nw in enumerate(networks):
            a = [np.ascontiguousarray(nw.src, np.int32), np.ascontiguousarray(nw.tgt, np.int32),
                 np.ascontiguousarray(nw.start, np.int64), np.ascontiguousarray(nw.end, np.int64)]
            keep.append(a)
            arr[k] = _Network(nw.n_nodes, None, len(a[0]), _ptr(a[0], C.c_int32),
                              _ptr(a[1], C.c_int32), _ptr(a[2], C.c_int64), _ptr(a[3], C.c_int64),
                              nw.initial_infected, nw.infect_prob, nw.exposed_duration,
                              nw.infectious_duration, nw.resistant_duration, nw.stream,
                              getattr(nw, "M", 0))
        ctx = C.c_void_p()
        p = params._c()
        rc = L.replay_open_batch(len(networks), arr, C.byref(p), C.byref(ctx))
        if rc != 0:
            raise ReplayError(rc, L.replay_last_error(None).decode())
        return cls(ctx, params)

    # lifecycle ----------------------------------------------------------
    def close(self) -> None:
        if self._ctx:
            load_library().replay_close(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise ReplayError(rc, load_library().replay_last_error(self._ctx).decode())

    # queries ------------------------------------------------------------
    def info(self) -> dict:
        inf = _Info()
        self._check(load_library().replay_get_info(self._ctx, C.byref(inf)))
        out = {name: getattr(inf, name) for name, _ in _Info._fields_}
        out["tune_ms"] = list(inf.tune_ms)
        return out

    def reset(self) -> None:
        self._check(load_library().replay_reset(self._ctx))

    def run(self, steps: int | None = None) -> None:
        if steps is None:
            steps = self.params.iterations - self.info()["steps_done"]
        self._check(load_library().replay_run(self._ctx, steps))

    def run_timed(self, steps: int | None = None) -> float:
        """Runs `steps` iterations and returns their device time in ms."""
        if steps is None:
            steps = self.params.iterations - self.info()["steps_done"]
        ms = C.c_double(0)
        self._check(load_library().replay_run_timed(self._ctx, steps, C.byref(ms)))
        return ms.value

    def sync(self) -> None:
        self._check(load_library().replay_sync(self._ctx))

    @property
    def steps_done(self) -> int:
        return self.info()["steps_done"]

    def _all_counts(self) -> np.ndarray:
        k = self.steps_done
        out = np.zeros((k + 1, int(self.replica_offsets[-1]), 4), np.int32)
        self._check(load_library().replay_get_counts(self._ctx, _ptr(out, C.c_int32), out.size))
        return out

    def counts(self) -> np.ndarray:
        """(steps_done+1, M, 4) int32 for one network, (steps_done+1,
        n_networks, M, 4) for a batch: exposed, infectious, resistant,
        susceptible.  A batch whose networks have different replica counts
        has no such array: use network_counts."""
        if not self.uniform:
            raise ValueError("networks have different replica counts: use network_counts(k)")
        out = self._all_counts().reshape(-1, self.n_networks, self.M, 4)
        return out[:, 0] if self.n_networks == 1 else out

    def network_counts(self, k: int) -> np.ndarray:
        """(steps_done+1, M_k, 4) int32 counts of network k's replicas."""
        r0, r1 = int(self.replica_offsets[k]), int(self.replica_offsets[k + 1])
        return self._all_counts()[:, r0:r1].copy()

    def network_state(self, k: int) -> np.ndarray:
        """N_k*M_k int32 countdowns of network k, column-major (node + sim*N_k)."""
        n = int(self.node_offsets[k + 1] - self.node_offsets[k])
        out = np.zeros(n * int(self.replica_offsets[k + 1] - self.replica_offsets[k]), np.int32)
        self._check(load_library().replay_get_network_state(self._ctx, k, _ptr(out, C.c_int32),
                                                            out.size))
        return out

    def autotune(self, steps: int = 16) -> int:
        v = C.c_int32(0)
        self._check(load_library().replay_autotune(self._ctx, steps, C.byref(v)))
        return v.value

    def set_variant(self, variant: int) -> None:
        self._check(load_library().replay_set_variant(self._ctx, variant))

    def state(self) -> np.ndarray:
        """N*M int32 countdowns, column-major (idx = node + sim*N)."""
        out = np.zeros(self.n_nodes * self.M, np.int32)
        self._check(load_library().replay_get_state(self._ctx, _ptr(out, C.c_int32), out.size))
        return out

    def iteration_time(self, k: int) -> int:
        return load_library().replay_iteration_time(self._ctx, k)

    def set_uniforms(self, u: np.ndarray | None) -> None:
        L = load_library()
        if u is None:
            self._check(L.replay_set_uniforms(self._ctx, None, 0))
            return
        u = np.ascontiguousarray(u, dtype=np.float32).ravel()
        self._check(L.replay_set_uniforms(self._ctx, _ptr(u, C.c_float), u.size))

    def slice(self, k: int):
        """CSR (rowptr, colind, val) of iteration k's window, built on the device."""
        L = load_library()
        rowptr = np.zeros(self.n_nodes + 1, np.int32)
        nnz = C.c_int64(0)
        L.replay_get_slice(self._ctx, k, _ptr(rowptr, C.c_int32), None, None, 0, C.byref(nnz))
        n = nnz.value
        col = np.zeros(max(n, 1), np.int32)
        val = np.zeros(max(n, 1), np.float32)
        self._check(L.replay_get_slice(self._ctx, k, _ptr(rowptr, C.c_int32),
                                       _ptr(col, C.c_int32), _ptr(val, C.c_float), n,
                                       C.byref(nnz)))
        return rowptr, col[:n], val[:n]

    def slice_nnz(self, k: int) -> int:
        """Nonzeros of iteration k's window CSR (merged columns), counted on
        the device without building the CSR."""
        rowptr = np.zeros(self.n_nodes + 1, np.int32)
        nnz = C.c_int64(0)
        self._check(load_library().replay_get_slice(self._ctx, k, _ptr(rowptr, C.c_int32), None,
                                                    None, -1, C.byref(nnz)))
        return nnz.value

    def exposure(self, k: int, noalloc: bool = False) -> np.ndarray:
        """Exposure seconds (N*M, column-major) of the current state in window
        k; noalloc stages it in the library's static device array instead of
        a buffer allocated for the call (N*M up to 2^24)."""
        out = np.zeros(self.n_nodes * self.M, np.float32)
        L = load_library()
        fn = L.replay_get_exposure_noalloc if noalloc else L.replay_get_exposure
        self._check(fn(self._ctx, k, _ptr(out, C.c_float), out.size))
        return out


def write_summary_csv(path: str, sim: Simulation) -> None:
    """Summary CSV in the reference format (cu:707-709, 558-561)."""
    counts = sim.counts()
    with open(path, "w") as f:
        f.write("time,simulation_id,exposed,infectious,resistant,susceptible\n")
        for k in range(counts.shape[0]):
            t = sim.iteration_time(k - 1)
            for s in range(counts.shape[1]):
                c = counts[k, s]
                f.write(f"{t},{s},{c[0]},{c[1]},{c[2]},{c[3]}\n")


def generate_groups(path: str, base_time: int, steps: int, step_size: int, population: int,
                    groups: int, prob_inner: float, prob_inter: float, seed: int = 0,
                    verbose: bool = False) -> int:
    """Writes the group-structured contact graph of
    generate_networks/graph_generator.c:39-157 for these arguments (base_time
    in UTC seconds) to path; returns the number of contact lines.  Needs no
    GPU."""
    L = load_library()
    n = C.c_int64(0)
    rc = L.replay_generate_groups(int(base_time), int(steps), int(step_size), int(population),
                                  int(groups), float(prob_inner), float(prob_inter), int(seed),
                                  path.encode(), C.byref(n), 1 if verbose else 0)
    if rc != 0:
        raise ReplayError(rc, L.replay_error_string(rc).decode())
    return n.value


def contact_graph(commands_path: str, out_path: str, quiet: bool = False) -> int:
    """Writes the contact graph of an ADDNODE / ADDEDGE command file, as
    contactgraphgen/contactgraphgen.py does, and returns the number of
    contact lines.  Messages for rejected command lines go to stderr unless
    quiet.  Raises ReplayError when the command file cannot be read, the
    graph cannot be written or a time passes 9999-12-31T23:59:59Z.  Needs no
    GPU."""
    L = load_library()
    n = C.c_int64(0)
    rc = L.replay_contactgraph(commands_path.encode(), out_path.encode(), C.byref(n),
                               1 if quiet else 0)
    if rc != 0:
        what = {-2: "the command file cannot be read or the graph cannot be written",
                -3: "the command file is not UTF-8",
                -9: "the commands expand to more than 2^30 contact occurrences",
                -11: f"a time passes 9999-12-31T23:59:59Z; {out_path} ends before it "
                     f"({n.value} contact lines)"}
        raise ReplayError(rc, what.get(rc, L.replay_error_string(rc).decode()))
    return n.value


__all__ = ["Params", "Network", "Simulation", "ReplayError", "load_library", "library_path",
           "write_summary_csv", "generate_groups", "contact_graph", "EXPORTED_SYMBOLS", "CLI_PATH",
           "GEN_PATH", "CGG_PATH", "LIB_PATH"]
