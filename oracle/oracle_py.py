"""ctypes bindings of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP product.  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg import this module.

Also binds the reference's own ingestion (tool/temporal_contact_matrix.c and
tool/ncls/ncls/src/intervaldb.c compiled into oracle/_ref/libref_ingest.so by
oracle/Makefile) so the restatement can be checked against it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_ingest.so")


class Interval(C.Structure):
    """Layout of the reference IntervalMap (intervaldb.h:11-18)."""

    _fields_ = [("start", C.c_int64), ("end", C.c_int64),
                ("target_id", C.c_int64), ("sublist", C.c_int64)]


class Sublist(C.Structure):
    _fields_ = [("start", C.c_int), ("len", C.c_int)]


class Graph(C.Structure):
    _fields_ = [("n_nodes", C.c_int), ("n_edges", C.c_int), ("n_indexed", C.c_int),
                ("iv", C.POINTER(Interval)), ("ncls_built", C.c_int),
                ("n_top", C.c_int), ("nlists", C.c_int),
                ("sub", C.POINTER(Sublist)), ("index_to_id", C.POINTER(C.c_int64))]


class Params(C.Structure):
    _fields_ = [("M", C.c_int32), ("step_size", C.c_float), ("iterations", C.c_int32),
                ("initial_infected", C.c_float), ("infect_prob", C.c_float),
                ("exposed_duration", C.c_int32), ("infectious_duration", C.c_int32),
                ("resistant_duration", C.c_int32), ("start_time", C.c_int64),
                ("static_network_duration", C.c_int64), ("seed", C.c_uint64),
                ("threads", C.c_int32), ("use_libm_exp", C.c_int32),
                ("uniforms", C.POINTER(C.c_float)), ("stream", C.c_uint32)]


_lib = None


def build(quiet: bool = True) -> None:
    """Compile liboracle.so / replay_cpu, and oracle/_ref when the reference is present."""
    out = subprocess.DEVNULL if quiet else None
    subprocess.run(["make", "-B", "-C", HERE, "all"], check=True, stdout=out)
    if os.path.isdir("/root/reference/tool"):
        subprocess.run(["make", "-B", "-C", HERE, "ref"], check=True, stdout=out)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.or_parse_file.argtypes = [C.c_char_p, P(P(Graph))]
        L.or_graph_from_contacts.argtypes = [C.c_int, P(C.c_int64), C.c_int, P(C.c_int32),
                                             P(C.c_int32), P(C.c_int64), P(C.c_int64),
                                             P(P(Graph))]
        L.or_graph_free.argtypes = [P(Graph)]
        L.or_build_ncls.argtypes = [P(Graph)]
        L.or_find_intervals.argtypes = [P(Graph), C.c_int64, C.c_int64, C.c_int,
                                        P(Interval), P(C.c_int)]
        L.or_step_csr.argtypes = [P(Graph), P(Params), C.c_int, P(C.c_int32), P(C.c_int32),
                                  P(C.c_float), P(C.c_int64), P(C.c_int64), P(C.c_int64),
                                  P(C.c_float)]
        L.or_build_csr.argtypes = [P(Interval), C.c_int, C.c_int, C.c_int64, C.c_int64,
                                   C.c_float, P(C.c_int32), P(C.c_int32), P(C.c_float),
                                   P(C.c_int64)]
        L.or_parse_iso8601.argtypes = [C.c_char_p]
        L.or_parse_iso8601.restype = C.c_int64
        L.or_expf.argtypes = [C.c_float]
        L.or_expf.restype = C.c_float
        L.or_philox4x32_10.argtypes = [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
        L.or_philox_word.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        L.or_philox_word.restype = C.c_uint32
        L.or_uniform.argtypes = [P(Params), C.c_int, C.c_int, C.c_int, C.c_int]
        L.or_uniform.restype = C.c_float
        L.or_simulate.argtypes = [P(Graph), P(Params), P(C.c_int32), P(C.c_int32),
                                  C.c_void_p, C.c_void_p, P(C.c_double)]
        L.or_simulate_csv.argtypes = [P(Graph), P(Params), C.c_char_p, C.c_char_p,
                                      P(C.c_int32), P(C.c_double)]
        L.or_spmm.argtypes = [P(C.c_int32), P(C.c_int32), P(C.c_float), C.c_int64, C.c_int,
                              P(C.c_float), P(C.c_float), C.c_int, C.c_int]
        L.or_spmm_f64.argtypes = [P(C.c_int32), P(C.c_int32), P(C.c_float), C.c_int,
                                  P(C.c_float), P(C.c_double), C.c_int]
        L.or_state_class.argtypes = [P(Params), C.c_int32]
        L.or_infection_probability.argtypes = [P(Params), C.c_int, P(C.c_int32), P(C.c_float)]
        L.or_infection_probability.restype = None
        L.or_window.argtypes = [P(Params), C.c_void_p, C.c_int, P(C.c_int64), P(C.c_int64),
                                P(C.c_int64), P(C.c_float)]
        _lib = L
    return _lib


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def make_params(M=4, step_size=3600.0, iterations=10, initial_infected=0.0,
                infect_prob=0.0, exposed_duration=0, infectious_duration=0,
                resistant_duration=0, start_time=946713600,
                static_network_duration=3600, seed=0, threads=1,
                use_libm_exp=0, uniforms: np.ndarray | None = None, stream=0) -> Params:
    p = Params(M, step_size, iterations, initial_infected, infect_prob, exposed_duration,
               infectious_duration, resistant_duration, start_time,
               static_network_duration, seed, threads, use_libm_exp, None, stream)
    if uniforms is not None:
        p._uniforms_keepalive = uniforms
        p.uniforms = _ptr(uniforms, C.c_float)
    return p


class OracleGraph:
    """Parsed graph + NCLS, owned by the oracle library."""

    def __init__(self, ptr):
        self.ptr = ptr

    @classmethod
    def from_file(cls, path: str, build_ncls: bool = True) -> "OracleGraph":
        g = C.POINTER(Graph)()
        rc = lib().or_parse_file(path.encode(), C.byref(g))
        if rc != 0:
            raise ValueError(f"oracle parse failed: {rc}")
        obj = cls(g)
        if build_ncls:
            obj.build_ncls()
        return obj

    @classmethod
    def from_contacts(cls, n_nodes: int, src, tgt, start, end, node_list=None,
                      build_ncls: bool = True) -> "OracleGraph":
        src = np.ascontiguousarray(src, dtype=np.int32)
        tgt = np.ascontiguousarray(tgt, dtype=np.int32)
        start = np.ascontiguousarray(start, dtype=np.int64)
        end = np.ascontiguousarray(end, dtype=np.int64)
        nl = None if node_list is None else np.ascontiguousarray(node_list, dtype=np.int64)
        g = C.POINTER(Graph)()
        rc = lib().or_graph_from_contacts(
            n_nodes, None if nl is None else _ptr(nl, C.c_int64), len(src),
            _ptr(src, C.c_int32), _ptr(tgt, C.c_int32), _ptr(start, C.c_int64),
            _ptr(end, C.c_int64), C.byref(g))
        if rc != 0:
            raise ValueError(f"oracle graph build failed: {rc}")
        obj = cls(g)
        if build_ncls:
            obj.build_ncls()
        return obj

    def build_ncls(self):
        rc = lib().or_build_ncls(self.ptr)
        if rc != 0:
            raise ValueError(f"NCLS build failed: {rc}")

    def __del__(self):
        # at interpreter exit the module's globals may already be cleared:
        # the process releases the graph then
        if getattr(self, "ptr", None) and callable(lib):
            lib().or_graph_free(self.ptr)
            self.ptr = None

    @property
    def n_nodes(self) -> int:
        return self.ptr.contents.n_nodes

    @property
    def n_edges(self) -> int:
        return self.ptr.contents.n_edges

    @property
    def n_top(self) -> int:
        return self.ptr.contents.n_top

    @property
    def nlists(self) -> int:
        return self.ptr.contents.nlists

    def intervals(self) -> np.ndarray:
        n = self.n_edges
        buf = (Interval * n).from_address(C.addressof(self.ptr.contents.iv.contents))
        return np.frombuffer(buf, dtype=np.int64).reshape(n, 4).copy()

    def find(self, ws: int, we: int, nbuf: int) -> np.ndarray:
        cap = max(nbuf, 1, self.n_edges)
        buf = (Interval * cap)()
        n = C.c_int(0)
        rc = lib().or_find_intervals(self.ptr, ws, we, nbuf, buf, C.byref(n))
        if rc != 0:
            raise ValueError(rc)
        return np.frombuffer(buf, dtype=np.int64).reshape(cap, 4)[: n.value].copy()

    def step_csr(self, params: Params, it: int):
        N = self.n_nodes
        cap = 2 * max(self.nlists, self.n_top, 1)
        rowptr = np.zeros(N + 1, np.int32)
        col = np.zeros(cap, np.int32)
        val = np.zeros(cap, np.float32)
        nnz = C.c_int64(0)
        ws = C.c_int64(0)
        we = C.c_int64(0)
        sc = C.c_float(0)
        rc = lib().or_step_csr(self.ptr, C.byref(params), it, _ptr(rowptr, C.c_int32),
                               _ptr(col, C.c_int32), _ptr(val, C.c_float), C.byref(nnz),
                               C.byref(ws), C.byref(we), C.byref(sc))
        if rc != 0:
            raise ValueError(rc)
        k = nnz.value
        return rowptr, col[:k].copy(), val[:k].copy(), (ws.value, we.value, sc.value)

    def simulate(self, params: Params, want_state: bool = False):
        N, M, T = self.n_nodes, params.M, params.iterations
        counts = np.zeros((T + 1, M, 4), np.int32)
        state = np.zeros(N * M, np.int32) if want_state else None
        secs = C.c_double(0)
        rc = lib().or_simulate(self.ptr, C.byref(params), _ptr(counts, C.c_int32),
                               None if state is None else _ptr(state, C.c_int32),
                               None, None, C.byref(secs))
        if rc != 0:
            raise ValueError(f"oracle simulate failed: {rc}")
        return counts, state, secs.value


def expf(x: float) -> float:
    return lib().or_expf(x)


def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().or_philox4x32_10(c, k, o)
    return list(o)


def uniform(params: Params, slot: int, node: int, replica: int, n_nodes: int) -> float:
    return lib().or_uniform(C.byref(params), slot, node, replica, n_nodes)


def state_class(params: Params, c: int) -> int:
    return lib().or_state_class(C.byref(params), c)


def infection_probability(params: Params, c, E, library: C.CDLL | None = None) -> np.ndarray:
    """P per (node, replica) from countdowns c and exposure seconds E, both
    N x M column-major (cu:439-458), by the oracle (or by `library`, a build
    of oracle.c loaded elsewhere, with the same signature)."""
    L = library if library is not None else lib()
    c = np.ascontiguousarray(c, dtype=np.int32)
    out = np.array(E, dtype=np.float32, copy=True)
    assert c.size == out.size and c.size % params.M == 0
    L.or_infection_probability(C.byref(params), c.size // params.M, _ptr(c, C.c_int32),
                               _ptr(out, C.c_float))
    return out


# ---------------------------------------------------------------------------
# the reference's own ingestion, compiled from its sources (oracle/_ref)
# ---------------------------------------------------------------------------
class _DBW(C.Structure):
    """IntervalDBWrapper (temporal_contact_matrix.h:45-50)."""

    _fields_ = [("im", C.POINTER(Interval)), ("subheader", C.POINTER(Sublist)),
                ("n", C.c_int), ("nlists", C.c_int)]


_ref = None


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref_lib() -> C.CDLL:
    global _ref
    if _ref is None:
        R = C.CDLL(REF_PATH)
        R.parse_edgelist_build_intervals.argtypes = [C.c_char_p, C.POINTER(C.c_int)]
        R.parse_edgelist_build_intervals.restype = C.POINTER(Interval)
        R.get_total_node_count.restype = C.c_int
        R.build_interval_db_wrapper.argtypes = [C.POINTER(Interval), C.c_int, C.POINTER(_DBW)]
        R.interval_iterator_alloc.restype = C.c_void_p
        R.free_interval_iterator.argtypes = [C.c_void_p]
        R.find_intervals.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.POINTER(Interval),
                                     C.c_int, C.POINTER(Sublist), C.c_int,
                                     C.POINTER(Interval), C.c_int, C.POINTER(C.c_int),
                                     C.POINTER(C.c_void_p)]
        R.find_intervals.restype = C.c_int64
        R.parse_iso8601_to_unix.argtypes = [C.c_char_p]
        R.parse_iso8601_to_unix.restype = C.c_int64
        R.free_node_map.restype = None
        _ref = R
    return _ref


class RefIngest:
    """Reference parse + NCLS of one file (process-global state in the reference)."""

    def __init__(self, path: str):
        R = ref_lib()
        R.free_node_map()
        cnt = C.c_int(0)
        im = R.parse_edgelist_build_intervals(path.encode(), C.byref(cnt))
        if not im or cnt.value == 0:
            raise ValueError("reference parse failed")
        self.n_nodes = R.get_total_node_count()
        self.n_edges = cnt.value
        self.file_order = np.frombuffer(
            (Interval * cnt.value).from_address(C.addressof(im.contents)),
            dtype=np.int64).reshape(cnt.value, 4).copy()
        self.dbw = _DBW()
        if not R.build_interval_db_wrapper(im, cnt.value, C.byref(self.dbw)):
            raise ValueError("reference NCLS build failed")
        self.n_top = self.dbw.n
        self.nlists = self.dbw.nlists
        self.ncls_order = np.frombuffer(
            (Interval * cnt.value).from_address(C.addressof(self.dbw.im.contents)),
            dtype=np.int64).reshape(cnt.value, 4).copy()

    def find(self, ws: int, we: int, nbuf: int) -> np.ndarray:
        """find_intervals exactly as cu:973-989 calls it (buffer max(nlists, n))."""
        R = ref_lib()
        cap = max(self.nlists if self.nlists > 0 else self.n_top, 1)
        cap = max(cap, nbuf, 1)
        buf = (Interval * cap)()
        it = R.interval_iterator_alloc()
        n = C.c_int(0)
        it_ret = C.c_void_p(it)
        R.find_intervals(it, ws, we, self.dbw.im, self.dbw.n, self.dbw.subheader,
                         self.dbw.nlists, buf, nbuf, C.byref(n), C.byref(it_ret))
        R.free_interval_iterator(it)
        return np.frombuffer(buf, dtype=np.int64).reshape(cap, 4)[: n.value].copy()
