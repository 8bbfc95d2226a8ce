"""ctypes binding of liblouvain_amd.so (the C ABI declared in include/lv_api.h).

Host-side mirror of the reference's interface for tests and benchmarks:
``Graph`` is GraphHost (graph.cu:17-87), ``Community`` is Community
(community.h:15-107) with ``optimization_phase`` / ``aggregation_phase`` /
``refine_phase`` standing for OptimizationPhase::run, AggregationPhase::run
and RefinePhase::run, and ``louvain`` is ModularityAlgorithms::Louvain
(modularityAlgorithms.h:8-81).  Every call goes to the HIP library; there is
no CPU fallback: a missing library raises at import.
"""
import ctypes as C
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblouvain_amd.so")

HASH, SORT, ADAPTIVE_SPEED, ADAPTIVE_MEMORY = 0, 1, 2, 3
SWEEP_FAST, SWEEP_SORT, SWEEP_HASH = 0, 1, 2
GAIN_EXACT, GAIN_REF_F32 = 0, 1
TIE_LOWEST, TIE_REFERENCE = 0, 1
MODES = {"hash": HASH, "sort": SORT, "adaptive_speed": ADAPTIVE_SPEED,
         "adaptive_memory": ADAPTIVE_MEMORY}


class LouvainError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(msg)
        self.status = status


class Params(C.Structure):
    _fields_ = [("mode", C.c_int32), ("gain", C.c_int32), ("tie", C.c_int32),
                ("chunk_emulation", C.c_int32), ("chunk_entries", C.c_uint64),
                ("sweep_threshold", C.c_double), ("level_threshold", C.c_double),
                ("max_sweeps", C.c_int64), ("max_levels", C.c_int32), ("profile", C.c_int32),
                ("autotune", C.c_int32), ("refine", C.c_int32),
                ("bucket_max_deg", C.c_uint32 * 8)]


class SweepRec(C.Structure):
    _fields_ = [("level", C.c_int32), ("sweep", C.c_int32), ("kind", C.c_int32),
                ("pre_decided", C.c_int32), ("n", C.c_uint32), ("n_active", C.c_uint32),
                ("n_moved", C.c_uint32), ("n_considered", C.c_uint32),
                ("E", C.c_uint64), ("e_active", C.c_uint64), ("sigma_in", C.c_int64),
                ("sum_tot_sq", C.c_uint64), ("q", C.c_double), ("comm_hash", C.c_uint64),
                ("move_ms", C.c_double), ("sweep_ms", C.c_double), ("candidates", C.c_uint64),
                ("move_bytes", C.c_uint64), ("commit_ms", C.c_double),
                ("commit_bytes", C.c_uint64), ("mover_entries", C.c_uint64)]


class LevelRec(C.Structure):
    _fields_ = [("level", C.c_int32), ("sweeps", C.c_int32), ("n", C.c_uint32),
                ("n_comm", C.c_uint32), ("E", C.c_uint64), ("q_start", C.c_double),
                ("q_end", C.c_double), ("best_hash", C.c_uint64), ("contract_ms", C.c_double),
                ("E_new", C.c_uint64), ("contract_bytes", C.c_uint64)]


class Result(C.Structure):
    _fields_ = [("modularity", C.c_double), ("n_of_best_communities", C.c_uint32),
                ("levels", C.c_int32), ("total_sweeps", C.c_int32), ("hash_trace", C.c_int32),
                ("best_communities", C.POINTER(C.c_uint32)),
                ("sweeps", C.POINTER(SweepRec)), ("sweeps_cap", C.c_int32),
                ("level_recs", C.POINTER(LevelRec)), ("levels_cap", C.c_int32),
                ("seconds", C.c_double), ("traversed_entries", C.c_uint64)]


class KernelStats(C.Structure):
    _fields_ = [("move_launches", C.c_uint64), ("move_ms", C.c_double),
                ("move_bytes", C.c_uint64), ("move_entries", C.c_uint64),
                ("sweep_ms", C.c_double), ("sweeps", C.c_uint64),
                ("variant_launches", C.c_uint64 * 96), ("variant_ms", C.c_double * 96),
                ("variant_bytes", C.c_uint64 * 96), ("variant_entries", C.c_uint64 * 96),
                ("bucket_launches", C.c_uint64 * 18), ("bucket_ms", C.c_double * 18),
                ("bucket_bytes", C.c_uint64 * 18), ("bucket_entries", C.c_uint64 * 18),
                ("commit_passes", C.c_uint64), ("commit_ms", C.c_double),
                ("commit_bytes", C.c_uint64), ("commit_entries", C.c_uint64),
                ("contractions", C.c_uint64), ("contract_ms", C.c_double),
                ("contract_bytes", C.c_uint64), ("contract_entries", C.c_uint64)]


class BatchResult(C.Structure):
    _fields_ = [("modularity", C.POINTER(C.c_double)),
                ("n_of_best_communities", C.POINTER(C.c_uint32)),
                ("levels", C.POINTER(C.c_int32)), ("sweeps", C.POINTER(C.c_int32)),
                ("best_communities", C.POINTER(C.c_uint32)),
                ("union_levels", C.c_int32), ("union_sweeps", C.c_int32),
                ("seconds", C.c_double), ("traversed_entries", C.c_uint64)]


class JobResult(C.Structure):
    _fields_ = [("status", C.c_int), ("modularity", C.c_double),
                ("n_of_best_communities", C.c_uint32), ("n", C.c_uint32),
                ("levels", C.c_int32), ("sweeps", C.c_int32),
                ("seconds_queued", C.c_double), ("seconds_total", C.c_double)]


class IngestInfo(C.Structure):
    _fields_ = [("min_id", C.c_int64), ("n", C.c_uint32), ("E", C.c_uint64),
                ("total_weight", C.c_double), ("seconds_read_upload", C.c_double),
                ("seconds_tokenize", C.c_double), ("seconds_build", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("admitted", C.c_uint64), ("retired", C.c_uint64),
                ("uploaded_bytes", C.c_uint64), ("seconds_assemble", C.c_double),
                ("seconds_sweeps", C.c_double), ("seconds_aggregate", C.c_double)]


# name -> (restype, argtypes); every symbol include/lv_api.h declares
_vp = C.c_void_p
_u32p = C.POINTER(C.c_uint32)
SIGNATURES = {
    "lv_default_params": (None, [C.POINTER(Params)]),
    "lv_status_string": (C.c_char_p, [C.c_int]),
    "lv_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "lv_graph_load_edge": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.POINTER(_vp)]),
    "lv_graph_from_edges": (C.c_int, [C.c_uint64, _vp, _vp, _vp, C.POINTER(_vp)]),
    "lv_graph_from_csr": (C.c_int, [C.c_uint32, C.c_uint64, _vp, _vp, _vp, C.c_double,
                                    C.POINTER(_vp)]),
    "lv_graph_free": (None, [_vp]),
    "lv_graph_info": (C.c_int, [_vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                                C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "lv_graph_copy_csr": (C.c_int, [_vp, _vp, _vp, _vp]),
    "lv_community_create": (C.c_int, [_vp, _vp, C.POINTER(_vp)]),
    "lv_community_destroy": (None, [_vp]),
    "lv_optimization_phase": (C.c_int, [_vp, C.POINTER(Params)]),
    "lv_aggregation_phase": (C.c_int, [_vp, C.c_int, C.POINTER(Params)]),
    "lv_refine_phase": (C.c_int, [_vp]),
    "lv_refine": (C.c_int, [_vp, C.POINTER(Params), C.POINTER(Result)]),
    "lv_modularity": (C.c_double, [_vp]),
    "lv_best_communities": (C.c_int, [_vp, _vp, C.c_size_t, C.POINTER(C.c_uint32)]),
    "lv_louvain": (C.c_int, [_vp, C.POINTER(Params), C.POINTER(Result)]),
    "lv_community_reset": (C.c_int, [_vp]),
    "lv_community_run": (C.c_int, [_vp, C.POINTER(Params), C.POINTER(Result)]),
    "lv_community_level_info": (C.c_int, [_vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                                          C.POINTER(C.c_int32)]),
    "lv_community_get_partition": (C.c_int, [_vp, _vp, C.c_size_t]),
    "lv_community_set_partition": (C.c_int, [_vp, _vp, C.c_size_t]),
    "lv_community_copy_graph": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "lv_best_moves": (C.c_int, [_vp, C.c_int, C.POINTER(Params), _vp, _vp,
                                C.POINTER(C.c_uint64)]),
    "lv_community_modularity_parts": (C.c_int, [_vp, C.POINTER(C.c_int64),
                                                C.POINTER(C.c_uint64), C.POINTER(C.c_double)]),
    "lv_modularity_parts_noalloc": (C.c_int, [C.c_uint32, C.c_uint64, _vp, _vp, _vp, _vp, _vp,
                                              C.c_double, _vp, C.POINTER(C.c_int64),
                                              C.POINTER(C.c_uint64), C.POINTER(C.c_double)]),
    "lv_community_get_tiles": (C.c_int, [_vp, _vp]),
    "lv_community_get_buckets": (C.c_int, [_vp, _vp]),
    "lv_community_set_tiles": (C.c_int, [_vp, _vp]),
    "lv_kernel_stats_get": (C.c_int, [_vp, C.POINTER(KernelStats)]),
    "lv_last_run_kernel_stats": (C.c_int, [C.POINTER(KernelStats)]),
    "lv_kernel_variant_name": (C.c_int, [C.c_int, C.c_int, C.c_char_p, C.c_size_t]),
    "lv_debug_mark_bytes": (C.c_int, [C.c_int32]),
    "lv_bandwidth_probe": (C.c_int, [C.c_uint64, C.c_int, C.POINTER(C.c_double)]),
    "lv_gather_probe": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.POINTER(C.c_double)]),
    "lv_batch_create": (C.c_int, [C.POINTER(_vp), C.c_uint32, _vp, C.POINTER(_vp)]),
    "lv_batch_destroy": (None, [_vp]),
    "lv_batch_info": (C.c_int, [_vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                                C.POINTER(C.c_uint64)]),
    "lv_batch_louvain": (C.c_int, [_vp, C.POINTER(Params), C.POINTER(BatchResult)]),
    "lv_community_load_edge": (C.c_int, [C.c_char_p, C.c_int, C.c_int, _vp, C.POINTER(_vp),
                                         C.POINTER(IngestInfo)]),
    "lv_graph_load_edge_gpu": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.POINTER(_vp),
                                         C.POINTER(IngestInfo)]),
    "lv_write_edge": (C.c_int, [C.c_char_p, _vp, _vp, _vp, C.c_uint64]),
    "lv_server_create": (C.c_int, [C.POINTER(Params), C.c_uint64, _vp, C.POINTER(_vp)]),
    "lv_server_destroy": (None, [_vp]),
    "lv_server_submit": (C.c_int, [_vp, _vp, C.POINTER(C.c_uint64)]),
    "lv_server_step": (C.c_int, [_vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_uint32)]),
    "lv_server_take_finished": (C.c_int, [_vp, _vp, C.c_size_t, C.POINTER(C.c_size_t)]),
    "lv_server_result": (C.c_int, [_vp, C.c_uint64, C.POINTER(JobResult), _vp, C.c_size_t,
                                   C.c_int]),
    "lv_server_footprint": (C.c_int, [_vp, C.POINTER(C.c_uint64)]),
    "lv_server_stats_get": (C.c_int, [_vp, C.POINTER(ServerStats)]),
    "lv_gen_rmat": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint64,
                              _vp, _vp, C.c_uint64, C.POINTER(C.c_uint64),
                              C.POINTER(C.c_uint32)]),
    "lv_gen_ba": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint64, _vp, _vp, C.c_uint64,
                            C.POINTER(C.c_uint64)]),
    "lv_gen_planted": (C.c_int, [C.c_uint32, C.c_double, C.c_double, C.c_uint32, C.c_uint32,
                                 C.c_double, C.c_uint32, C.c_uint32, C.c_double, C.c_uint32,
                                 C.c_uint32, C.c_uint64, _vp, _vp, _vp, C.c_uint64,
                                 C.POINTER(C.c_uint64), _vp]),
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `make -C {_HERE}` (hipcc, gfx950)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def _check(rc):
    if rc != 0:
        raise LouvainError(rc, "louvain_amd: " + lib.lv_status_string(rc).decode())


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def device_count():
    n = C.c_int()
    _check(lib.lv_device_count(C.byref(n)))
    return n.value


def params(**kw):
    p = Params()
    lib.lv_default_params(C.byref(p))
    names = {f[0] for f in Params._fields_}
    for k, v in kw.items():
        if k not in names:
            raise TypeError(f"unknown lv_params field {k!r}")
        if k == "mode" and isinstance(v, str):
            v = MODES[v]
        if k == "bucket_max_deg":
            v = (C.c_uint32 * 8)(*v)
        setattr(p, k, v)
    return p


class Graph:
    """GraphHost: host CSR built with the reference's .edge semantics."""

    def __init__(self, handle):
        self.h = handle

    def __del__(self):
        if getattr(self, "h", None):
            lib.lv_graph_free(self.h)
            self.h = None

    @classmethod
    def load_edge(cls, path, weighted=False, skip=0):
        h = C.c_void_p()
        _check(lib.lv_graph_load_edge(str(path).encode(), int(weighted), int(skip), C.byref(h)))
        return cls(h)

    @classmethod
    def load_edge_gpu(cls, path, weighted=False, skip=0):
        """The .edge reader run on the GPU; returns (graph, ingest info)."""
        h = C.c_void_p()
        info = IngestInfo()
        _check(lib.lv_graph_load_edge_gpu(str(path).encode(), int(weighted), int(skip),
                                          C.byref(h), C.byref(info)))
        return cls(h), _rec(info)

    @classmethod
    def from_edges(cls, a, b, w=None):
        a = np.ascontiguousarray(a, dtype=np.int64)
        b = np.ascontiguousarray(b, dtype=np.int64)
        w = None if w is None else np.ascontiguousarray(w, dtype=np.uint32)
        h = C.c_void_p()
        _check(lib.lv_graph_from_edges(len(a), _p(a), _p(b), _p(w), C.byref(h)))
        return cls(h)

    @classmethod
    def from_csr(cls, row, col, w=None, total_weight=0.0):
        row = np.ascontiguousarray(row, dtype=np.uint64)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        w = None if w is None else np.ascontiguousarray(w, dtype=np.uint32)
        h = C.c_void_p()
        _check(lib.lv_graph_from_csr(len(row) - 1, len(col), _p(row), _p(col), _p(w),
                                     float(total_weight), C.byref(h)))
        return cls(h)

    def info(self):
        n = C.c_uint32(); E = C.c_uint64(); tw = C.c_double(); mi = C.c_int64()
        _check(lib.lv_graph_info(self.h, C.byref(n), C.byref(E), C.byref(tw), C.byref(mi)))
        return {"n": n.value, "E": E.value, "total_weight": tw.value, "min_id": mi.value}

    @property
    def n(self):
        return self.info()["n"]

    @property
    def E(self):
        return self.info()["E"]

    def csr(self):
        i = self.info()
        row = np.empty(i["n"] + 1, np.uint64); col = np.empty(i["E"], np.uint32)
        w = np.empty(i["E"], np.uint32)
        _check(lib.lv_graph_copy_csr(self.h, _p(row), _p(col), _p(w)))
        return row, col, w

    def louvain(self, prm=None, sweeps_cap=100000, levels_cap=1000, hash_trace=False,
                best=True):
        """ModularityAlgorithms::Louvain: upload, level loop, result."""
        return _run_result(lambda r: lib.lv_louvain(self.h, C.byref(prm or params()), r),
                           self.n, sweeps_cap, levels_cap, hash_trace, best)


class Community:
    """Device-resident partition state of one level (Community, community.h)."""

    def __init__(self, graph, stream=None):
        h = C.c_void_p()
        if graph is not None:
            _check(lib.lv_community_create(graph.h, stream, C.byref(h)))
        self.h = h

    @classmethod
    def load_edge(cls, path, weighted=False, skip=0, stream=None):
        """A community straight from an .edge file parsed on the GPU (no host
        CSR); returns (community, ingest info)."""
        c = cls(None)
        info = IngestInfo()
        _check(lib.lv_community_load_edge(str(path).encode(), int(weighted), int(skip), stream,
                                          C.byref(c.h), C.byref(info)))
        return c, _rec(info)

    def close(self):
        if getattr(self, "h", None):
            lib.lv_community_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def optimization_phase(self, prm=None):
        _check(lib.lv_optimization_phase(self.h, C.byref(prm or params())))

    def aggregation_phase(self, aggregate, prm=None):
        _check(lib.lv_aggregation_phase(self.h, int(bool(aggregate)), C.byref(prm or params())))

    def refine_phase(self):
        _check(lib.lv_refine_phase(self.h))

    @property
    def modularity(self):
        return lib.lv_modularity(self.h)

    def level_info(self):
        n = C.c_uint32(); E = C.c_uint64(); lvl = C.c_int32()
        _check(lib.lv_community_level_info(self.h, C.byref(n), C.byref(E), C.byref(lvl)))
        return n.value, E.value, lvl.value

    def best_communities(self, n0):
        out = np.empty(n0, np.uint32)
        nc = C.c_uint32()
        _check(lib.lv_best_communities(self.h, _p(out), n0, C.byref(nc)))
        return out, nc.value

    def partition(self):
        n, _, _ = self.level_info()
        out = np.empty(n, np.uint32)
        _check(lib.lv_community_get_partition(self.h, _p(out), n))
        return out

    def set_partition(self, comm):
        comm = np.ascontiguousarray(comm, dtype=np.uint32)
        _check(lib.lv_community_set_partition(self.h, _p(comm), len(comm)))

    def graph_csr(self):
        n, E, _ = self.level_info()
        row = np.empty(n + 1, np.uint64); col = np.empty(E, np.uint32)
        w = np.empty(E, np.uint32); k = np.empty(n, np.uint32)
        _check(lib.lv_community_copy_graph(self.h, _p(row), _p(col), _p(w), _p(k)))
        return row, col, w, k

    def best_moves(self, kind, prm=None, active=None):
        n, _, _ = self.level_info()
        tgt = np.empty(n, np.uint32)
        act = None if active is None else np.ascontiguousarray(active, dtype=np.uint8)
        ea = C.c_uint64()
        _check(lib.lv_best_moves(self.h, int(kind), C.byref(prm or params()), _p(act), _p(tgt),
                                 C.byref(ea)))
        return tgt, ea.value

    def buckets(self):
        """Degree bounds of the best-move buckets in use (lv_community_get_buckets)."""
        out = np.zeros(8, np.uint32)
        _check(lib.lv_community_get_buckets(self.h, _p(out)))
        return out.tolist()

    def gather_probe(self, width=8, reps=5):
        """Gathers per second of arr[col[j]] over this community's input graph
        (lv_gather_probe): the ceiling of the best-move pass's dependent access."""
        g = C.c_double()
        _check(lib.lv_gather_probe(self.h, int(width), int(reps), C.byref(g)))
        return g.value

    def modularity_parts(self):
        si = C.c_int64(); s = C.c_uint64(); q = C.c_double()
        _check(lib.lv_community_modularity_parts(self.h, C.byref(si), C.byref(s), C.byref(q)))
        return q.value, si.value, s.value

    def reset(self):
        _check(lib.lv_community_reset(self.h))

    def refine(self, n0, prm=None, sweeps_cap=100000, hash_trace=False, best=True):
        """The opt-in refinement after a run (n0 = vertices of the level-0 graph)."""
        return _run_result(lambda r: lib.lv_refine(self.h, C.byref(prm or params()), r),
                           n0, sweeps_cap, 1, hash_trace, best)

    def run(self, prm=None, sweeps_cap=100000, levels_cap=1000, hash_trace=False, best=True,
            n0=None):
        """Level loop on the resident state (call reset() first to start at level 0)."""
        n0 = n0 if n0 is not None else self.level_info()[0]
        return _run_result(lambda r: lib.lv_community_run(self.h, C.byref(prm or params()), r),
                           n0, sweeps_cap, levels_cap, hash_trace, best)

    def tiles(self):
        t = np.zeros(16, np.uint8)
        _check(lib.lv_community_get_tiles(self.h, _p(t)))
        return t

    def set_tiles(self, sel):
        t = np.ascontiguousarray(sel, dtype=np.uint8)
        assert t.shape == (16,)
        _check(lib.lv_community_set_tiles(self.h, _p(t)))

    def kernel_stats(self):
        k = KernelStats()
        _check(lib.lv_kernel_stats_get(self.h, C.byref(k)))
        return _rec(k)


class Batch:
    """Many graphs run together as one disjoint union (lv_batch_*)."""

    def __init__(self, graphs, stream=None):
        self.graphs = list(graphs)
        arr = (_vp * len(self.graphs))(*[g.h for g in self.graphs])
        h = _vp()
        _check(lib.lv_batch_create(arr, len(self.graphs), stream, C.byref(h)))
        self.h = h
        self.sizes = [g.n for g in self.graphs]

    def close(self):
        if getattr(self, "h", None):
            lib.lv_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def info(self):
        G = C.c_uint32(); n = C.c_uint64(); E = C.c_uint64()
        _check(lib.lv_batch_info(self.h, C.byref(G), C.byref(n), C.byref(E)))
        return G.value, n.value, E.value

    def louvain(self, prm=None, best=True):
        G = len(self.graphs)
        q = np.empty(G, np.float64); nc = np.empty(G, np.uint32)
        lev = np.empty(G, np.int32); sw = np.empty(G, np.int32)
        out = np.empty(sum(self.sizes), np.uint32) if best else None
        r = BatchResult()
        r.modularity = q.ctypes.data_as(C.POINTER(C.c_double))
        r.n_of_best_communities = nc.ctypes.data_as(C.POINTER(C.c_uint32))
        r.levels = lev.ctypes.data_as(C.POINTER(C.c_int32))
        r.sweeps = sw.ctypes.data_as(C.POINTER(C.c_int32))
        if best:
            r.best_communities = out.ctypes.data_as(C.POINTER(C.c_uint32))
        _check(lib.lv_batch_louvain(self.h, C.byref(prm or params()), C.byref(r)))
        bests = None
        if best:
            offs = np.concatenate([[0], np.cumsum(self.sizes)])
            bests = [out[offs[i]:offs[i + 1]] for i in range(G)]
        return {"modularity": q, "n_comm": nc, "levels": lev, "sweeps": sw, "best": bests,
                "union_levels": r.union_levels, "union_sweeps": r.union_sweeps,
                "seconds": r.seconds, "traversed_entries": r.traversed_entries}


class Server:
    """Served Louvain with continuous batching (lv_server_*).  Submitted
    graphs are referenced until their job is retired."""

    def __init__(self, prm=None, hbm_budget=200 << 30, stream=None):
        h = _vp()
        _check(lib.lv_server_create(C.byref(prm or params()), int(hbm_budget), stream, C.byref(h)))
        self.h = h
        self._graphs = {}

    def close(self):
        if getattr(self, "h", None):
            lib.lv_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def submit(self, graph):
        jid = C.c_uint64()
        _check(lib.lv_server_submit(self.h, graph.h, C.byref(jid)))
        self._graphs[jid.value] = graph
        return jid.value

    def step(self):
        """One union level; returns (active, queued, retired)."""
        a = C.c_uint32(); q = C.c_uint32(); r = C.c_uint32()
        _check(lib.lv_server_step(self.h, C.byref(a), C.byref(q), C.byref(r)))
        return a.value, q.value, r.value

    def take_finished(self, cap=4096):
        ids = np.empty(cap, np.uint64)
        k = C.c_size_t()
        _check(lib.lv_server_take_finished(self.h, _p(ids), cap, C.byref(k)))
        return [int(x) for x in ids[:k.value]]

    def stats(self):
        st = ServerStats()
        _check(lib.lv_server_stats_get(self.h, C.byref(st)))
        return _rec(st)

    def result(self, job_id, best=True, release=True):
        g = self._graphs.get(job_id)
        n = g.n if (g is not None and best) else 0
        out = np.empty(max(n, 1), np.uint32) if best else None
        r = JobResult()
        _check(lib.lv_server_result(self.h, job_id, C.byref(r),
                                    _p(out) if best else None, n, int(release)))
        if release:
            self._graphs.pop(job_id, None)
        return {"status": r.status, "modularity": r.modularity,
                "n_comm": r.n_of_best_communities, "n": r.n, "levels": r.levels,
                "sweeps": r.sweeps, "seconds_queued": r.seconds_queued,
                "seconds_total": r.seconds_total, "best": out[:r.n] if best else None}


def write_edge(path, a, b, w=None):
    a = np.ascontiguousarray(a, dtype=np.int64)
    b = np.ascontiguousarray(b, dtype=np.int64)
    w = None if w is None else np.ascontiguousarray(w, dtype=np.uint32)
    _check(lib.lv_write_edge(str(path).encode(), _p(a), _p(b), _p(w), len(a)))


def footprint(graph):
    b = C.c_uint64()
    _check(lib.lv_server_footprint(graph.h, C.byref(b)))
    return b.value


_rec_bufs = threading.local()


def _record_buffers(sweeps_cap, levels_cap):
    """Per-thread record arrays, grown on demand and reused: a zeroed array
    of 100,000 sweep records costs about a millisecond to allocate, as much as
    a whole small run.  The records are copied out before the next call."""
    sw, lv = getattr(_rec_bufs, "sw", None), getattr(_rec_bufs, "lv", None)
    if sw is None or len(sw) < sweeps_cap:
        sw = _rec_bufs.sw = (SweepRec * sweeps_cap)()
    if lv is None or len(lv) < levels_cap:
        lv = _rec_bufs.lv = (LevelRec * levels_cap)()
    return sw, lv


def _run_result(call, n, sweeps_cap, levels_cap, hash_trace, best):
    out = np.empty(n, np.uint32) if best else None
    sw, lv = _record_buffers(sweeps_cap, levels_cap)
    r = Result()
    r.hash_trace = int(hash_trace)
    if best:
        r.best_communities = out.ctypes.data_as(C.POINTER(C.c_uint32))
    r.sweeps = C.cast(sw, C.POINTER(SweepRec)); r.sweeps_cap = sweeps_cap
    r.level_recs = C.cast(lv, C.POINTER(LevelRec)); r.levels_cap = levels_cap
    _check(call(C.byref(r)))
    sweeps = [_rec(sw[i]) for i in range(min(r.total_sweeps, sweeps_cap))]
    levels = [_rec(lv[i]) for i in range(min(r.levels, levels_cap))]
    return {"modularity": r.modularity, "n_comm": r.n_of_best_communities,
            "levels": r.levels, "total_sweeps": r.total_sweeps, "best": out,
            "sweeps": sweeps, "level_recs": levels, "seconds": r.seconds,
            "traversed_entries": r.traversed_entries}


def last_run_kernel_stats():
    k = KernelStats()
    _check(lib.lv_last_run_kernel_stats(C.byref(k)))
    return _rec(k)


def _rec(s):
    out = {}
    for name, _ in s._fields_:
        if name.startswith("_"):
            continue
        v = getattr(s, name)
        out[name] = list(v) if isinstance(v, C.Array) else v
    return out


def debug_mark_bytes(on):
    """Test hook: commits keep their marks in the byte array (the path of
    graphs with 2^31 or more vertices) while `on`; process-wide."""
    _check(lib.lv_debug_mark_bytes(1 if on else 0))


def bandwidth_probe(nbytes=1 << 30, reps=5):
    """GB/s (read + write) of the plain device copy kernel k_stream_copy."""
    g = C.c_double()
    _check(lib.lv_bandwidth_probe(int(nbytes), int(reps), C.byref(g)))
    return g.value


def kernel_variant_name(variant, gain=GAIN_EXACT):
    buf = C.create_string_buffer(96)
    _check(lib.lv_kernel_variant_name(int(variant), int(gain), buf, 96))
    return buf.value.decode()


def gen_rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=2):
    """R-MAT edge list (symmetrised, no self-loops or duplicates, dense ids)."""
    m = C.c_uint64(); n = C.c_uint32()
    _check(lib.lv_gen_rmat(scale, edge_factor, a, b, c, seed, None, None, 0, C.byref(m),
                           C.byref(n)))
    ea = np.empty(m.value, np.int64); eb = np.empty(m.value, np.int64)
    _check(lib.lv_gen_rmat(scale, edge_factor, a, b, c, seed, _p(ea), _p(eb), m.value,
                           C.byref(m), C.byref(n)))
    return ea, eb, n.value


def gen_ba(n, m_attach, seed=1):
    m = C.c_uint64()
    _check(lib.lv_gen_ba(n, m_attach, seed, None, None, 0, C.byref(m)))
    ea = np.empty(m.value, np.int64); eb = np.empty(m.value, np.int64)
    _check(lib.lv_gen_ba(n, m_attach, seed, _p(ea), _p(eb), m.value, C.byref(m)))
    return ea, eb


def gen_planted(n=1_000_000, deg_exp=2.5, mean_deg=20.0, max_deg=5000, n_comm=2000,
                size_exp=1.5, size_lo=50, size_hi=5000, mu=0.3, w_lo=1, w_hi=10, seed=3):
    m = C.c_uint64()
    args = (n, deg_exp, mean_deg, max_deg, n_comm, size_exp, size_lo, size_hi, mu, w_lo, w_hi,
            seed)
    _check(lib.lv_gen_planted(*args, None, None, None, 0, C.byref(m), None))
    ea = np.empty(m.value, np.int64); eb = np.empty(m.value, np.int64)
    ew = np.empty(m.value, np.uint32); planted = np.empty(n, np.uint32)
    _check(lib.lv_gen_planted(*args, _p(ea), _p(eb), _p(ew), m.value, C.byref(m),
                              _p(planted)))
    nv = int(max(ea.max(), eb.max()) + 1) if len(ea) else 0
    return ea, eb, ew, planted[:nv]
