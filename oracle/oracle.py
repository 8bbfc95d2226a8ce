"""ctypes wrapper of the CPU oracle (oracle/louvain_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the timed CPU
baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg import this module; the product path never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblv_oracle.so")

HASH, SORT, ADAPTIVE_SPEED, ADAPTIVE_MEMORY = 0, 1, 2, 3
SWEEP_FAST, SWEEP_SORT, SWEEP_HASH = 0, 1, 2
GAIN_EXACT, GAIN_REF_F32 = 0, 1
TIE_LOWEST, TIE_REFERENCE = 0, 1
ERR = {0: "ok", 1: "bad argument", 2: "out of memory", 3: "parse error",
       4: "io error", 5: "ids not dense", 6: "out of range"}


class Params(C.Structure):
    _fields_ = [("mode", C.c_int), ("gain", C.c_int), ("tie", C.c_int),
                ("chunk_emulation", C.c_int), ("chunk_entries", C.c_uint64),
                ("sweep_threshold", C.c_double), ("level_threshold", C.c_double),
                ("max_sweeps", C.c_int64), ("max_total_sweeps", C.c_int64),
                ("max_levels", C.c_int32), ("refine", C.c_int32)]


class SweepRec(C.Structure):
    _fields_ = [("level", C.c_int32), ("sweep", C.c_int32), ("kind", C.c_int32),
                ("_pad", C.c_int32), ("n", C.c_uint32), ("n_active", C.c_uint32),
                ("n_moved", C.c_uint32), ("n_considered", C.c_uint32),
                ("E", C.c_uint64), ("e_active", C.c_uint64), ("sigma_in", C.c_int64),
                ("sum_tot_sq", C.c_uint64), ("q", C.c_double), ("comm_hash", C.c_uint64)]


class LevelRec(C.Structure):
    _fields_ = [("level", C.c_int32), ("sweeps", C.c_int32), ("n", C.c_uint32),
                ("n_comm", C.c_uint32), ("E", C.c_uint64), ("q_start", C.c_double),
                ("q_end", C.c_double), ("best_hash", C.c_uint64)]


class Result(C.Structure):
    _fields_ = [("modularity", C.c_double), ("n_of_best_communities", C.c_uint32),
                ("levels", C.c_int32), ("total_sweeps", C.c_int32), ("_pad", C.c_int32),
                ("best_communities", C.POINTER(C.c_uint32)),
                ("sweeps", C.POINTER(SweepRec)), ("sweeps_cap", C.c_int32),
                ("level_recs", C.POINTER(LevelRec)), ("levels_cap", C.c_int32)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        L.lvo_default_params.argtypes = [C.POINTER(Params)]
        L.lvo_graph_from_edges.argtypes = [C.c_uint64, vp, vp, vp, C.POINTER(vp)]
        L.lvo_graph_load_edge.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(vp)]
        L.lvo_graph_from_csr.argtypes = [C.c_uint32, C.c_uint64, vp, vp, vp, C.c_double, C.POINTER(vp)]
        L.lvo_graph_free.argtypes = [vp]
        L.lvo_graph_n.argtypes = [vp]; L.lvo_graph_n.restype = C.c_uint32
        L.lvo_graph_E.argtypes = [vp]; L.lvo_graph_E.restype = C.c_uint64
        L.lvo_graph_total_weight.argtypes = [vp]; L.lvo_graph_total_weight.restype = C.c_double
        L.lvo_graph_min_id.argtypes = [vp]; L.lvo_graph_min_id.restype = C.c_int64
        L.lvo_graph_copy.argtypes = [vp, vp, vp, vp, vp]
        L.lvo_modularity.argtypes = [vp, vp, vp, vp]; L.lvo_modularity.restype = C.c_double
        L.lvo_best_moves.argtypes = [vp, vp, vp, vp, C.c_int, C.POINTER(Params), vp, vp, vp]
        L.lvo_best_moves.restype = C.c_int64
        L.lvo_aggregate.argtypes = [vp, vp, vp, vp, C.POINTER(vp)]
        L.lvo_louvain.argtypes = [vp, C.POINTER(Params), C.POINTER(Result)]
        L.lvo_fnv1a_u32.argtypes = [vp, C.c_uint64]; L.lvo_fnv1a_u32.restype = C.c_uint64
        L.lvo_dq_ref_f32.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_int64, C.c_int64,
                                     C.c_double]
        L.lvo_dq_ref_f32.restype = C.c_float
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def params(**kw):
    p = Params()
    lib().lvo_default_params(C.byref(p))
    names = {f[0] for f in Params._fields_}
    for k, v in kw.items():
        if k not in names:
            raise TypeError(f"unknown lvo_params field {k!r}")
        setattr(p, k, v)
    return p


class Graph:
    """Level graph in CSR form (row[n+1] u64, col[E] u32, w[E] u32, k[n] u64)."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.lvo_graph_free(self.h)
            self.h = None

    @staticmethod
    def _check(rc):
        if rc != 0:
            raise RuntimeError("oracle: " + ERR.get(rc, str(rc)))

    @classmethod
    def from_edges(cls, a, b, w=None):
        a = np.ascontiguousarray(a, dtype=np.int64)
        b = np.ascontiguousarray(b, dtype=np.int64)
        wa = None if w is None else np.ascontiguousarray(w, dtype=np.uint32)
        h = C.c_void_p()
        cls._check(lib().lvo_graph_from_edges(len(a), _p(a), _p(b), _p(wa), C.byref(h)))
        return cls(h)

    @classmethod
    def load_edge(cls, path, weighted=False, skip=0):
        h = C.c_void_p()
        cls._check(lib().lvo_graph_load_edge(path.encode(), int(weighted), int(skip), C.byref(h)))
        return cls(h)

    @classmethod
    def from_csr(cls, row, col, w=None, total_weight=0.0):
        row = np.ascontiguousarray(row, dtype=np.uint64)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        wa = None if w is None else np.ascontiguousarray(w, dtype=np.uint32)
        h = C.c_void_p()
        cls._check(lib().lvo_graph_from_csr(len(row) - 1, len(col), _p(row), _p(col), _p(wa),
                                            float(total_weight), C.byref(h)))
        return cls(h)

    @property
    def n(self):
        return lib().lvo_graph_n(self.h)

    @property
    def E(self):
        return lib().lvo_graph_E(self.h)

    @property
    def total_weight(self):
        return lib().lvo_graph_total_weight(self.h)

    @property
    def min_id(self):
        return lib().lvo_graph_min_id(self.h)

    def arrays(self):
        n, E = self.n, self.E
        row = np.empty(n + 1, np.uint64); col = np.empty(E, np.uint32)
        w = np.empty(E, np.uint32); k = np.empty(n, np.uint64)
        lib().lvo_graph_copy(self.h, _p(row), _p(col), _p(w), _p(k))
        return row, col, w, k

    def modularity(self, comm):
        comm = np.ascontiguousarray(comm, dtype=np.uint32)
        si = C.c_int64(); s = C.c_uint64()
        q = lib().lvo_modularity(self.h, _p(comm), C.byref(si), C.byref(s))
        return q, si.value, s.value

    def best_moves(self, comm, active, kind, prm=None, tot=None):
        comm = np.ascontiguousarray(comm, dtype=np.uint32)
        active = np.ascontiguousarray(active, dtype=np.uint8)
        if tot is None:
            _, _, _, k = self.arrays()
            tot = np.bincount(comm, weights=k.astype(np.float64), minlength=self.n).astype(np.int64)
        tot = np.ascontiguousarray(tot, dtype=np.int64)
        prm = prm or params()
        target = np.empty(self.n, np.uint32)
        ea = C.c_uint64(); nc = C.c_uint32()
        moved = lib().lvo_best_moves(self.h, _p(comm), _p(tot), _p(active), int(kind),
                                     C.byref(prm), _p(target), C.byref(ea), C.byref(nc))
        return target, int(moved), ea.value, nc.value

    def aggregate(self, comm):
        comm = np.ascontiguousarray(comm, dtype=np.uint32)
        new_id = np.empty(self.n, np.uint32)
        nn = C.c_uint32(); h = C.c_void_p()
        self._check(lib().lvo_aggregate(self.h, _p(comm), _p(new_id), C.byref(nn), C.byref(h)))
        return new_id, nn.value, Graph(h)

    def louvain(self, prm=None, sweeps_cap=100000, levels_cap=1000):
        prm = prm or params()
        best = np.empty(self.n, np.uint32)
        sw = (SweepRec * sweeps_cap)()
        lv = (LevelRec * levels_cap)()
        r = Result()
        r.best_communities = best.ctypes.data_as(C.POINTER(C.c_uint32))
        r.sweeps = C.cast(sw, C.POINTER(SweepRec)); r.sweeps_cap = sweeps_cap
        r.level_recs = C.cast(lv, C.POINTER(LevelRec)); r.levels_cap = levels_cap
        self._check(lib().lvo_louvain(self.h, C.byref(prm), C.byref(r)))
        sweeps = [_rec(sw[i]) for i in range(min(r.total_sweeps, sweeps_cap))]
        levels = [_rec(lv[i]) for i in range(levels_cap) if lv[i].sweeps > 0]
        return {"modularity": r.modularity, "n_comm": r.n_of_best_communities,
                "levels": r.levels, "total_sweeps": r.total_sweeps,
                "best": best, "sweeps": sweeps, "level_recs": levels}


def _rec(s):
    return {name: getattr(s, name) for name, _ in s._fields_ if not name.startswith("_")}


def fnv1a_u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return lib().lvo_fnv1a_u32(_p(a), len(a))
