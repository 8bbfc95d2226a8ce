"""lv_modularity_parts_noalloc (the entry point that allocates nothing)
against lv_community_modularity_parts on the same graphs and partitions.

The inputs live in device memory the test owns (hipMalloc through the HIP
runtime the library itself links, libamdhip64.so.7); two calls in
a row, the second on a smaller graph, must each give the existing entry
point's exact sigma_in and sum of tot^2 (and so the same Q).  The argument
checks run on the CPU.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import louvain_amd as L

MAX_N = 1 << 24


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def call(n, E, row, col, w, k, comm, tw):
    si = C.c_int64(); s = C.c_uint64(); q = C.c_double()
    rc = L.lib.lv_modularity_parts_noalloc(n, E, row, col, w, k, comm, tw, None, C.byref(si),
                                           C.byref(s), C.byref(q))
    return rc, (q.value, si.value, s.value)


def test_noalloc_argument_checks():
    # the checks come before any device access, so they hold without a GPU
    assert call(MAX_N + 1, 0, 1, 1, None, 1, 1, 1.0)[0] == 6          # LV_ERR_RANGE
    assert call(0, 0, 1, 1, None, 1, 1, 1.0)[0] == 1                  # LV_ERR_ARG
    assert call(4, 4, None, 1, None, 1, 1, 1.0)[0] == 1
    assert call(4, 4, 1, None, None, 1, 1, 1.0)[0] == 1
    assert call(4, 4, 1, 1, None, 1, None, 1.0)[0] == 1
    assert call(4, 4, 1, 1, None, 1, 1, 0.0)[0] == 1


class DeviceBuffers:
    """Device copies of host arrays, made and freed with the HIP runtime."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(a.nbytes, 16)) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # H2D
        return p.value

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _device_inputs(buf, c, comm):
    row, col, w, k = c.graph_csr()
    t = {"row": buf.put(row), "col": buf.put(col), "k": buf.put(k),
         "comm": buf.put(np.ascontiguousarray(comm, np.uint32)),
         "w": None if np.all(w == 1) else buf.put(w)}
    return len(row) - 1, len(col), t


@pytest.mark.gpu
def test_noalloc_matches_community_parts(fixture_path):
    if L.device_count() < 1:
        pytest.skip("no GPU visible")
    buf = DeviceBuffers()
    rng = np.random.default_rng(11)
    # first call: a weighted R-MAT graph (2^16 vertices), random partition;
    # second call: the smaller bundled fixture, unit weights, another partition
    ea, eb, n_big = L.gen_rmat(16, 16, 0.57, 0.19, 0.19, seed=7)
    wts = rng.integers(1, 9, len(ea)).astype(np.uint32)
    cases = [("rmat16 weighted", L.Graph.from_edges(ea, eb, wts), 300),
             ("fixture", L.Graph.load_edge(fixture_path), 40),
             ("rmat16 weighted, identity", L.Graph.from_edges(ea, eb, wts), 0)]
    for name, g, ncomm in cases:
        c = L.Community(g)
        n = g.n
        comm = (rng.integers(0, ncomm, n) if ncomm else np.arange(n)).astype(np.uint32)
        c.set_partition(comm)
        want = c.modularity_parts()
        nn, E, t = _device_inputs(buf, c, comm)
        assert nn == n
        tw = g.info()["total_weight"]
        say(f"no-alloc: {name}, n={n}, E={E}")
        rc, got = call(n, E, t["row"], t["col"], t["w"], t["k"], t["comm"], tw)
        assert rc == 0, rc
        assert got == want, (name, got, want)
        c.close()
    # a community id outside [0, n) is reported, not written
    g = L.Graph.load_edge(fixture_path)
    c = L.Community(g)
    bad = np.zeros(g.n, np.uint32)
    bad[17] = g.n
    _, E, t = _device_inputs(buf, c, bad)
    rc, _ = call(g.n, E, t["row"], t["col"], None, t["k"], t["comm"], g.info()["total_weight"])
    assert rc == 1                                     # LV_ERR_ARG
    c.close()
    buf.free()
