"""GPU .edge ingest against the oracle's reader (graph.cu:17-87 semantics).

Bar: the same CSR (rows, destinations, weights), vertex count, entries,
total weight and id shift, and the same error on inputs the reader rejects,
as the oracle's .edge reader (oracle/louvain_oracle.c, lvo_graph_load_edge)
and the library's host reader.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import louvain_amd as L
from oracle import oracle as O
from test_gpu_parity import REC_KEYS, compare_runs

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if L.device_count() < 1:
        pytest.skip("no GPU visible")


def _load(loader, path, weighted, skip):
    try:
        g = loader(path, weighted, skip)
        if isinstance(g, tuple):
            g = g[0]
        return g, None
    except L.LouvainError as e:
        return None, e.status


def oracle_load(path, weighted, skip):
    """The oracle's reader: (graph info, (row, col, w)) or (None, status)."""
    h = C.c_void_p()
    rc = O.lib().lvo_graph_load_edge(str(path).encode(), int(weighted), int(skip), C.byref(h))
    if rc != 0:
        return None, rc
    g = O.Graph(h)
    info = {"n": g.n, "E": g.E, "total_weight": g.total_weight, "min_id": g.min_id}
    return info, g.arrays()[:3]


def _same(path, weighted=False, skip=0):
    """GPU ingest == oracle reader (and == the host reader); returns the
    error status they agree on, or None."""
    oi, oa = oracle_load(path, weighted, skip)
    gd, ed = _load(L.Graph.load_edge_gpu, path, weighted, skip)
    gh, eh = _load(L.Graph.load_edge, path, weighted, skip)
    eo = None if oi is not None else oa
    assert ed == eo, f"{path}: oracle error {eo}, GPU error {ed}"
    assert eh == eo, f"{path}: oracle error {eo}, host error {eh}"
    if oi is None:
        return eo
    assert gd.info() == oi, (gd.info(), oi)
    assert gh.info() == oi, (gh.info(), oi)
    for x, y, z in zip(oa, gd.csr()[:3], gh.csr()[:3]):
        assert np.array_equal(x, y) and np.array_equal(x, z)
    return None


CASES = {
    "basic": ("1 2\n2 3\n3 1\n", False, 0),
    "crlf_tabs": ("  1\t2\r\n2 3\r\n\r\n3 1", False, 0),
    "skip": ("# header line\nfoo bar baz\n1 2\n2 3\n3 1\n", False, 2),
    "junk_stops": ("1 2\n2 3\nabc\n3 4\n", False, 0),
    "odd_tokens": ("1 2 3", False, 0),
    "sign_after_digit": ("1-2 3 4\n", False, 0),
    "signs": ("+1 +2\n-1 1\n0 2\n", False, 0),
    "bad_sign": ("- 5\n1 2\n", False, 0),
    "overflow_first": ("99999999999 1\n", False, 0),
    "overflow_later": ("1 2\n2147483648 1\n", False, 0),
    "int_limits": ("-2147483648 -2147483647\n", False, 0),
    "dups_loops": ("1 1\n1 2\n2 1\n1 1\n2 2\n", False, 0),
    "weighted": ("1 2 5\n2 3 7\n1 2 9\n3 1 -1\n2 2 3\n", True, 0),
    "weighted_units": ("1 2 1\n2 3 1\n3 1 1\n", True, 0),
    "weighted_partial": ("1 2 4\n2 3\n", True, 0),
    "long_line": (" " * 200 + "0000000000000000000000000000001" + " " * 130 + "2\n2 3\n", False, 0),
    "no_newline_end": ("5 6\n6 7", False, 0),
    "not_dense": ("1 2\n5 6\n", False, 0),
    "empty": ("", False, 0),
    "only_junk": ("hello\n", False, 0),
}


@pytest.mark.parametrize("name", list(CASES))
def test_edge_cases_match_oracle_reader(tmp_path, name):
    text, weighted, skip = CASES[name]
    p = tmp_path / f"{name}.edge"
    p.write_bytes(text.encode())
    say(f"ingest case {name}")
    _same(p, weighted, skip)


def test_missing_file(tmp_path):
    assert _same(tmp_path / "nope.edge") == 4          # LV_ERR_IO


def test_fixture_and_large_generated_files(fixture_path, tmp_path):
    assert _same(fixture_path) is None
    ea, eb, n = L.gen_rmat(18, seed=3)
    rng = np.random.default_rng(5)
    w = rng.integers(1, 11, len(ea)).astype(np.uint32)
    # shuffle, flip and repeat some edges so that dedupe and first-wins matter
    perm = rng.permutation(len(ea))
    a, b = ea[perm], eb[perm]
    flip = rng.random(len(a)) < 0.5
    a, b = np.where(flip, b, a), np.where(flip, a, b)
    a = np.concatenate([a, a[:5000]])
    b = np.concatenate([b, b[:5000]])
    w = np.concatenate([w, rng.integers(1, 11, 5000).astype(np.uint32)])
    for weighted in (False, True):
        p = tmp_path / f"rmat18_{int(weighted)}.edge"
        L.write_edge(p, a + 7, b + 7, w if weighted else None)
        say(f"ingest rmat18 weighted={weighted}: {p.stat().st_size} bytes")
        assert _same(p, weighted) is None


def test_community_from_gpu_ingest_runs_like_host(fixture_path):
    c, info = L.Community.load_edge(fixture_path)
    assert (info["n"], info["E"], info["min_id"]) == (10000, 39992, 0)
    ref = L.Graph.load_edge(fixture_path).louvain(L.params(), hash_trace=True)
    c.reset()
    compare_runs(c.run(L.params(), hash_trace=True), ref, "gpu ingest community")
