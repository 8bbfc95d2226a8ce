"""CPU tests of the C ABI library: it loads, exports every declared entry
point, and its host-side parts (.edge parsing, CSR construction, generators,
error mapping) agree with the oracle.  No kernel is launched here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import louvain_amd as L
from oracle import oracle as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(REPO, "include", "lv_api.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set(re.findall(r"\b(lv_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 30
    for name in names:
        assert hasattr(L.lib, name), name
    assert set(names) == set(L.SIGNATURES), set(names) ^ set(L.SIGNATURES)


def test_default_params_mirror_reference_constants():
    p = L.params()
    assert (p.mode, p.gain, p.tie, p.chunk_emulation) == (L.HASH, L.GAIN_EXACT, L.TIE_LOWEST, 0)
    assert p.chunk_entries == 50_000_000          # STEP_ROUND (constants.h:18)
    assert p.sweep_threshold == p.level_threshold == 0.001   # constants.h:15
    assert p.max_sweeps == 1_000_000_000          # EARLY_STOP_LIMIT (constants.h:16)


def csr_equal(g, og):
    row, col, w = g.csr()
    r2, c2, w2, _ = og.arrays()
    return np.array_equal(row, r2) and np.array_equal(col, c2) and np.array_equal(w, w2)


def test_fixture_csr_matches_oracle(fixture_path):
    g = L.Graph.load_edge(fixture_path)
    og = O.Graph.load_edge(fixture_path)
    assert g.info() == {"n": 10000, "E": 39992, "total_weight": 19996.0, "min_id": 0}
    assert csr_equal(g, og)


@pytest.mark.parametrize("seed", range(5))
def test_random_edge_lists_match_oracle(seed):
    rng = np.random.default_rng(seed)
    n = 300
    m = 2000
    off = int(rng.integers(-50, 50))
    a = np.concatenate([rng.integers(0, n, m), np.arange(n), rng.integers(0, n, 7)]) + off
    b = np.concatenate([rng.integers(0, n, m), (np.arange(n) + 1) % n,
                        rng.integers(0, n, 7)]) + off
    b[-3:] = a[-3:]                                   # self-loops
    w = rng.integers(1, 20, len(a)).astype(np.uint32) if seed % 2 else None
    g = L.Graph.from_edges(a, b, w)
    og = O.Graph.from_edges(a, b, w)
    assert csr_equal(g, og)
    i = g.info()
    assert i["min_id"] == og.min_id and i["total_weight"] == og.total_weight


def test_weighted_file_and_errors(tmp_path):
    p = tmp_path / "w.edge"
    p.write_text("c comment\n0 1 4\n1 2 3\n1 0 9\n2 2 5\n")
    g = L.Graph.load_edge(str(p), weighted=True, skip=1)
    og = O.Graph.load_edge(str(p), weighted=True, skip=1)
    assert csr_equal(g, og)
    with pytest.raises(L.LouvainError) as e:
        L.Graph.from_edges([0, 5], [1, 6])
    assert e.value.status == 5                        # LV_ERR_NOT_DENSE
    empty = tmp_path / "e.edge"
    empty.write_text("nothing here\n")
    with pytest.raises(L.LouvainError) as e:
        L.Graph.load_edge(str(empty))
    assert e.value.status == 3                        # LV_ERR_PARSE
    with pytest.raises(L.LouvainError) as e:
        L.Graph.load_edge(str(tmp_path / "missing.edge"))
    assert e.value.status == 4                        # LV_ERR_IO


def test_community_without_device_reports_no_device(fixture_path):
    if L.device_count() > 0:
        pytest.skip("a GPU is visible")
    g = L.Graph.load_edge(fixture_path)
    with pytest.raises(L.LouvainError) as e:
        L.Community(g)
    assert e.value.status == 8                        # LV_ERR_NO_DEVICE
    with pytest.raises(L.LouvainError) as e:
        g.louvain()
    assert e.value.status == 8
    with pytest.raises(L.LouvainError) as e:
        L.Batch([g, g])
    assert e.value.status == 8
    with pytest.raises(L.LouvainError) as e:
        L.Server()
    assert e.value.status == 8
    g8 = C.c_double()
    assert L.lib.lv_gather_probe(None, 8, 1, C.byref(g8)) == 1       # LV_ERR_ARG


def test_server_footprint_and_argument_errors(fixture_path):
    import ctypes as C
    g = L.Graph.load_edge(fixture_path)
    # 48 bytes per entry + 96 per vertex + 1 MiB
    assert L.footprint(g) == 48 * 39992 + 96 * 10000 + (1 << 20)
    h = C.c_void_p()
    assert L.lib.lv_server_create(None, 0, None, C.byref(h)) == 1      # LV_ERR_ARG
    assert L.lib.lv_server_step(None, None, None, None) == 1
    assert L.lib.lv_server_submit(None, g.h, None) == 1
    assert L.lib.lv_server_result(None, 1, None, None, 0, 0) == 1


def test_batch_argument_errors():
    import ctypes as C
    h = C.c_void_p()
    assert L.lib.lv_batch_create(None, 3, None, C.byref(h)) == 1          # LV_ERR_ARG
    arr = (C.c_void_p * 2)(None, None)
    assert L.lib.lv_batch_create(arr, 2, None, C.byref(h)) == 1
    assert L.lib.lv_batch_create(arr, 0, None, C.byref(h)) == 1
    assert L.lib.lv_batch_louvain(None, None, None) == 1
    assert L.lib.lv_batch_info(None, None, None, None) == 1


def test_rmat_generator_properties():
    ea, eb, n = L.gen_rmat(12, seed=2)
    ea2, eb2, n2 = L.gen_rmat(12, seed=2)
    assert np.array_equal(ea, ea2) and np.array_equal(eb, eb2) and n == n2
    assert np.all(ea < eb)                            # normalised, no self-loops
    key = ea * (1 << 32) + eb
    assert np.all(np.diff(key) > 0)                   # sorted, no duplicates
    used = np.zeros(n, bool); used[ea] = True; used[eb] = True
    assert used.all() and ea.max() < n and eb.max() < n   # dense compacted ids
    ea3, _, _ = L.gen_rmat(12, seed=3)
    assert not np.array_equal(ea, ea3[:len(ea)]) or len(ea) != len(ea3)


def test_ba_generator_properties():
    n, m = 2000, 3
    ea, eb = L.gen_ba(n, m, seed=1)
    assert len(ea) == (n - m) * m
    pairs = set(zip(np.minimum(ea, eb).tolist(), np.maximum(ea, eb).tolist()))
    assert len(pairs) == len(ea)                      # simple graph
    deg = np.bincount(np.concatenate([ea, eb]), minlength=n)
    assert deg[m:].min() >= m and deg.max() > 10 * m   # preferential attachment hubs


def test_planted_generator_properties():
    ea, eb, ew, planted = L.gen_planted(n=20000, n_comm=40, size_lo=50, size_hi=2000,
                                        max_deg=500, seed=3)
    assert ew.min() >= 1 and ew.max() <= 10
    n = int(max(ea.max(), eb.max())) + 1
    assert len(planted) == n
    og = O.Graph.from_edges(ea, eb, ew)
    q, _, _ = og.modularity(planted)
    inside = np.mean(planted[ea] == planted[eb])
    assert 0.55 < inside < 0.85                       # mixing mu = 0.3
    assert q > 0.5


def test_ba_10000_oracle_q_matches_fp64_recompute():
    """The CPU-path workload (BA n=10,000, 2 edges per vertex, seed 1): the
    oracle's final Q equals an independent FP64 recomputation and networkx."""
    import networkx as nx
    from oracle import oracle as O
    from qcheck import modularity_fp64
    ea, eb = L.gen_ba(10000, 2, seed=1)
    assert len(ea) == 19996 and int(max(ea.max(), eb.max())) == 9999
    og = O.Graph.from_edges(ea, eb)
    r = og.louvain(O.params(sweep_threshold=1e-6, level_threshold=1e-6))
    row, col, _ = og.arrays()[:3]
    assert abs(r["modularity"] - modularity_fp64(row, col, None, r["best"])) < 1e-9
    G = nx.Graph()
    G.add_edges_from(zip(ea.tolist(), eb.tolist()))
    parts = {}
    for v, c in enumerate(r["best"]):
        parts.setdefault(int(c), set()).add(v)
    assert abs(r["modularity"] - nx.community.modularity(G, list(parts.values()))) < 1e-9


def test_params_layout_and_tile_variant_names():
    import ctypes as C
    assert C.sizeof(L.Params) == 96
    with pytest.raises(TypeError):
        L.params(no_such_field=1)
    assert L.params(autotune=1).autotune == 1
    assert list(L.params().bucket_max_deg) == [0] * 8
    assert list(L.params(bucket_max_deg=[2, 4, 8, 16, 32, 128, 512, 2048]).bucket_max_deg)[-1] == 2048
    # variant = ((bucket*2 + fast)*2 + weighted)*3 + tile
    assert L.kernel_variant_name(((0 * 2 + 0) * 2 + 0) * 3 + 1) == "k_move_table<1, 8, 0, false, true>"
    assert L.kernel_variant_name(((5 * 2 + 0) * 2 + 1) * 3 + 1) == "k_move_table<32, 388, 0, true, true>"
    assert L.kernel_variant_name(((3 * 2 + 1) * 2 + 0) * 3 + 2) == "k_move_fast<32, 0, false>"
    assert L.kernel_variant_name(((6 * 2 + 0) * 2 + 0) * 3 + 0) == "k_move_table<256, 1540, 0, false, true>"
    assert L.kernel_variant_name(((7 * 2 + 0) * 2 + 0) * 3 + 0) == "k_move_table<256, 4960, 0, false, false>"
    with pytest.raises(L.LouvainError):
        L.kernel_variant_name(((6 * 2 + 0) * 2 + 0) * 3 + 1)      # bucket 6 has one tile


def test_write_edge_round_trip(tmp_path):
    rng = np.random.default_rng(17)
    n = 500
    a = rng.integers(-50, n - 50, 4000)
    b = rng.integers(-50, n - 50, 4000)
    a = np.concatenate([a, np.arange(-50, n - 50)])
    b = np.concatenate([b, np.roll(np.arange(-50, n - 50), 1)])
    w = rng.integers(0, 9, len(a)).astype(np.uint32)
    for weighted in (False, True):
        p = tmp_path / f"g{int(weighted)}.edge"
        L.write_edge(p, a, b, w if weighted else None)
        g = L.Graph.load_edge(p, weighted=weighted)
        ref = L.Graph.from_edges(a, b, w if weighted else None)
        r1, c1, w1 = g.csr()[:3]
        r2, c2, w2 = ref.csr()[:3]
        assert np.array_equal(r1, r2) and np.array_equal(c1, c2) and np.array_equal(w1, w2)
        assert g.info()["min_id"] == -50


def test_traffic_profile_names_the_roofline_kernel():
    """bench.py reads the hardware traffic of its roofline kernel (the degree
    257-1024 bucket's default hash-sweep variant on unit weights) from the
    committed counter profile; the profile must name that kernel."""
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "profiles",
                        "k_move_table_256_1540_rmat22_traffic.json")
    t = json.load(open(path))
    assert t["kernel"] == L.kernel_variant_name(((6 * 2 + 0) * 2 + 0) * 3 + 0)
    # the KiB figures are stored rounded to 1/100 KiB
    assert abs(t["traffic_bytes_per_launch"] -
               (t["fetch_kib_per_dispatch"] * t["fetch_scale"] +
                t["write_kib_per_dispatch"] * t["write_scale"]) * 1024) < 64



@pytest.mark.parametrize("name", ["basic", "crlf_tabs", "skip", "junk_stops", "odd_tokens",
                                  "sign_after_digit", "signs", "bad_sign", "overflow_first",
                                  "overflow_later", "int_limits", "dups_loops", "weighted",
                                  "weighted_units", "weighted_partial", "long_line",
                                  "no_newline_end", "not_dense", "empty", "only_junk"])
def test_host_reader_edge_cases_match_oracle(tmp_path, name):
    """The library's host .edge reader and the oracle's agree on every edge
    case the GPU ingest test uses (same CSR, or the same error status)."""
    from test_gpu_ingest import CASES, oracle_load
    text, weighted, skip = CASES[name]
    p = tmp_path / f"{name}.edge"
    p.write_bytes(text.encode())
    oi, oa = oracle_load(p, weighted, skip)
    try:
        g = L.Graph.load_edge(p, weighted, skip)
        err = None
    except L.LouvainError as e:
        g, err = None, e.status
    if oi is None:
        assert err == oa, (err, oa)
        return
    assert err is None and g.info() == oi, (err, oi)
    for x, y in zip(oa, g.csr()[:3]):
        assert np.array_equal(x, y)


def test_from_csr_total_weight_must_be_half_the_weight_sum():
    row = np.array([0, 1, 2], np.uint64)
    col = np.array([1, 0], np.uint32)
    w = np.array([3, 3], np.uint32)
    assert L.Graph.from_csr(row, col, w).info()["total_weight"] == 3.0
    assert L.Graph.from_csr(row, col, w, total_weight=3.0).info()["total_weight"] == 3.0
    for bad in (2.5, 4.0, 1e12):
        with pytest.raises(L.LouvainError) as e:
            L.Graph.from_csr(row, col, w, total_weight=bad)
        assert e.value.status == 1                    # LV_ERR_ARG
