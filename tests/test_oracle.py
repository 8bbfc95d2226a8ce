"""CPU tests that pin the oracle (oracle/louvain_oracle.c).

The reference ships no tests or expected outputs, so the oracle is pinned by:
  * closed forms on the bundled fixture (singleton modularity, sums);
  * brute force: every per-vertex decision equals the argmax of the modularity
    change computed by full recomputation in exact rational arithmetic;
  * networkx's modularity on random partitions;
  * an independent numpy emulation of the reference's Thrust pipeline
    (tests/ref_pipeline.py), trace for trace;
  * textbook answers (ring of cliques) and invariants (aggregation).
"""
import itertools
from fractions import Fraction

import numpy as np
import pytest

import ref_pipeline as R
from oracle import oracle as O


def fixture_graph(path):
    return O.Graph.load_edge(path)


def test_fixture_closed_forms(fixture_path):
    g = fixture_graph(fixture_path)
    assert (g.n, g.E, g.total_weight, g.min_id) == (10000, 39992, 19996.0, 0)
    row, col, w, k = g.arrays()
    assert int(k.sum()) == 39992 and int((k.astype(np.int64) ** 2).sum()) == 800164
    q, si, s = g.modularity(np.arange(g.n))
    assert si == 0 and s == 800164
    assert q == -800164 / (4 * 19996.0 * 19996.0)
    assert abs(Fraction(q) - Fraction(-200041, 399840016)) < Fraction(1, 10**18)
    # all in one community: Q = 0
    q1, si1, s1 = g.modularity(np.zeros(g.n))
    assert si1 == 39992 and s1 == 39992 ** 2 and q1 == 0.0


def exact_q(edges_w, n, comm):
    """Newman modularity from scratch in rationals (independent formula)."""
    m2 = sum(2 * w for (_, _, w) in edges_w)
    tot = {}
    inside = Fraction(0)
    for a, b, w in edges_w:
        tot[comm[a]] = tot.get(comm[a], 0) + w
        tot[comm[b]] = tot.get(comm[b], 0) + w
        if comm[a] == comm[b]:
            inside += 2 * w
    return inside / m2 - sum(Fraction(t * t, m2 * m2) for t in tot.values())


def small_graph(rng, n, m, weighted):
    pairs = set()
    while len(pairs) < m:
        a, b = rng.integers(0, n, 2)
        if a != b:
            pairs.add((min(a, b), max(a, b)))
    for v in range(n):   # keep ids dense
        pairs.add((v, (v + 1) % n) if v + 1 < n else (0, n - 1))
    pairs = sorted(pairs)
    w = [int(rng.integers(1, 6)) if weighted else 1 for _ in pairs]
    return [(int(a), int(b), wi) for (a, b), wi in zip(pairs, w)]


@pytest.mark.parametrize("seed", range(12))
def test_bruteforce_best_moves(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(4, 11))
    edges = small_graph(rng, n, int(rng.integers(n, 2 * n + 3)), weighted=seed % 2 == 1)
    a = np.array([e[0] for e in edges]); b = np.array([e[1] for e in edges])
    w = np.array([e[2] for e in edges], np.uint32)
    g = O.Graph.from_edges(a, b, w)
    for trial in range(4):
        comm = rng.integers(0, max(1, n // (trial + 1)), n).astype(np.uint32)
        adj = {v: set() for v in range(n)}
        for x, y, _ in edges:
            adj[x].add(y); adj[y].add(x)
        q0 = exact_q(edges, n, comm)
        want_low = comm.copy()
        want_high = comm.copy()
        for v in range(n):
            cands = sorted({int(comm[x]) for x in adj[v]} - {int(comm[v])})
            gains = []
            for c in cands:
                cc = comm.copy(); cc[v] = c
                gains.append((exact_q(edges, n, cc) - q0, c))
            if not gains:
                continue
            best = max(d for d, _ in gains)
            if best > 0:
                tied = [c for d, c in gains if d == best]
                want_low[v] = min(tied)
                want_high[v] = max(tied)
        act = np.ones(n, np.uint8)
        got, _, _, _ = g.best_moves(comm, act, O.SWEEP_SORT, O.params())
        assert np.array_equal(got, want_low), (seed, trial)
        got_h, _, _, _ = g.best_moves(comm, act, O.SWEEP_HASH, O.params(tie=O.TIE_REFERENCE))
        assert np.array_equal(got_h, want_high), (seed, trial)
        got_hl, _, _, _ = g.best_moves(comm, act, O.SWEEP_HASH, O.params())
        assert np.array_equal(got_hl, want_low), (seed, trial)
    # fast sweep at the identity partition: argmax over neighbours, move iff v < c
    ident = np.arange(n, dtype=np.uint32)
    q0 = exact_q(edges, n, ident)
    got, _, _, _ = g.best_moves(ident, np.ones(n, np.uint8), O.SWEEP_FAST, O.params())
    for v in range(n):
        nb = sorted({x for x, y, _ in edges if y == v} | {y for x, y, _ in edges if x == v})
        gains = []
        for c in nb:
            cc = ident.copy(); cc[v] = c
            gains.append((exact_q(edges, n, cc) - q0, c))
        best = max(d for d, _ in gains)
        c_best = min(c for d, c in gains if d == best)
        assert got[v] == (c_best if best > 0 and v < c_best else v)


def test_modularity_matches_networkx():
    nx = pytest.importorskip("networkx")
    rng = np.random.default_rng(3)
    for trial in range(5):
        n = 200
        edges = small_graph(rng, n, 600, weighted=trial % 2 == 0)
        G = nx.Graph()
        G.add_nodes_from(range(n))
        for a, b, w in edges:
            G.add_edge(a, b, weight=w)
        g = O.Graph.from_edges([e[0] for e in edges], [e[1] for e in edges],
                               np.array([e[2] for e in edges], np.uint32))
        comm = rng.integers(0, 15, n).astype(np.uint32)
        parts = [set(np.flatnonzero(comm == c).tolist()) for c in np.unique(comm)]
        want = nx.community.modularity(G, parts, weight="weight")
        q, _, _ = g.modularity(comm)
        assert abs(q - want) < 1e-12
        assert abs(Fraction(q) - exact_q(edges, n, comm)) < Fraction(1, 10**14)


COMBOS = list(itertools.product(["hash", "sort"], ["exact", "f32"], ["lowest", "reference"]))
ENUM = {"hash": O.HASH, "sort": O.SORT, "exact": O.GAIN_EXACT, "f32": O.GAIN_REF_F32,
        "lowest": O.TIE_LOWEST, "reference": O.TIE_REFERENCE}


def trace_of(res):
    return [(s["level"], s["sweep"], s["kind"], s["n_moved"], s["e_active"], s["sigma_in"],
             s["sum_tot_sq"], s["q"]) for s in res["sweeps"]]


@pytest.mark.parametrize("mode,gain,tie", COMBOS)
def test_oracle_matches_pipeline_on_fixture(fixture_path, mode, gain, tie):
    g = fixture_graph(fixture_path)
    row, col, w, _ = g.arrays()
    for thr in (1e-3, 1e-6):
        a = R.louvain(row, col, w, g.total_weight, mode, gain, tie, thr, thr)
        b = g.louvain(O.params(mode=ENUM[mode], gain=ENUM[gain], tie=ENUM[tie],
                               sweep_threshold=thr, level_threshold=thr))
        assert trace_of(b) == a["trace"]
        assert np.array_equal(a["best"], b["best"]) and a["modularity"] == b["modularity"]


@pytest.mark.parametrize("seed", range(6))
def test_oracle_matches_pipeline_random(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(50, 400))
    m = int(rng.integers(n, 5 * n))
    a = rng.integers(0, n, m); b = rng.integers(0, n, m)
    # self-loops and duplicates on purpose
    a = np.concatenate([a, np.arange(n), rng.integers(0, n, 5)])
    b = np.concatenate([b, (np.arange(n) + 1) % n, a[-5:] if m >= 5 else rng.integers(0, n, 5)])
    w = rng.integers(1, 10, len(a)).astype(np.uint32) if seed % 2 else None
    g = O.Graph.from_edges(a, b, w)
    row, col, ww, _ = g.arrays()
    for mode, gain, tie in COMBOS:
        ra = R.louvain(row, col, ww, g.total_weight, mode, gain, tie, 1e-6, 1e-6)
        rb = g.louvain(O.params(mode=ENUM[mode], gain=ENUM[gain], tie=ENUM[tie],
                                sweep_threshold=1e-6, level_threshold=1e-6))
        assert trace_of(rb) == ra["trace"], (seed, mode, gain, tie)
        assert np.array_equal(ra["best"], rb["best"])


def test_ring_of_cliques_finds_cliques():
    k, s = 10, 6   # ring of 10 cliques of 6
    a, b = [], []
    for c in range(k):
        base = c * s
        for i, j in itertools.combinations(range(s), 2):
            a.append(base + i); b.append(base + j)
        a.append(base); b.append(((c + 1) % k) * s + 1)
    g = O.Graph.from_edges(a, b)
    r = g.louvain(O.params(sweep_threshold=1e-9, level_threshold=1e-9))
    clique = np.arange(k * s) // s
    # the partition into cliques, and its closed-form modularity
    assert len(set(zip(r["best"].tolist(), clique.tolist()))) == k == r["n_comm"]
    m = k * (s * (s - 1) // 2) + k
    q_cliques = k * ((s * (s - 1) // 2) / m - ((s * (s - 1) + 2) / (2 * m)) ** 2)
    assert abs(r["modularity"] - q_cliques) < 1e-12


def test_aggregation_invariants(fixture_path):
    g = fixture_graph(fixture_path)
    rng = np.random.default_rng(9)
    row, col, w, k = g.arrays()
    for ncomm in (10000, 500, 7, 1):
        comm = rng.integers(0, ncomm, g.n).astype(np.uint32)
        new_id, n_new, h = g.aggregate(comm)
        used = np.unique(comm)
        assert n_new == len(used)
        assert np.array_equal(new_id[used], np.arange(n_new))         # order preserving
        r2, c2, w2, k2 = h.arrays()
        assert int(w2.sum()) == int(w.sum())                          # total weight kept
        assert h.total_weight == g.total_weight
        assert h.modularity(np.arange(n_new))[0] == g.modularity(comm)[0]
        tot = np.bincount(comm, weights=k.astype(np.float64), minlength=g.n)
        assert np.array_equal(k2, tot[used].astype(np.uint64))
        # rows sorted by destination, no duplicates
        for v in range(n_new):
            seg = c2[r2[v]:r2[v + 1]]
            assert np.all(np.diff(seg.astype(np.int64)) > 0)


def test_sweep_schedule(fixture_path):
    g = fixture_graph(fixture_path)
    kinds = {}
    for mode in (O.HASH, O.SORT, O.ADAPTIVE_SPEED, O.ADAPTIVE_MEMORY):
        r = g.louvain(O.params(mode=mode, sweep_threshold=-1.0, max_sweeps=7, max_levels=1))
        kinds[mode] = [s["kind"] for s in r["sweeps"]]
    assert kinds[O.HASH] == [0] + [2] * 7
    assert kinds[O.ADAPTIVE_MEMORY] == [0] + [2] * 7
    assert kinds[O.SORT] == [0] + [1] * 7
    assert kinds[O.ADAPTIVE_SPEED] == [0, 1, 1, 1, 1, 2, 2, 2]


def test_chunk_emulation_early_return():
    # path graph; only high-id vertices active: the first chunk selects nothing,
    # so the fast/sort sweeps return before reaching the active chunk
    n = 200
    a = np.arange(n - 1); b = a + 1
    g = O.Graph.from_edges(a, b)
    comm = np.arange(n, dtype=np.uint32)
    act = np.zeros(n, np.uint8); act[150:] = 1
    p_on = O.params(chunk_emulation=1, chunk_entries=100)
    p_off = O.params()
    t_off, moved_off, _, _ = g.best_moves(comm, act, O.SWEEP_SORT, p_off)
    t_on, moved_on, _, _ = g.best_moves(comm, act, O.SWEEP_SORT, p_on)
    assert moved_off > 0 and moved_on == 0 and np.array_equal(t_on, comm)
    t_h, moved_h, _, _ = g.best_moves(comm, act, O.SWEEP_HASH, p_on)   # no early return
    assert np.array_equal(t_h, t_off)
    # chunks larger than the graph change nothing
    t_big, _, _, _ = g.best_moves(comm, act, O.SWEEP_SORT, O.params(chunk_emulation=1))
    assert np.array_equal(t_big, t_off)


def test_edge_parsing(tmp_path):
    p = tmp_path / "g.edge"
    p.write_text("% header\n% more\n5 6\n6 5\n7 7\n6 8\n5 6\nbad 9 9\n8 5\n")
    g = O.Graph.load_edge(str(p), skip=2)
    row, col, w, k = g.arrays()
    assert g.min_id == 5 and g.n == 4
    # (5,6) dedup, self-loop (7,7) -> two entries, parsing stops at "bad"
    assert g.E == 2 * 2 + 2
    assert row.tolist() == [0, 1, 3, 5, 6]
    assert col.tolist() == [1, 0, 3, 2, 2, 1]
    assert g.total_weight == 3.0
    pw = tmp_path / "w.edge"
    pw.write_text("0 1 4\n1 2 3\n1 0 9\n")
    gw = O.Graph.load_edge(str(pw), weighted=True)
    _, _, w2, k2 = gw.arrays()
    assert w2.tolist() == [4, 4, 3, 3] and k2.tolist() == [4, 7, 3]   # first weight wins
    with pytest.raises(RuntimeError, match="dense"):
        O.Graph.from_edges([0, 5], [1, 6])
    empty = tmp_path / "e.edge"
    empty.write_text("x\n")
    with pytest.raises(RuntimeError, match="parse"):
        O.Graph.load_edge(str(empty))


@pytest.mark.parametrize("seed", range(4))
def test_oracle_refine_matches_pipeline(fixture_path, seed):
    """Opt-in refinement: the final partition swept again on the level-0 graph
    (no fast sweep) and renumbered -- the oracle against the independent numpy
    pipeline, trace for trace; its modularity equals a from-scratch
    recomputation and, on these inputs, is no lower than without refinement."""
    if seed == 0:
        g = fixture_graph(fixture_path)
    else:
        rng = np.random.default_rng(300 + seed)
        n = int(rng.integers(100, 600))
        a = np.concatenate([rng.integers(0, n, 4 * n), np.arange(n)])
        b = np.concatenate([rng.integers(0, n, 4 * n), (np.arange(n) + 1) % n])
        w = rng.integers(1, 8, len(a)).astype(np.uint32) if seed % 2 else None
        g = O.Graph.from_edges(a, b, w)
    row, col, w, _ = g.arrays()
    for mode, gain, tie in COMBOS[:4]:
        ra = R.louvain(row, col, w, g.total_weight, mode, gain, tie, 1e-6, 1e-6, refine=True)
        rb = g.louvain(O.params(mode=ENUM[mode], gain=ENUM[gain], tie=ENUM[tie],
                                sweep_threshold=1e-6, level_threshold=1e-6, refine=1))
        assert trace_of(rb) == ra["trace"], (seed, mode, gain, tie)
        assert np.array_equal(ra["best"], rb["best"]) and ra["n_comm"] == rb["n_comm"]
        assert rb["modularity"] == g.modularity(rb["best"])[0]
        plain = g.louvain(O.params(mode=ENUM[mode], gain=ENUM[gain], tie=ENUM[tie],
                                   sweep_threshold=1e-6, level_threshold=1e-6))
        assert rb["modularity"] >= plain["modularity"]
        assert rb["sweeps"][:len(plain["sweeps"])] == plain["sweeps"]


@pytest.mark.parametrize("kind", ["star", "random"])
def test_from_edges_with_three_radix_digits(kind):
    """Ids past 65,535 with every pair's smaller end below it: the oracle's CSR
    build sorts keys (min << 32 | max) whose set bits fill three 16-bit digits,
    an odd number of radix passes, so the sorted keys come back from the
    scratch arrays (a Barabasi-Albert graph of 66,123 vertices took this path).
    The CSR equals numpy's: both directions, sorted by (source, destination),
    duplicates kept once."""
    n = 70001
    if kind == "star":
        a = np.concatenate([[0], np.ones(n - 2, np.int64)])
        b = np.concatenate([[1], np.arange(2, n, dtype=np.int64)])
    else:
        rng = np.random.default_rng(7)
        m = 200_000
        a = rng.integers(0, 60_000, m)
        b = rng.integers(0, n, m)
        keep = a != b
        a, b = a[keep], b[keep]
        # every id in use (the graph must be dense): a path over the low ids,
        # each higher id tied to a low one
        hi_ids = np.arange(60_000, n)
        a = np.concatenate([a, np.arange(0, 59_999), rng.integers(0, 60_000, len(hi_ids))])
        b = np.concatenate([b, np.arange(1, 60_000), hi_ids])
        a, b = np.where(rng.random(len(a)) < 0.5, (a, b), (b, a))
    assert np.minimum(a, b).max() < 65536 <= np.maximum(a, b).max()
    row, col, w, k = O.Graph.from_edges(a, b).arrays()
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    pairs = np.unique(lo.astype(np.int64) << 32 | hi)
    lo, hi = pairs >> 32, pairs & 0xFFFFFFFF
    src = np.concatenate([lo, hi])
    dst = np.concatenate([hi, lo])
    order = np.lexsort((dst, src))
    assert np.array_equal(col, dst[order])
    assert np.array_equal(row, np.concatenate([[0], np.cumsum(np.bincount(src, minlength=n))]))
    assert np.array_equal(k, np.bincount(src, minlength=n))
