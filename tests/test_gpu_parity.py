"""GPU parity: the HIP library (through the C ABI) against the CPU oracle.

Bar: bit-exact per-vertex best moves, partitions (FNV hash per sweep), exact
modularity parts (sigma_in, sum of tot^2), renumbered ids and contracted
graphs; modularity (f64 combine of exact integers) identical.
"""
import itertools
import sys
import zlib

import numpy as np
import pytest

import louvain_amd as L
from oracle import oracle as O
from partition_mix import fmix32

pytestmark = pytest.mark.gpu

MODES = [("hash", L.HASH, O.HASH), ("sort", L.SORT, O.SORT),
         ("adaptive_speed", L.ADAPTIVE_SPEED, O.ADAPTIVE_SPEED),
         ("adaptive_memory", L.ADAPTIVE_MEMORY, O.ADAPTIVE_MEMORY)]
GAINS = [("exact", L.GAIN_EXACT, O.GAIN_EXACT), ("f32", L.GAIN_REF_F32, O.GAIN_REF_F32)]
TIES = [("lowest", L.TIE_LOWEST, O.TIE_LOWEST), ("reference", L.TIE_REFERENCE, O.TIE_REFERENCE)]
REC_KEYS = ["level", "sweep", "kind", "n", "n_active", "n_moved", "n_considered", "E",
            "e_active", "sigma_in", "sum_tot_sq", "q", "comm_hash"]


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if L.device_count() < 1:
        pytest.skip("no GPU visible")


def compare_runs(a, b, label):
    sa = [{k: s[k] for k in REC_KEYS} for s in a["sweeps"]]
    sb = [{k: s[k] for k in REC_KEYS} for s in b["sweeps"]]
    for i, (x, y) in enumerate(zip(sa, sb)):
        assert x == y, f"{label}: sweep record {i} differs\nGPU   {x}\noracle {y}"
    assert len(sa) == len(sb), f"{label}: {len(sa)} vs {len(sb)} sweeps"
    assert a["modularity"] == b["modularity"], label
    assert a["n_comm"] == b["n_comm"], label
    assert np.array_equal(a["best"], b["best"]), label


def run_both(edges, gprm, oprm, w=None):
    ea, eb = edges
    g = L.Graph.from_edges(ea, eb, w)
    og = O.Graph.from_edges(ea, eb, w)
    return g.louvain(gprm, hash_trace=True), og.louvain(oprm)


@pytest.mark.parametrize("mode,gain,tie", list(itertools.product(MODES, GAINS, TIES)),
                         ids=lambda x: x[0])
def test_fixture_full_run(fixture_path, mode, gain, tie):
    for thr in (1e-3, 1e-6):
        label = f"fixture {mode[0]} {gain[0]} {tie[0]} thr={thr}"
        say(label)
        g = L.Graph.load_edge(fixture_path)
        og = O.Graph.load_edge(fixture_path)
        a = g.louvain(L.params(mode=mode[1], gain=gain[1], tie=tie[1], sweep_threshold=thr,
                               level_threshold=thr), hash_trace=True)
        b = og.louvain(O.params(mode=mode[2], gain=gain[2], tie=tie[2], sweep_threshold=thr,
                                level_threshold=thr))
        compare_runs(a, b, label)


def _random_graph(rng, n, m, hubs=0, hub_deg=0, selfloops=0):
    a = rng.integers(0, n, m)
    b = rng.integers(0, n, m)
    ea, eb = [a], [b]
    for h in range(hubs):
        nb = rng.choice(n, size=min(hub_deg, n - 1), replace=False)
        ea.append(np.full(len(nb), h)); eb.append(nb)
    if selfloops:
        s = rng.integers(0, n, selfloops)
        ea.append(s); eb.append(s)
    # ring keeps ids dense
    ea.append(np.arange(n)); eb.append((np.arange(n) + 1) % n)
    return np.concatenate(ea), np.concatenate(eb)


GRAPHS = {
    "er_small": dict(n=300, m=900),
    "er_selfloops": dict(n=500, m=2000, selfloops=40),
    "hubs": dict(n=20000, m=40000, hubs=3, hub_deg=9000),
    "mid": dict(n=5000, m=60000),
    # rows of ~4,050 distinct neighbours: the degree 1025-4096 bucket's
    # best-move tables (4,960 slots) filled to ~82%, their fullest
    "bucket7_full": dict(n=30000, m=45000, hubs=30, hub_deg=4050),
}


@pytest.mark.parametrize("name", list(GRAPHS))
@pytest.mark.parametrize("weighted", [False, True])
def test_random_graph_full_run(name, weighted):
    rng = np.random.default_rng(zlib.crc32(name.encode()) + weighted)
    ea, eb = _random_graph(rng, **GRAPHS[name])
    w = rng.integers(1, 11, len(ea)).astype(np.uint32) if weighted else None
    for (mn, gm, om), (gn, gg, og_) in itertools.product(MODES[:2], GAINS):
        label = f"{name} w={weighted} {mn} {gn}"
        say(label)
        a, b = run_both((ea, eb), L.params(mode=gm, gain=gg, sweep_threshold=1e-6,
                                           level_threshold=1e-6),
                        O.params(mode=om, gain=og_, sweep_threshold=1e-6,
                                 level_threshold=1e-6), w)
        compare_runs(a, b, label)


def test_heavy_weights_full_run():
    """Weights up to 200,000 with the total weight just under the 2^31 bound
    of 2m (exact scores 2m*w - k*tot in int64, totals in u32): full runs in
    the hash and sort modes and both gains equal the oracle's traces."""
    rng = np.random.default_rng(77)
    ea, eb = _random_graph(rng, n=3000, m=7000, hubs=1, hub_deg=2500, selfloops=10)
    w = rng.integers(1, 200_001, len(ea)).astype(np.uint64)
    # scale so that 2m stays below 2^31 (self-loops count twice)
    two_m = int(2 * w.sum())
    if two_m >= (1 << 31) - 1:
        w = np.maximum(1, (w * ((1 << 31) - 2) // (two_m + 1))).astype(np.uint64)
    w = w.astype(np.uint32)
    assert 2 * int(w.sum()) < (1 << 31)
    for (mn, gm, om), (gn, gg, og_) in itertools.product(MODES[:2], GAINS):
        label = f"heavy weights {mn} {gn}"
        say(label)
        a, b = run_both((ea, eb), L.params(mode=gm, gain=gg, sweep_threshold=1e-6,
                                           level_threshold=1e-6),
                        O.params(mode=om, gain=og_, sweep_threshold=1e-6,
                                 level_threshold=1e-6), w)
        compare_runs(a, b, label)


@pytest.mark.parametrize("kind", [L.SWEEP_SORT, L.SWEEP_HASH])
def test_best_moves_snapshots(kind):
    rng = np.random.default_rng(7 + kind)
    ea, eb = _random_graph(rng, n=12000, m=50000, hubs=2, hub_deg=6000, selfloops=10)
    w = rng.integers(1, 6, len(ea)).astype(np.uint32)
    g = L.Graph.from_edges(ea, eb, w)
    og = O.Graph.from_edges(ea, eb, w)
    n = g.n
    c = L.Community(g)
    for trial in range(6):
        ncomm = [n, 5000, 300, 20, 3, 1][trial]
        comm = rng.integers(0, ncomm, n).astype(np.uint32)
        active = (rng.random(n) < [1.0, 0.5, 0.9, 0.1, 1.0, 1.0][trial]).astype(np.uint8)
        c.set_partition(comm)
        for (gn, gg, og_), (tn, gt, ot) in itertools.product(GAINS, TIES):
            say(f"snapshot kind={kind} trial={trial} {gn} {tn}")
            tg, ea_g = c.best_moves(kind, L.params(gain=gg, tie=gt), active)
            to, moved, ea_o, _ = og.best_moves(comm, active, kind, O.params(gain=og_, tie=ot))
            assert np.array_equal(tg, to), f"trial {trial} {gn} {tn}: " \
                f"{np.flatnonzero(tg != to)[:10]}"
            assert ea_g == ea_o
        q, si, s = c.modularity_parts()
        qo, sio, so = og.modularity(comm)
        assert (si, s) == (sio, so) and q == qo


def test_hub_partition_ranges():
    """Hub rows are summed per community hash partition in 4096-slot LDS
    tables, range by range of the mix bits below the partition; a range with
    more than 3072 communities is halved and redone.  Hub A's 7000 neighbour
    communities all mix into partition 0 and the lowest 1/8 of its range, so
    its workgroup halves the range repeatedly; hub B spreads 20000 communities
    over one partition's many ranges; hub C sees 12 communities; hub D has
    self-loops and neighbours in its own community."""
    rng = np.random.default_rng(41)
    ids = np.arange(1, 4_000_000, dtype=np.uint64)
    mix = fmix32(ids)
    part, sub = mix >> 27, mix & ((1 << 27) - 1)
    packed = ids[(part == 0) & (sub < (1 << 24))][:7000].astype(np.int64)
    spread = ids[part == 5][:20000].astype(np.int64)
    assert len(packed) == 7000 and len(spread) == 20000
    n = int(max(packed.max(), spread.max())) + 2
    hubs = [n - 1, n - 2, n - 3, n - 4]
    leaves = rng.permutation(np.arange(8, n - 8))[:40000]
    la, lb, lc, ld = leaves[:7000], leaves[7000:27000], leaves[27000:33000], leaves[33000:39000]
    chain = np.arange(n - 1)   # every id in use (the library wants dense ids)
    ea = np.concatenate([np.full(7000, hubs[0]), np.full(20000, hubs[1]), np.full(6000, hubs[2]),
                         np.full(6000, hubs[3]), [hubs[3]] * 2, la[:500], chain])
    eb = np.concatenate([la, lb, lc, ld, [hubs[3]] * 2, lb[:500], chain + 1])
    w = rng.integers(1, 7, len(ea)).astype(np.uint32)
    g = L.Graph.from_edges(ea, eb, w)
    og = O.Graph.from_edges(ea, eb, w)
    assert g.n == n
    comm = np.arange(n, dtype=np.uint32)
    comm[la] = packed
    comm[lb] = spread
    comm[lc] = rng.integers(0, 12, len(lc)) + 100
    comm[ld[:2500]] = hubs[3]
    comm[ld[2500:]] = rng.integers(0, n, 3500)
    # hub A's communities: one partition, the lowest eighth of its mix range,
    # far above the 3072 an LDS range holds
    mix_a = fmix32(comm[la])
    assert np.all(mix_a >> 27 == 0) and len(np.unique(comm[la])) > 2 * 3072
    c = L.Community(g)
    c.set_partition(comm)
    active = np.ones(n, np.uint8)
    for (gn, gg, og_), (tn, gt, ot) in itertools.product(GAINS, TIES):
        say(f"hub partitions {gn} {tn}")
        tg, ea_g = c.best_moves(L.SWEEP_HASH, L.params(gain=gg, tie=gt), active)
        to, moved, ea_o, _ = og.best_moves(comm, active, O.SWEEP_HASH, O.params(gain=og_, tie=ot))
        assert np.array_equal(tg[hubs], to[hubs]), (gn, tn, tg[hubs], to[hubs])
        assert np.array_equal(tg, to), f"{gn} {tn}: {np.flatnonzero(tg != to)[:10]}"
        assert ea_g == ea_o


def test_fast_sweep_identity():
    rng = np.random.default_rng(11)
    ea, eb = _random_graph(rng, n=15000, m=60000, hubs=2, hub_deg=7000)
    g = L.Graph.from_edges(ea, eb)
    og = O.Graph.from_edges(ea, eb)
    c = L.Community(g)
    n = g.n
    ident = np.arange(n, dtype=np.uint32)
    for gn, gg, og_ in GAINS:
        say(f"fast identity {gn}")
        tg, _ = c.best_moves(L.SWEEP_FAST, L.params(gain=gg))
        to, _, _, _ = og.best_moves(ident, np.ones(n, np.uint8), O.SWEEP_FAST, O.params(gain=og_))
        assert np.array_equal(tg, to)


def test_commit_mark_bytes_path():
    """Graphs below 2^31 vertices keep the commit's marks in bit 63 of each
    vertex's (community, target) word; larger ones use the byte array.  The
    test hook forces the byte path: full runs still equal the oracle."""
    rng = np.random.default_rng(23)
    ea, eb = _random_graph(rng, n=20000, m=60000, hubs=2, hub_deg=7000, selfloops=20)
    w = rng.integers(1, 8, len(ea)).astype(np.uint32)
    L.debug_mark_bytes(True)
    try:
        for (mn, gm, om), (gn, gg, og_) in itertools.product(MODES[:2], GAINS):
            label = f"mark bytes {mn} {gn}"
            say(label)
            a, b = run_both((ea, eb), L.params(mode=gm, gain=gg, sweep_threshold=1e-6,
                                               level_threshold=1e-6),
                            O.params(mode=om, gain=og_, sweep_threshold=1e-6,
                                     level_threshold=1e-6), w)
            compare_runs(a, b, label)
    finally:
        L.debug_mark_bytes(False)


def _table_slots(d):
    """Slots of the best-move LDS table of a row of d entries (d + d/2 + 1,
    rounded up to 4)."""
    return (d + d // 2 + 1 + 3) & ~3


def _first_slot(c, t):
    """First probe of key c in a table of t slots (multiplicative hash scaled
    by a high multiply)."""
    return ((c.astype(np.uint64) * 2654435761) & 0xFFFFFFFF) * t >> 32


def test_tables_full_at_bucket_limits():
    """A vertex at the largest degree of every LDS bucket (4 ... 4096) and one
    past it, each neighbour in its own community: the tables hold as many
    distinct keys as they ever do (two thirds of the slots), and every key
    hashes into the last quarter of the table, so the probe runs wrap past its
    end.  Best moves of the hash sweep equal the oracle's in every gain and
    tie mode, weighted and not."""
    rng = np.random.default_rng(5)
    degs = [4, 5, 8, 9, 16, 17, 32, 33, 64, 65, 256, 257, 1024, 1025, 4096]
    n = 400_000
    centers = np.arange(len(degs)) + 10
    ids = np.arange(1000, n, dtype=np.uint64)
    ea, eb, leaf_comm = [], [], {}
    next_leaf = n - 1
    for v, d in zip(centers, degs):
        # d - 2 leaves; the centre's two chain neighbours complete its degree
        t = _table_slots(d)
        pool = ids[_first_slot(ids, t) >= (3 * t) // 4]
        comms = rng.choice(pool, size=d - 2, replace=False)
        leaves = np.arange(next_leaf - (d - 2) + 1, next_leaf + 1)
        next_leaf -= d - 2
        ea.append(np.full(d - 2, v)); eb.append(leaves)
        leaf_comm.update(zip(leaves.tolist(), comms.tolist()))
    chain = np.arange(n - 1)   # every id in use
    ea = np.concatenate(ea + [chain]); eb = np.concatenate(eb + [chain + 1])
    comm = np.arange(n, dtype=np.uint32)
    lv = np.fromiter(leaf_comm.keys(), np.int64)
    comm[lv] = np.fromiter(leaf_comm.values(), np.int64)
    active = np.zeros(n, np.uint8)
    active[centers] = 1
    for weighted in (False, True):
        w = rng.integers(1, 9, len(ea)).astype(np.uint32) if weighted else None
        g = L.Graph.from_edges(ea, eb, w)
        og = O.Graph.from_edges(ea, eb, w)
        assert g.n == n
        row, _, _ = g.csr()
        assert np.array_equal(np.diff(row.astype(np.int64))[centers], degs)
        c = L.Community(g)
        c.set_partition(comm)
        for (gn, gg, og_), (tn, gt, ot) in itertools.product(GAINS, TIES):
            say(f"full tables weighted={weighted} {gn} {tn}")
            tg, ea_g = c.best_moves(L.SWEEP_HASH, L.params(gain=gg, tie=gt), active)
            to, _, ea_o, _ = og.best_moves(comm, active, O.SWEEP_HASH,
                                           O.params(gain=og_, tie=ot))
            assert np.array_equal(tg[centers], to[centers]), (gn, tn, tg[centers], to[centers])
            assert np.array_equal(tg, to)
            assert ea_g == ea_o


def test_aggregation_matches_oracle():
    rng = np.random.default_rng(5)
    ea, eb = _random_graph(rng, n=8000, m=40000, hubs=1, hub_deg=5000, selfloops=20)
    w = rng.integers(1, 9, len(ea)).astype(np.uint32)
    g = L.Graph.from_edges(ea, eb, w)
    og = O.Graph.from_edges(ea, eb, w)
    n = g.n
    for ncomm in (n, 3000, 100, 1):
        say(f"aggregate ncomm={ncomm}")
        comm = rng.integers(0, ncomm, n).astype(np.uint32)
        c = L.Community(g)
        c.set_partition(comm)
        c.aggregation_phase(True)
        new_id, n_new, h = og.aggregate(comm)
        row, col, ww, k = c.graph_csr()
        r2, c2, w2, k2 = h.arrays()
        assert len(row) - 1 == n_new
        assert np.array_equal(row, r2) and np.array_equal(col, c2) and np.array_equal(ww, w2)
        assert np.array_equal(k.astype(np.uint64), k2)
        best, nb = c.best_communities(n)
        assert nb == n_new
        assert np.array_equal(best, new_id[comm])
        # Q is preserved by contraction
        q_new, _, _ = c.modularity_parts()
        assert q_new == og.modularity(comm)[0]
        c.close()


def _packed_hub_graph(rng):
    """Four hubs whose leaves are ids chosen by their community mix: hub A's
    7000 leaves all mix into the lowest 1/128 of the top-bit-0 half (one
    partition for up to 2 partitions, and far above one LDS piece), hub B's
    20000 leaves into one of 32 partitions; a chain keeps every id in use."""
    ids = np.arange(1, 4_000_000, dtype=np.uint64)
    mix = fmix32(ids)
    part, sub = mix >> 27, mix & ((1 << 27) - 1)
    packed = ids[(part == 0) & (sub < (1 << 24))][:7000].astype(np.int64)
    spread = ids[part == 5][:20000].astype(np.int64)
    n = int(max(packed.max(), spread.max())) + 2
    hubs = [n - 1, n - 2, n - 3, n - 4]
    others = rng.permutation(np.arange(8, n - 8))[:12000]
    chain = np.arange(n - 1)
    ea = np.concatenate([np.full(7000, hubs[0]), np.full(20000, hubs[1]), np.full(6000, hubs[2]),
                         np.full(6000, hubs[3]), [hubs[3]] * 2, chain])
    eb = np.concatenate([packed, spread, others[:6000], others[6000:], [hubs[3]] * 2, chain + 1])
    return ea, eb, n, hubs, packed


def test_aggregation_giant_partitions():
    """Contraction of giant communities (above 4096 virtual entries) through
    destination hash partitions and LDS sums: the identity partition (hub A's
    row overflows an LDS piece again and again), blocks of 1000 vertices,
    one community, three communities, and all even vertices merged (a giant
    with ~900 K distinct destinations over 256 partitions)."""
    rng = np.random.default_rng(43)
    ea, eb, n, hubs, packed = _packed_hub_graph(rng)
    w = rng.integers(1, 5, len(ea)).astype(np.uint32)
    g = L.Graph.from_edges(ea, eb, w)
    og = O.Graph.from_edges(ea, eb, w)
    assert g.n == n
    v = np.arange(n, dtype=np.uint32)
    evens = np.where(v % 2 == 0, 0, v).astype(np.uint32)
    parts = {"identity": v, "blocks": v // 1000, "one": np.zeros(n, np.uint32),
             "three": v % 3, "evens": evens}
    for name, comm in parts.items():
        say(f"aggregate giants {name}")
        c = L.Community(g)
        c.set_partition(comm)
        c.aggregation_phase(True)
        new_id, n_new, h = og.aggregate(comm)
        row, col, ww, k = c.graph_csr()
        r2, c2, w2, k2 = h.arrays()
        assert len(row) - 1 == n_new, name
        assert np.array_equal(row, r2), name
        assert np.array_equal(col, c2) and np.array_equal(ww, w2), name
        assert np.array_equal(k.astype(np.uint64), k2), name
        q_new, _, _ = c.modularity_parts()
        assert q_new == og.modularity(comm)[0], name
        c.close()


def test_rmat_scale16_parity():
    """R-MAT scale 16 in all four modes, and the hash mode with the
    reference's tie rule, trace for trace against the oracle."""
    ea, eb, n = L.gen_rmat(16, seed=2)
    runs = [(m, TIES[0]) for m in MODES] + [(MODES[0], TIES[1])]
    for (mn, gm, om), (tn, gt, ot) in runs:
        say(f"rmat16 {mn} {tn}")
        a, b = run_both((ea, eb), L.params(mode=gm, tie=gt, sweep_threshold=1e-6,
                                           level_threshold=1e-6),
                        O.params(mode=om, tie=ot, sweep_threshold=1e-6, level_threshold=1e-6))
        compare_runs(a, b, f"rmat16 {mn} {tn}")


def test_rmat_scale18_parity():
    """R-MAT scale 18 (262 K vertices, 7 M entries: hub rows, giant
    communities, every LDS bucket) at the benchmark's thresholds, trace for
    trace against the oracle in the hash and sort modes and both gains."""
    ea, eb, n = L.gen_rmat(18, 16, 0.57, 0.19, 0.19, seed=18)
    for (mn, gm, om), (gn, gg, og_) in [(MODES[0], GAINS[0]), (MODES[1], GAINS[0]),
                                         (MODES[0], GAINS[1])]:
        label = f"rmat18 {mn} {gn}"
        say(label)
        a, b = run_both((ea, eb), L.params(mode=gm, gain=gg), O.params(mode=om, gain=og_))
        compare_runs(a, b, label)


RMAT22_TRACES = [("rmat22_hash_trace.json", L.GAIN_EXACT, L.TIE_LOWEST),
                 ("rmat22_hash_ref_trace.json", L.GAIN_REF_F32, L.TIE_REFERENCE)]


@pytest.mark.parametrize("trace,gain,tie", RMAT22_TRACES, ids=["exact", "reference_method"])
def test_rmat22_full_size_golden_trace(trace, gain, tie):
    """The benchmark workload at full size (R-MAT scale 22, 2.4 M vertices,
    128 M entries) against the oracle's trace of the same run, computed once
    by tests/golden/make_rmat22_trace.py: every sweep record (active, moved,
    entries, exact sigma_in and sum of tot^2, Q, FNV hash of the partition),
    the final modularity, communities and the final partition's SHA-256.  Two
    traces: exact gains with lowest-id ties, and the reference's own method
    (f32-rounded delta-Q, operatorsDelta.h:19-20, with the hash sweeps'
    highest-id ties, phaseOptimizationHash.cu:23-35)."""
    import hashlib
    import json
    import os
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", trace)))
    ea, eb, n = L.gen_rmat(22, 16, 0.57, 0.19, 0.19, seed=2)
    g = L.Graph.from_edges(ea, eb)
    assert g.n == gold["n"]
    say(f"rmat22 full size: {trace}")
    r = g.louvain(L.params(mode=L.HASH, gain=gain, tie=tie), hash_trace=True)
    got = [{k: s[k] for k in REC_KEYS} for s in r["sweeps"]]
    assert len(got) == gold["total_sweeps"]
    for i, (x, y) in enumerate(zip(got, gold["sweeps"])):
        assert x == y, f"sweep record {i} differs\nGPU    {x}\noracle {y}"
    assert r["modularity"] == gold["modularity"]
    assert r["n_comm"] == gold["n_comm"] and r["levels"] == gold["levels"]
    assert hashlib.sha256(np.ascontiguousarray(r["best"], np.uint32).tobytes()).hexdigest() \
        == gold["best_sha256"]


def test_chunk_emulation_parity(fixture_path):
    for mode in (MODES[0], MODES[1]):
        say(f"chunk emulation {mode[0]}")
        g = L.Graph.load_edge(fixture_path)
        og = O.Graph.load_edge(fixture_path)
        a = g.louvain(L.params(mode=mode[1], chunk_emulation=1, chunk_entries=3000),
                      hash_trace=True)
        b = og.louvain(O.params(mode=mode[2], chunk_emulation=1, chunk_entries=3000))
        compare_runs(a, b, f"chunk {mode[0]}")


def test_phase_api_matches_louvain(fixture_path):
    """OptimizationPhase / AggregationPhase driven one call at a time."""
    g = L.Graph.load_edge(fixture_path)
    ref = g.louvain(L.params())
    c = L.Community(g)
    prm = L.params()
    while True:
        old = c.modularity
        c.optimization_phase(prm)
        c.refine_phase()
        cont = c.modularity - old > prm.level_threshold
        c.aggregation_phase(cont, prm)
        if not cont:
            break
    best, nc = c.best_communities(g.n)
    assert np.array_equal(best, ref["best"]) and nc == ref["n_comm"]
    assert c.modularity == ref["modularity"]


@pytest.mark.parametrize("gain", GAINS, ids=lambda x: x[0])
def test_first_sweep_decided_by_contraction(gain):
    """lv_community_run decides each level's first (fast) sweep while the
    contraction builds the level (LDS-table rows and giant partitions alike):
    those sweeps are flagged pre_decided, and the run equals the phase API's,
    whose levels run that sweep's best-move pass (R-MAT scale 18: hub rows and
    giant communities at levels 1-2), and the oracle's trace."""
    ea, eb, n = L.gen_rmat(18, 16, 0.57, 0.19, 0.19, seed=18)
    say(f"first sweep decided by the contraction, {gain[0]}")
    g = L.Graph.from_edges(ea, eb)
    prm = L.params(gain=gain[1])
    a = g.louvain(prm, hash_trace=True)
    assert a["levels"] >= 3
    for sw in a["sweeps"]:
        assert sw["pre_decided"] == (1 if sw["sweep"] == 0 and sw["level"] > 0 else 0), sw
    b = O.Graph.from_edges(ea, eb).louvain(O.params(gain=gain[2]))
    compare_runs(a, b, f"rmat18 first sweep {gain[0]}")
    c = L.Community(g)
    while True:
        old = c.modularity
        c.optimization_phase(prm)
        cont = c.modularity - old > prm.level_threshold
        c.aggregation_phase(cont, prm)
        if not cont:
            break
    best, nc = c.best_communities(g.n)
    assert np.array_equal(best, a["best"]) and nc == a["n_comm"]
    assert c.modularity == a["modularity"]
    c.close()


def test_chunk_emulation_early_return_snapshot():
    """An empty first chunk stops the fast/sort sweeps but not the hash sweep."""
    n = 200
    a = np.arange(n - 1)
    b = a + 1
    g = L.Graph.from_edges(a, b)
    og = O.Graph.from_edges(a, b)
    c = L.Community(g)
    comm = np.arange(n, dtype=np.uint32)
    act = np.zeros(n, np.uint8)
    act[150:] = 1
    for kind in (L.SWEEP_SORT, L.SWEEP_HASH):
        for emu in (0, 1):
            say(f"chunk snapshot kind={kind} emu={emu}")
            tg, _ = c.best_moves(kind, L.params(chunk_emulation=emu, chunk_entries=100), act)
            to, _, _, _ = og.best_moves(comm, act, kind,
                                        O.params(chunk_emulation=emu, chunk_entries=100))
            assert np.array_equal(tg, to)
    tg, _ = c.best_moves(L.SWEEP_SORT, L.params(chunk_emulation=1, chunk_entries=100), act)
    assert np.array_equal(tg, comm)          # nothing moves: the early return hit chunk 0


def test_sort_sweep_rmat_snapshot():
    """Sort-variant decisions (radix sort + runs) equal the oracle on R-MAT with hubs."""
    ea, eb, n = L.gen_rmat(15, seed=9)
    g = L.Graph.from_edges(ea, eb)
    og = O.Graph.from_edges(ea, eb)
    c = L.Community(g)
    rng = np.random.default_rng(3)
    for ncomm in (n, 2000, 50):
        comm = rng.integers(0, ncomm, n).astype(np.uint32)
        act = (rng.random(n) < 0.7).astype(np.uint8)
        c.set_partition(comm)
        for kind in (L.SWEEP_SORT, L.SWEEP_HASH):
            say(f"rmat15 snapshot ncomm={ncomm} kind={kind}")
            tg, eg = c.best_moves(kind, L.params(), act)
            to, _, eo, _ = og.best_moves(comm, act, kind, O.params())
            assert np.array_equal(tg, to) and eg == eo


def test_every_tile_choice_and_the_tuner_give_identical_runs():
    """The best-move kernels' tile (lanes per vertex) changes speed only: each
    candidate of every bucket, and the autotuner's pick, give the oracle's run."""
    rng = np.random.default_rng(41)
    ea, eb = _random_graph(rng, n=30000, m=150000, hubs=2, hub_deg=6000, selfloops=20)
    w = rng.integers(1, 6, len(ea)).astype(np.uint32)
    g = L.Graph.from_edges(ea, eb, w)
    ref = O.Graph.from_edges(ea, eb, w).louvain(O.params(sweep_threshold=1e-6,
                                                          level_threshold=1e-6))
    prm = L.params(sweep_threshold=1e-6, level_threshold=1e-6)
    c = L.Community(g)
    for k in range(3):
        # candidate k where the bucket has one, else the default
        sel = np.array([k if k < n else 0 for n in (3, 3, 3, 3, 2, 2, 1, 1) for _ in (0, 1)],
                       np.uint8)
        c.set_tiles(sel)
        assert np.array_equal(c.tiles(), sel)
        c.reset()
        say(f"tiles {sel.tolist()}")
        compare_runs(c.run(prm, hash_trace=True), ref, f"tile candidate {k}")
    t = L.Community(g)
    t.reset()
    compare_runs(t.run(L.params(sweep_threshold=1e-6, level_threshold=1e-6, autotune=1),
                       hash_trace=True), ref, "autotuned")
    say(f"tuned tiles {t.tiles().tolist()}")


@pytest.mark.parametrize("mode,gain", list(itertools.product(MODES[:2], GAINS)),
                         ids=lambda x: x[0])
def test_refine_matches_oracle(fixture_path, mode, gain):
    """The opt-in refinement (final partition swept again on the level-0
    graph), trace for trace against the oracle; lv_refine after a plain run
    gives the same partition as refine=1."""
    rng = np.random.default_rng(77)
    ea, eb = _random_graph(rng, n=6000, m=30000, hubs=1, hub_deg=5000, selfloops=10)
    w = rng.integers(1, 6, len(ea)).astype(np.uint32)
    for label, g, og in (("fixture", L.Graph.load_edge(fixture_path), O.Graph.load_edge(fixture_path)),
                         ("random", L.Graph.from_edges(ea, eb, w), O.Graph.from_edges(ea, eb, w))):
        kw = dict(sweep_threshold=1e-6, level_threshold=1e-6)
        say(f"refine {label} {mode[0]} {gain[0]}")
        a = g.louvain(L.params(mode=mode[1], gain=gain[1], refine=1, **kw), hash_trace=True)
        b = og.louvain(O.params(mode=mode[2], gain=gain[2], refine=1, **kw))
        compare_runs(a, b, f"refine {label} {mode[0]} {gain[0]}")
        c = L.Community(g)
        plain = c.run(L.params(mode=mode[1], gain=gain[1], **kw))
        assert a["modularity"] >= plain["modularity"]
        r = c.refine(g.n, L.params(mode=mode[1], gain=gain[1], **kw))
        assert r["modularity"] == a["modularity"] and np.array_equal(r["best"], a["best"])


BOUND_SETS = [[2, 6, 12, 24, 48, 128, 512, 2048], [4, 8, 16, 32, 64, 256, 1024, 1024],
              [1, 3, 9, 27, 64, 200, 1000, 4000], [3, 6, 12, 24, 48, 192, 768, 3072],
              [4, 8, 16, 32, 64, 512, 1024, 4096], [2, 6, 12, 24, 48, 300, 300, 4000]]


def test_perturbed_bucket_bounds_equal_oracle(fixture_path):
    """Runs with other degree bounds of the best-move buckets (vertices moved
    between LDS tables of other sizes and tiles, and between the LDS buckets
    and the hub path) are trace for trace the oracle's: the results do not
    depend on the bucketing.  The extreme set sends every row of more than
    two entries down the hub path."""
    ea, eb, n = L.gen_rmat(16, 16, 0.57, 0.19, 0.19, seed=31)
    g = L.Graph.from_edges(ea, eb)
    og = O.Graph.from_edges(ea, eb)
    for mode in MODES[:2]:
        ref = og.louvain(O.params(mode=mode[2]))
        for bs in BOUND_SETS:
            label = f"rmat16 {mode[0]} bounds {bs}"
            say(label)
            compare_runs(g.louvain(L.params(mode=mode[1], bucket_max_deg=bs), hash_trace=True),
                         ref, label)
    fg = L.Graph.load_edge(fixture_path)
    ofg = O.Graph.load_edge(fixture_path)
    for gain in GAINS:
        for tie in TIES:
            label = f"fixture hubs everywhere {gain[0]} {tie[0]}"
            say(label)
            a = fg.louvain(L.params(gain=gain[1], tie=tie[1], bucket_max_deg=[0] * 7 + [2]),
                           hash_trace=True)
            compare_runs(a, ofg.louvain(O.params(gain=gain[2], tie=tie[2])), label)
    for bad in ([5, 8, 16, 32, 64, 256, 1024, 4096], [4, 8, 16, 32, 64, 256, 1024, 1],
                [4, 8, 16, 8, 64, 256, 1024, 4096], [4, 8, 16, 32, 64, 256, 1024, 5000],
                [4, 8, 16, 32, 64, 513, 1024, 4096]):
        with pytest.raises(L.LouvainError) as e:
            fg.louvain(L.params(bucket_max_deg=bad))
        assert e.value.status == 1                    # LV_ERR_ARG


def test_bucket_bound_tuner_keeps_results():
    """autotune = 1 times the three bound sets on the community's first run
    and keeps one; the run and a second one equal the oracle's."""
    ea, eb, n = L.gen_rmat(17, 16, 0.57, 0.19, 0.19, seed=32)
    g = L.Graph.from_edges(ea, eb)
    ref = O.Graph.from_edges(ea, eb).louvain(O.params())
    c = L.Community(g)
    for rep in range(2):
        c.reset()
        r = c.run(L.params(autotune=1), hash_trace=True, n0=g.n)
        compare_runs(r, ref, f"tuned run {rep}")
    say(f"tuned bounds: {c.buckets()}")
    assert c.buckets() in ([4, 8, 16, 32, 64, 256, 1024, 4096], [3, 6, 12, 24, 48, 192, 768, 3072],
                           [2, 4, 8, 16, 32, 128, 512, 2048])


def test_commit_and_contraction_records():
    """The commit's count of mover entries and the per-sweep / per-level
    measurement records.  The first sweep runs at the identity partition with
    every vertex listed, so its movers are the vertices whose oracle fast-sweep
    target is not themselves, and its mover entries their row lengths (hubs
    of 7,000 entries go through the large-row commit)."""
    rng = np.random.default_rng(31)
    ea, eb = _random_graph(rng, n=15000, m=60000, hubs=2, hub_deg=7000)
    g = L.Graph.from_edges(ea, eb)
    og = O.Graph.from_edges(ea, eb)
    n = g.n
    row = g.csr()[0].astype(np.int64)
    to, _, _, _ = og.best_moves(np.arange(n, dtype=np.uint32), np.ones(n, np.uint8),
                                O.SWEEP_FAST, O.params())
    movers = to != np.arange(n, dtype=np.uint32)
    want = int(np.diff(row)[movers].sum())
    c = L.Community(g)
    r = c.run(L.params(profile=1), best=False, n0=n)
    s0 = r["sweeps"][0]
    say(f"sweep 0: movers {int(movers.sum())}, mover entries {s0['mover_entries']} (want {want})")
    assert s0["n_moved"] == int(movers.sum())
    assert s0["mover_entries"] == want
    for s in r["sweeps"]:
        assert s["mover_entries"] <= s["E"]
        assert (s["mover_entries"] > 0) == (s["n_moved"] > 0)
        assert s["commit_ms"] > 0 and s["commit_bytes"] > 0
    lv = r["level_recs"]
    for a, b in zip(lv[:-1], lv[1:]):
        assert a["E_new"] == b["E"] and a["contract_ms"] > 0 and a["contract_bytes"] > 0
    assert lv[-1]["contract_ms"] == 0 and lv[-1]["E_new"] == 0
    k = c.kernel_stats()
    assert k["contractions"] == len(lv) - 1 and k["commit_passes"] == len(r["sweeps"])
    assert k["commit_entries"] == sum(s["mover_entries"] for s in r["sweeps"])
