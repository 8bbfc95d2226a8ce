"""Batches: many graphs in one disjoint-union launch sequence.

Bar: every graph's result (modularity, community count, levels, sweeps,
best partition) is bit-identical to the same graph run alone on the GPU and
to the oracle, for mixed sizes, weights, self-loops and hub vertices.
"""
import sys

import numpy as np
import pytest

import louvain_amd as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if L.device_count() < 1:
        pytest.skip("no GPU visible")


def _mixed_graphs(seed, count):
    """BA graphs of log-uniform size, weighted random graphs with self-loops,
    and one graph with hub vertices (degree > 4096)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        kind = i % 4
        if kind in (0, 1):
            n = int(np.exp(rng.uniform(np.log(200), np.log(4000))))
            ea, eb = L.gen_ba(n, int(rng.integers(2, 9)), seed=int(rng.integers(1 << 30)))
            w = rng.integers(1, 11, len(ea)).astype(np.uint32) if kind == 1 else None
        elif kind == 2:
            n = int(rng.integers(50, 1500))
            m = int(n * rng.uniform(1.5, 6))
            a = np.concatenate([rng.integers(0, n, m), np.arange(n), rng.integers(0, n, 5)])
            b = np.concatenate([rng.integers(0, n, m), (np.arange(n) + 1) % n, np.zeros(5, int)])
            b[-5:] = a[-5:]                       # self-loops
            ea, eb = a, b
            w = rng.integers(1, 6, len(ea)).astype(np.uint32)
        else:
            n = 600 if i != 3 else 9000
            a = [np.arange(n), rng.integers(0, n, 3 * n)]
            b = [(np.arange(n) + 1) % n, rng.integers(0, n, 3 * n)]
            if i == 3:                            # hubs
                for h in range(2):
                    nb = rng.choice(n, size=5000, replace=False)
                    a.append(np.full(len(nb), h)); b.append(nb)
            ea, eb = np.concatenate(a), np.concatenate(b)
            w = None
        out.append((ea, eb, w))
    return out


CASES = [
    ("hash exact lowest 1e-3", dict(mode=L.HASH), dict(mode=O.HASH)),
    ("hash f32 reference 1e-6",
     dict(mode=L.HASH, gain=L.GAIN_REF_F32, tie=L.TIE_REFERENCE, sweep_threshold=1e-6,
          level_threshold=1e-6),
     dict(mode=O.HASH, gain=O.GAIN_REF_F32, tie=O.TIE_REFERENCE, sweep_threshold=1e-6,
          level_threshold=1e-6)),
    ("sort exact 1e-6", dict(mode=L.SORT, sweep_threshold=1e-6, level_threshold=1e-6),
     dict(mode=O.SORT, sweep_threshold=1e-6, level_threshold=1e-6)),
    ("adaptive_speed caps", dict(mode=L.ADAPTIVE_SPEED, max_sweeps=3, max_levels=2),
     dict(mode=O.ADAPTIVE_SPEED, max_sweeps=3, max_levels=2)),
    ("hash autotuned", dict(mode=L.HASH, autotune=1), dict(mode=O.HASH)),
]


@pytest.fixture(scope="module")
def batch_graphs():
    return _mixed_graphs(11, 24)


@pytest.mark.parametrize("label,gp,op", CASES, ids=[c[0] for c in CASES])
def test_batch_equals_each_graph_alone(batch_graphs, label, gp, op):
    gs = [L.Graph.from_edges(a, b, w) for a, b, w in batch_graphs]
    B = L.Batch(gs)
    G, n, E = B.info()
    assert G == len(gs) and n == sum(g.n for g in gs) and E == sum(g.E for g in gs)
    say(f"batch {label}: {G} graphs, n={n}, E={E}")
    r = B.louvain(L.params(**gp))
    for i, (g, (a, b, w)) in enumerate(zip(gs, batch_graphs)):
        alone = g.louvain(L.params(**gp))
        orc = O.Graph.from_edges(a, b, w).louvain(O.params(**op))
        got = (r["modularity"][i], int(r["n_comm"][i]), int(r["levels"][i]), int(r["sweeps"][i]))
        for ref, name in ((alone, "alone"), (orc, "oracle")):
            exp = (ref["modularity"], ref["n_comm"], ref["levels"], ref["total_sweeps"])
            assert got == exp, f"{label} graph {i} vs {name}: {got} != {exp}"
            assert np.array_equal(r["best"][i], ref["best"]), f"{label} graph {i} vs {name}"
    # the batch runs again from its resident level 0 with the same result
    r2 = B.louvain(L.params(**gp))
    assert np.array_equal(r2["modularity"], r["modularity"])
    assert all(np.array_equal(x, y) for x, y in zip(r2["best"], r["best"]))


def test_batch_of_larger_graphs_equals_oracle():
    """Twelve Barabasi-Albert graphs of 20 K-200 K vertices (the batched
    configuration's upper range: rows in every LDS bucket, hubs, giant
    communities at the contraction) as one batch, graph by graph against the
    oracle."""
    rng = np.random.default_rng(404)
    graphs = []
    for i in range(12):
        n = int(np.exp(rng.uniform(np.log(20_000), np.log(200_000))))
        graphs.append(L.gen_ba(n, int(rng.integers(2, 9)), seed=int(rng.integers(1 << 30))))
    gs = [L.Graph.from_edges(a, b) for a, b in graphs]
    say(f"larger batch: n={sum(g.n for g in gs)}, E={sum(g.E for g in gs)}")
    r = L.Batch(gs).louvain(L.params(mode=L.HASH, autotune=1))
    for i, (a, b) in enumerate(graphs):
        orc = O.Graph.from_edges(a, b).louvain(O.params(mode=O.HASH))
        got = (r["modularity"][i], int(r["n_comm"][i]), int(r["levels"][i]), int(r["sweeps"][i]))
        exp = (orc["modularity"], orc["n_comm"], orc["levels"], orc["total_sweeps"])
        assert got == exp, f"graph {i}: {got} != {exp}"
        assert np.array_equal(r["best"][i], orc["best"]), f"graph {i}"


def test_batch_of_one_and_identical_copies():
    ea, eb = L.gen_ba(3000, 3, seed=5)
    g = L.Graph.from_edges(ea, eb)
    alone = g.louvain(L.params())
    r1 = L.Batch([g]).louvain(L.params())
    assert r1["modularity"][0] == alone["modularity"]
    assert np.array_equal(r1["best"][0], alone["best"])
    r = L.Batch([g] * 7).louvain(L.params())
    for i in range(7):
        assert r["modularity"][i] == alone["modularity"]
        assert np.array_equal(r["best"][i], alone["best"])
