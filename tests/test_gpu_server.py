"""Served Louvain: jobs submitted while others run, admitted by HBM budget.

Bar: every job's result (modularity, community count, levels, sweeps, best
partition) is bit-identical to the same graph run alone and to the oracle,
whatever the mix it ran in and whenever it was admitted.
"""
import sys

import numpy as np
import pytest

import louvain_amd as L
from oracle import oracle as O
from test_gpu_batch import _mixed_graphs

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if L.device_count() < 1:
        pytest.skip("no GPU visible")


@pytest.fixture(scope="module")
def jobs():
    return _mixed_graphs(23, 16)


def _drain(S, limit=10000):
    for _ in range(limit):
        a, q, _ = S.step()
        if a == 0 and q == 0:
            return
    raise AssertionError("server did not drain")


def _check(S, ids, graphs, edges, gp, op, label):
    done = set()
    for i, jid in enumerate(ids):
        r = S.result(jid)
        assert r["status"] == 0
        alone = graphs[i].louvain(L.params(**gp))
        a, b, w = edges[i]
        orc = O.Graph.from_edges(a, b, w).louvain(O.params(**op))
        got = (r["modularity"], r["n_comm"], r["levels"], r["sweeps"])
        for ref, name in ((alone, "alone"), (orc, "oracle")):
            exp = (ref["modularity"], ref["n_comm"], ref["levels"], ref["total_sweeps"])
            assert got == exp, f"{label} job {i} vs {name}: {got} != {exp}"
            assert np.array_equal(r["best"], ref["best"]), f"{label} job {i} vs {name}"
        done.add(jid)
    return done


@pytest.mark.parametrize("label,gp,op", [
    ("hash", dict(mode=L.HASH), dict(mode=O.HASH)),
    ("sort f32 reference 1e-6",
     dict(mode=L.SORT, gain=L.GAIN_REF_F32, tie=L.TIE_REFERENCE, sweep_threshold=1e-6,
          level_threshold=1e-6),
     dict(mode=O.SORT, gain=O.GAIN_REF_F32, tie=O.TIE_REFERENCE, sweep_threshold=1e-6,
          level_threshold=1e-6)),
])
def test_staggered_submissions_equal_alone(jobs, label, gp, op):
    graphs = [L.Graph.from_edges(a, b, w) for a, b, w in jobs]
    S = L.Server(L.params(**gp))
    ids = []
    # jobs arrive while others are at later levels
    plan = [3, 0, 4, 1, 0, 5, 3]
    k = 0
    for burst in plan:
        for _ in range(burst):
            ids.append(S.submit(graphs[k]))
            k += 1
        a, q, r = S.step()
        say(f"server {label}: active {a} queued {q} retired {r}")
    _drain(S)
    fin = set()
    while True:
        got = S.take_finished(5)
        if not got:
            break
        fin.update(got)
    assert fin == set(ids)
    _check(S, ids, graphs, jobs, gp, op, label)


def test_budget_admits_one_at_a_time():
    ea, eb = L.gen_ba(3000, 5, seed=3)
    g = L.Graph.from_edges(ea, eb)
    alone = g.louvain(L.params())
    S = L.Server(L.params(), hbm_budget=int(1.5 * L.footprint(g)))
    ids = [S.submit(g) for _ in range(5)]
    seen = 0
    for _ in range(10000):
        a, q, _ = S.step()
        assert a <= 1
        seen = max(seen, a)
        if a == 0 and q == 0:
            break
    assert seen == 1
    for jid in ids:
        r = S.result(jid)
        assert (r["modularity"], r["n_comm"]) == (alone["modularity"], alone["n_comm"])
        assert np.array_equal(r["best"], alone["best"])


def test_tight_budget_mixed_jobs(jobs):
    graphs = [L.Graph.from_edges(a, b, w) for a, b, w in jobs[:6]]
    S = L.Server(L.params(), hbm_budget=max(L.footprint(g) for g in graphs))
    ids = [S.submit(g) for g in graphs]
    _drain(S)
    _check(S, ids, graphs, jobs[:6], {}, {}, "tight budget")


def test_same_graph_many_times_and_errors():
    ea, eb = L.gen_ba(2000, 4, seed=9)
    g = L.Graph.from_edges(ea, eb)
    alone = g.louvain(L.params())
    S = L.Server(L.params())
    ids = [S.submit(g) for _ in range(9)]
    _drain(S)
    for jid in ids:
        r = S.result(jid)
        assert r["modularity"] == alone["modularity"]
        assert np.array_equal(r["best"], alone["best"])
    # unknown / already released job
    with pytest.raises(L.LouvainError) as e:
        S.result(ids[0])
    assert e.value.status == 9                        # LV_ERR_STATE
    # a job larger than the budget is refused at submission
    small = L.Server(L.params(), hbm_budget=1 << 20)
    with pytest.raises(L.LouvainError) as e:
        small.submit(g)
    assert e.value.status == 6                        # LV_ERR_RANGE
    # 2m >= 2^31 is refused at submission
    big = L.Graph.from_csr(np.array([0, 1, 2], np.uint64), np.array([1, 0], np.uint32),
                           np.array([1 << 30, 1 << 30], np.uint32))
    with pytest.raises(L.LouvainError) as e:
        S.submit(big)
    assert e.value.status == 6


def test_served_mix_with_rmat_jobs():
    """The served configuration's mix in small: Barabasi-Albert jobs of the
    batched configuration's distribution beside R-MAT jobs (scales 16-17,
    edge factor 16, (0.57, 0.19, 0.19, 0.05): stand-ins for the 5% of R-MAT-20
    requests, with hub rows above 4096 entries), submitted in waves while
    earlier jobs are at later levels; each job's result equals the oracle's."""
    rng = np.random.default_rng(55)
    jobs = []
    for i in range(10):
        if i in (2, 7):
            ea, eb, _ = L.gen_rmat(16 + (i == 7), 16, 0.57, 0.19, 0.19, seed=500 + i)
        else:
            n = int(np.exp(rng.uniform(np.log(200), np.log(200_000))))
            ea, eb = L.gen_ba(n, int(rng.integers(2, 9)), seed=5_000_000 + i)
        jobs.append((ea, eb, None))
    graphs = [L.Graph.from_edges(a, b, w) for a, b, w in jobs]
    say("served mix: " + ", ".join(f"n={g.n} E={g.E}" for g in graphs))
    S = L.Server(L.params())
    ids = []
    for k, g in enumerate(graphs):
        ids.append(S.submit(g))
        if k % 3 == 2:
            S.step()
    _drain(S)
    _check(S, ids, graphs, jobs, {}, {}, "served mix")
