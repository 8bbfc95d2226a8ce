"""GPU parity on the measured workloads' own generators.

- Barabasi-Albert, n = 10,000, 2 edges per new vertex, seed 1 (the CPU-path
  workload): every mode / gain, thresholds 1e-6 and 1e-3, trace for trace.
- Planted partition, n = 1,000,000, power-law degrees (exponent 2.5, mean 20,
  max 5,000), 2,000 communities (sizes exponent 1.5 in [50, 5000]), mixing
  0.3, integer weights 1..10, seed 3, at full size: trace for trace against
  the oracle, final Q against an FP64 numpy recomputation (1e-6 absolute),
  and the planted partition's Q reported beside it.
"""
import itertools
import sys

import numpy as np
import pytest

import louvain_amd as L
from oracle import oracle as O
from qcheck import modularity_fp64
from test_gpu_parity import GAINS, MODES, compare_runs

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if L.device_count() < 1:
        pytest.skip("no GPU visible")


@pytest.mark.parametrize("mode,gain", list(itertools.product(MODES, GAINS)), ids=lambda x: x[0])
def test_ba_10000(mode, gain):
    ea, eb = L.gen_ba(10000, 2, seed=1)
    g = L.Graph.from_edges(ea, eb)
    og = O.Graph.from_edges(ea, eb)
    row, col, _ = g.csr()[:3]
    for thr in (1e-6, 1e-3):
        label = f"BA-10000 {mode[0]} {gain[0]} thr={thr}"
        say(label)
        a = g.louvain(L.params(mode=mode[1], gain=gain[1], sweep_threshold=thr,
                               level_threshold=thr), hash_trace=True)
        b = og.louvain(O.params(mode=mode[2], gain=gain[2], sweep_threshold=thr,
                                level_threshold=thr))
        compare_runs(a, b, label)
        assert abs(a["modularity"] - modularity_fp64(row, col, None, a["best"])) < 1e-6


@pytest.fixture(scope="module")
def planted():
    say("generating planted partition n=1e6 seed 3")
    ea, eb, ew, lab = L.gen_planted()
    return ea, eb, ew, lab


@pytest.mark.parametrize("mode", ["hash", "sort"])
def test_planted_1m(planted, mode):
    ea, eb, ew, lab = planted
    lm, om = {"hash": (L.HASH, O.HASH), "sort": (L.SORT, O.SORT)}[mode]
    g = L.Graph.from_edges(ea, eb, ew)
    row, col, w = g.csr()[:3]
    say(f"planted 1e6 {mode}: GPU run")
    a = g.louvain(L.params(mode=lm), hash_trace=True)
    say(f"planted 1e6 {mode}: oracle run")
    b = O.Graph.from_edges(ea, eb, ew).louvain(O.params(mode=om))
    compare_runs(a, b, f"planted-1e6 {mode}")
    q64 = modularity_fp64(row, col, w, a["best"])
    q_planted = modularity_fp64(row, col, w, lab)
    say(f"planted 1e6 {mode}: Q={a['modularity']:.9f} fp64={q64:.9f} planted={q_planted:.9f} "
        f"communities={a['n_comm']}")
    assert abs(a["modularity"] - q64) < 1e-6
    assert a["modularity"] > 0.9 * q_planted


def test_planted_1m_reference_method(planted):
    """The reference's own method on the planted configuration (f32-rounded
    delta-Q, operatorsDelta.h:19-20, and highest-id ties in hash sweeps,
    phaseOptimizationHash.cu:23-35): the GPU against the oracle trace for
    trace, and the default exact-gain run's final Q within 0.5% of it."""
    ea, eb, ew, lab = planted
    g = L.Graph.from_edges(ea, eb, ew)
    row, col, w = g.csr()[:3]
    say("planted 1e6 reference method: GPU run")
    a = g.louvain(L.params(mode=L.HASH, gain=L.GAIN_REF_F32, tie=L.TIE_REFERENCE),
                  hash_trace=True)
    say("planted 1e6 reference method: oracle run")
    b = O.Graph.from_edges(ea, eb, ew).louvain(
        O.params(mode=O.HASH, gain=O.GAIN_REF_F32, tie=O.TIE_REFERENCE))
    compare_runs(a, b, "planted-1e6 reference method")
    q_ref = a["modularity"]
    assert abs(q_ref - modularity_fp64(row, col, w, a["best"])) < 1e-6
    q_exact = g.louvain(L.params(mode=L.HASH))["modularity"]
    say(f"planted 1e6: Q exact {q_exact:.9f}, Q reference method {q_ref:.9f}, "
        f"relative difference {(q_exact - q_ref) / abs(q_ref):+.3%}")
    assert abs(q_exact - q_ref) <= 0.005 * abs(q_ref)


def test_gather_probe_rates_are_sane(planted):
    """lv_gather_probe on the planted 1e6 graph: both widths give a finite
    positive rate below one gather per byte of the 8 TB/s peak, and widths
    other than 4 and 8 are refused."""
    import ctypes as C
    ea, eb, w, _ = planted
    c = L.Community(L.Graph.from_edges(ea, eb, w))
    g4, g8 = c.gather_probe(4), c.gather_probe(8)
    say(f"gather probe planted-1e6: {g4 / 1e9:.1f} G/s (4 B), {g8 / 1e9:.1f} G/s (8 B)")
    for g in (g4, g8):
        assert np.isfinite(g) and 1e9 < g < 8e12
    out = C.c_double()
    assert L.lib.lv_gather_probe(c.h, 3, 1, C.byref(out)) == 1
