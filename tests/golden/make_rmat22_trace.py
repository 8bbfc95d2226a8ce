"""Writes the oracle's sweep traces of the benchmark workload (R-MAT scale
22, edge factor 16, (0.57, 0.19, 0.19, 0.05), seed 2, HASH mode, thresholds
0.001) for the GPU tests that run the same graph at full size:

  rmat22_hash_trace.json      exact int64 gains, lowest-id ties (the default);
  rmat22_hash_ref_trace.json  the reference's own method: f32-rounded delta-Q
                              (operatorsDelta.h:19-20) and highest-id ties in
                              hash sweeps (phaseOptimizationHash.cu:23-35).

The oracle needs about a minute per trace, too long for a test; this script
is run once and its output committed.

    python tests/golden/make_rmat22_trace.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "louvain-fast-move-cuda_amd"))

import louvain_amd as L          # host-side generator and CSR only
from oracle import oracle as O

KEYS = ["level", "sweep", "kind", "n", "n_active", "n_moved", "n_considered", "E", "e_active",
        "sigma_in", "sum_tot_sq", "q", "comm_hash"]


RUNS = [("rmat22_hash_trace.json", "exact gains, lowest-id ties",
         dict(gain=O.GAIN_EXACT, tie=O.TIE_LOWEST)),
        ("rmat22_hash_ref_trace.json", "the reference's f32-rounded gains and hash-sweep "
         "highest-id ties", dict(gain=O.GAIN_REF_F32, tie=O.TIE_REFERENCE))]


def main():
    ea, eb, n = L.gen_rmat(22, 16, 0.57, 0.19, 0.19, seed=2)
    g = O.Graph.from_edges(ea, eb)
    only = sys.argv[1:]
    for name, desc, kw in RUNS:
        if only and name not in only:
            continue
        write(g, name, desc, O.params(mode=O.HASH, **kw))


def write(g, name, desc, prm):
    r = g.louvain(prm)
    out = {
        "workload": "R-MAT scale 22, edge factor 16, (0.57,0.19,0.19,0.05), seed 2, "
                    f"symmetrised, unit weights; HASH, {desc}, thresholds 0.001",
        "n": int(g.n),
        "modularity": r["modularity"],
        "n_comm": int(r["n_comm"]),
        "levels": int(r["levels"]),
        "total_sweeps": len(r["sweeps"]),
        "sweeps": [{k: s[k] for k in KEYS} for s in r["sweeps"]],
        "best_sha256": hashlib.sha256(np.ascontiguousarray(r["best"], np.uint32).tobytes()).hexdigest(),
    }
    with open(os.path.join(HERE, name), "w") as f:
        json.dump(out, f, indent=1)
    print(name, out["modularity"], out["n_comm"], out["levels"], out["total_sweeps"])


if __name__ == "__main__":
    main()
