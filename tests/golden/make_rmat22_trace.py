"""Writes rmat22_hash_trace.json: the oracle's sweep trace of the benchmark
workload (R-MAT scale 22, edge factor 16, (0.57, 0.19, 0.19, 0.05), seed 2,
HASH mode, exact gains, lowest-id ties, thresholds 0.001), for the GPU test
that runs the same graph at full size.  The oracle needs about a minute for
it, too long for a test; this script is run once and its output committed.

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


def main():
    ea, eb, n = L.gen_rmat(22, 16, 0.57, 0.19, 0.19, seed=2)
    g = O.Graph.from_edges(ea, eb)
    r = g.louvain(O.params(mode=O.HASH))
    out = {
        "workload": "R-MAT scale 22, edge factor 16, (0.57,0.19,0.19,0.05), seed 2, "
                    "symmetrised, unit weights; HASH, exact gains, lowest-id ties, "
                    "thresholds 0.001",
        "n": int(g.n),
        "modularity": r["modularity"],
        "n_comm": int(r["n_comm"]),
        "levels": int(r["levels"]),
        "total_sweeps": len(r["sweeps"]),
        "sweeps": [{k: s[k] for k in KEYS} for s in r["sweeps"]],
        "best_sha256": hashlib.sha256(np.ascontiguousarray(r["best"], np.uint32).tobytes()).hexdigest(),
    }
    with open(os.path.join(HERE, "rmat22_hash_trace.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out["modularity"], out["n_comm"], out["levels"], out["total_sweeps"])


if __name__ == "__main__":
    main()
