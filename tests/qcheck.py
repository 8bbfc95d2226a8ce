"""FP64 modularity of a partition, recomputed from scratch with numpy.

Q = sum_c [ in_c / (2m) - (tot_c / 2m)^2 ] over the CSR's directed entries
(a self-loop's entries count as in the CSR, i.e. twice per loop edge), the
definition `community.h:104` evaluates.  Used to check the library's and the
oracle's modularity independently of both.
"""
import numpy as np


def modularity_fp64(row, col, w, part):
    row = np.asarray(row, np.int64)
    col = np.asarray(col, np.int64)
    part = np.asarray(part, np.int64)
    n = len(row) - 1
    wt = np.ones(len(col), np.float64) if w is None else np.asarray(w, np.float64)
    src = np.repeat(np.arange(n), np.diff(row))
    two_m = wt.sum()
    same = part[src] == part[col]
    sigma_in = wt[same].sum()
    k = np.bincount(src, weights=wt, minlength=n)
    tot = np.bincount(part, weights=k)
    return float(sigma_in / two_m - np.dot(tot, tot) / (two_m * two_m))
