"""Step-by-step numpy emulation of the reference's Thrust pipeline.

Test infrastructure: an independent second restatement of the reference
method, written as the reference's own sequence of bulk primitives
(copy_if -> sort_by_key -> reduce_by_key -> transform(delta-Q) ->
reduce_by_key(GetMaxValue) -> update kernels -> update_changed_kernel ->
compute_modularity -> aggregation), so that it shares no code path with the C
oracle's per-vertex dense accumulator.  Both must produce identical traces.

References (paths relative to the reference repo):
  phaseOptimizationSort.cu:28-158 (fast), :161-379 (sort)
  phaseOptimizationHash.cu:23-162, hashmap.cuh:46-73 (hash)
  operatorsDelta.h:7-45 (delta-Q), operators.h:30-40 (GetMaxValue)
  phaseOptimizationUpdate.cu:14-132 (move rules, marking)
  community.h:68-106 (modularity), phaseAggregation.cuh:19-175 (aggregation)
"""
import numpy as np

FAST, SORT, HASH = 0, 1, 2


def csr_to_coo(row, col, w):
    n = len(row) - 1
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(row.astype(np.int64)))
    return src, col.astype(np.int64), w.astype(np.int64)


def modularity(src, dst, w, k, comm, tw):
    """community.h:68-106."""
    n = len(k)
    tot = np.bincount(comm, weights=k.astype(np.float64), minlength=n)
    sq = np.square(tot)
    sigma_in = int(w[comm[src] == comm[dst]].sum())
    s = int(np.rint(tot).astype(np.int64).__pow__(2).sum())
    return sigma_in / (2 * tw) - float(s) / (4 * tw * tw), sigma_in, s, sq


def dq_f32(wsum, self_w, k, tot_own, tot_c, tw):
    """operatorsDelta.h:19-20 / :38-39 with numpy dtypes matching C++ promotion."""
    a = (wsum.astype(np.float32) - self_w.astype(np.float32)).astype(np.float32)
    t1 = a.astype(np.float64) / tw
    t2 = (k.astype(np.float64) * ((tot_own.astype(np.float64) - k.astype(np.float64))
                                   - tot_c.astype(np.float64))) / (2 * tw * tw)
    return (t1 + t2).astype(np.float32)


def dq_exact(wsum, self_w, k, tot_own, tot_c, m2):
    return m2 * (wsum - self_w) + k * (tot_own - k - tot_c)


def sweep(kind, src, dst, w, k, comm, tot, active, tw, gain="exact", tie="lowest"):
    """One optimisation sweep + fast_move_update.  Returns (comm', tot', active', moved)."""
    n = len(k)
    m2 = int(round(2 * tw))
    # copy_if with TestTupleValue (operatorsCommunity.h:7-19)
    sel = active[src].astype(bool) & (src != dst)
    s = src[sel]
    c = comm[dst[sel]]
    ws = w[sel]
    if kind == FAST:
        # no reduction: one candidate per entry, self = 0 (phaseOptimizationSort.cu:98)
        ks, kc, kw = s, c, ws
        self_v = np.zeros(n, np.int64)
    elif kind == SORT:
        order = np.lexsort((kc_ := c, s))
        s2, c2, w2 = s[order], kc_[order], ws[order]
        if len(s2):
            brk = np.ones(len(s2), bool)
            brk[1:] = (s2[1:] != s2[:-1]) | (c2[1:] != c2[:-1])
            starts = np.flatnonzero(brk)
            ks, kc = s2[starts], c2[starts]
            kw = np.add.reduceat(w2, starts)
        else:
            ks = kc = kw = np.zeros(0, np.int64)
        # select_self_node (phaseOptimizationSort.cu:9-25)
        self_v = np.zeros(n, np.int64)
        own = kc == comm[ks]
        self_v[ks[own]] = kw[own]
    else:  # HASH: own-community weight goes to self, others into the table
        own_e = c == comm[s]
        self_v = np.bincount(s[own_e], weights=ws[own_e], minlength=n).astype(np.int64)
        keys = (s[~own_e] << 32) | c[~own_e]
        uk, inv = np.unique(keys, return_inverse=True)
        kw = np.bincount(inv, weights=ws[~own_e], minlength=len(uk)).astype(np.int64)
        ks, kc = uk >> 32, uk & 0xFFFFFFFF

    own_c = comm[ks]
    if gain == "exact":
        val = dq_exact(kw, self_v[ks], k[ks], tot[own_c], tot[kc], m2)
    else:
        val = dq_f32(kw, self_v[ks], k[ks], tot[own_c], tot[kc], tw)

    target = comm.copy()
    if len(ks):
        if kind == HASH:
            # update_best_kernel: atomicMax over (bits(dq) << 32 | c) for dq > 0
            pos = val > 0
            vs, vc, vv = ks[pos], kc[pos], val[pos]
            if len(vs):
                hi = tie == "reference"
                order = np.lexsort(((vc if hi else -vc), vv, vs))
                last = np.ones(len(order), bool)
                last[:-1] = vs[order][1:] != vs[order][:-1]
                pick = order[last]
                bv, bc, bval = vs[pick], vc[pick], vv[pick]
                mv = (comm[bv] != bc) & (bval > 0)
                target[bv[mv]] = bc[mv]
        else:
            # reduce_by_key with GetMaxValue: keep the earlier element on ties
            # (candidates are ordered by community within each vertex)
            order = np.lexsort((kc, ks))
            ks_o, kc_o, v_o = ks[order], kc[order], val[order]
            brk = np.ones(len(ks_o), bool)
            brk[1:] = ks_o[1:] != ks_o[:-1]
            starts = np.flatnonzero(brk)
            gmax = np.maximum.reduceat(v_o, starts)
            seg = np.cumsum(brk) - 1
            is_max = v_o == gmax[seg]
            # first max in each segment
            first = np.full(len(starts), -1, np.int64)
            idx = np.flatnonzero(is_max)
            seg_idx = seg[idx]
            uniq, fpos = np.unique(seg_idx, return_index=True)
            first[uniq] = idx[fpos]
            bv, bc, bval = ks_o[first], kc_o[first], v_o[first]
            mv = (comm[bv] != bc) & (bval > 0)
            if kind == FAST:
                mv &= ~(bv > bc)
            target[bv[mv]] = bc[mv]
    moved = target != comm
    tot = tot.copy()
    np.add.at(tot, target[moved], k[moved])
    np.subtract.at(tot, comm[moved], k[moved])
    comm = target
    # update_changed_kernel (phaseOptimizationUpdate.cu:117-132)
    hit = moved[dst] & (comm[dst] != comm[src])
    active = np.zeros(n, np.uint8)
    active[src[hit]] = 1
    return comm, tot, active, int(moved.sum()), int(sel.sum())


def aggregate(src, dst, w, k, comm):
    """phaseAggregation.cuh:19-175 (sort form)."""
    n = len(k)
    tot = np.bincount(comm, weights=k.astype(np.float64), minlength=n)
    cmap = np.cumsum(tot > 0).astype(np.int64)          # inclusive scan of CountNotZero
    ns, nd = cmap[comm[src]] - 1, cmap[comm[dst]] - 1   # MakeCommunityPair
    key = (ns << 32) | nd
    uk, inv = np.unique(key, return_inverse=True)
    nw = np.bincount(inv, weights=w, minlength=len(uk)).astype(np.int64)
    n_new = int(cmap[-1])
    nsrc, ndst = uk >> 32, uk & 0xFFFFFFFF
    nk = np.bincount(nsrc, weights=nw, minlength=n_new).astype(np.int64)
    return cmap - 1, n_new, nsrc, ndst, nw, nk


def louvain(row, col, w, tw, mode="hash", gain="exact", tie="lowest",
            sweep_thr=0.001, level_thr=0.001, refine=False):
    """modularityAlgorithms.h:14-81 with phaseOptimization.cuh:116-181.
    refine: afterwards, the partition on the level-0 graph swept again from
    sweep index 1 (no fast sweep) and renumbered (the opt-in refinement)."""
    src, dst, ww = csr_to_coo(np.asarray(row), np.asarray(col), np.asarray(w))
    src0, dst0, ww0 = src, dst, ww
    n0 = len(row) - 1
    k = np.bincount(src, weights=ww, minlength=n0).astype(np.int64)
    k0 = k
    best = np.arange(n0, dtype=np.int64)
    trace = []
    level = 0
    while True:
        n = len(k)
        comm = np.arange(n, dtype=np.int64)
        tot = k.copy()
        q0 = modularity(src, dst, ww, k, comm, tw)[0]
        q = q0
        active = np.ones(n, np.uint8)
        ex = 0
        while True:
            old = q
            if ex == 0:
                kind = FAST
            elif mode == "hash":
                kind = HASH
            else:
                kind = SORT
            comm, tot, active, moved, sel = sweep(kind, src, dst, ww, k, comm, tot, active, tw,
                                                  gain, tie)
            ex += 1
            q, si, s, _ = modularity(src, dst, ww, k, comm, tw)
            trace.append((level, ex - 1, kind, moved, sel, si, s, q))
            if not (q - old > sweep_thr):
                break
        cont = q - q0 > level_thr
        cmap, n_new, nsrc, ndst, nw, nk = aggregate(src, dst, ww, k, comm)
        best = cmap[comm[best]]
        if not cont:
            break
        src, dst, ww, k = nsrc, ndst, nw, nk
        level += 1
    if refine:
        level += 1
        comm = best.copy()
        tot = np.bincount(comm, weights=k0, minlength=n0).astype(np.int64)
        q = modularity(src0, dst0, ww0, k0, comm, tw)[0]
        active = np.ones(n0, np.uint8)
        ex = 1
        while True:
            old = q
            kind = HASH if mode == "hash" else SORT
            comm, tot, active, moved, sel = sweep(kind, src0, dst0, ww0, k0, comm, tot, active,
                                                  tw, gain, tie)
            ex += 1
            q, si, s, _ = modularity(src0, dst0, ww0, k0, comm, tw)
            trace.append((level, ex - 1, kind, moved, sel, si, s, q))
            if not (q - old > sweep_thr):
                break
        cmap, n_new, _, _, _, _ = aggregate(src0, dst0, ww0, k0, comm)
        best = cmap[comm]
    return {"modularity": q, "best": best, "n_comm": n_new, "trace": trace}
