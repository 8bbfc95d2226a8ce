"""Benchmark: plain Louvain on R-MAT scale 22 on one MI355X.

Workload (the single-GPU configuration): R-MAT scale 22, edge factor 16,
(a,b,c,d) = (0.57,0.19,0.19,0.05), seed 2, symmetrised, self-loops and
duplicates dropped, ids of non-isolated vertices compacted; unit weights.
One step = one full Louvain run to convergence (all levels, the reference's
0.001 thresholds, HASH mode) starting from the CSR already resident in HBM.

value = traversed directed entries per second over the whole run (sum over
sweeps of the entries of the vertices examined, divided by the step time).
The roofline block covers the dominant best-move kernel, timed with HIP
events on the stream it runs on, and the whole best-move pass, and sets
the pass beside the measured random-gather ceiling of the same graph
(`gather_ceiling`: a kernel that does nothing but the pass's dependent
neighbour-word gather); the CPU
baseline is the single-threaded oracle run to convergence on the same CSR.
`per_sweep` gives traversed entries per second of the best-move pass for
sweep 0 of level 0 and the entry-weighted mean over all sweeps; `buckets`
gives each degree bucket's achieved bandwidth from one extra run with the
pass on a single stream; `reference_method` is the same graph run with the
reference's f32-rounded gains and hash-sweep highest-id ties; `c1` is the
CPU-path configuration (the bundled 10,000-vertex fixture and a
Barabasi-Albert graph of the same size) on the oracle.  The `batched` sub-object
is the batched configuration (4,096 Barabasi-Albert graphs in one
disjoint-union launch sequence, graphs/s, against a loop of per-graph runs);
`--workload batched` prints it as its own line with value = graphs/s.  The
`served` sub-object is the served configuration (Poisson arrivals at 200
req/s, continuous batching, p50/p99 latency, graphs/s) over a 60 s arrival
window; `--workload served` prints it as its own line.
"""
import argparse
import json
import os
import platform
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "louvain-fast-move-cuda_amd"))
sys.path.insert(0, REPO)

METRIC = "Louvain s-to-converge + edges/s vs HBM roofline; batched graphs/s, p99 lat"
HBM_PEAK_GBS = 8000.0
FIXTURE = os.path.join(REPO, "tests", "golden", "graph-power-law-10000-2-99-1.edge")
BUCKET_NAMES = ["deg<=4", "deg<=8", "deg<=16", "deg<=32", "deg<=64", "deg<=256", "deg<=1024",
                "deg<=4096", "hub"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(row, col, mode, threshold):
    """The single-threaded oracle run to convergence on the same CSR (all
    levels, the same mode and thresholds): the CPU path's s-to-converge."""
    from oracle import oracle as O
    og = O.Graph.from_csr(row, col)
    mode_id = {"hash": O.HASH, "sort": O.SORT, "adaptive_speed": O.ADAPTIVE_SPEED,
               "adaptive_memory": O.ADAPTIVE_MEMORY}[mode]
    t0 = time.perf_counter()
    r = og.louvain(O.params(mode=mode_id, sweep_threshold=threshold, level_threshold=threshold))
    dt = time.perf_counter() - t0
    ent = sum(s["e_active"] for s in r["sweeps"])
    return {"value": ent / dt, "unit": "edges/s", "cores": 1, "kind": "port",
            "s_to_converge": dt, "modularity": r["modularity"], "n_communities": r["n_comm"],
            "sample": f"oracle (C, -O3, 1 thread) run to convergence on the same R-MAT CSR "
                      f"({r['levels']} levels, {len(r['sweeps'])} sweeps, {mode}, exact gains, "
                      f"thresholds {threshold:g}): {ent} traversed entries in {dt:.2f} s; "
                      f"CPU {cpu_model()}, nproc {os.cpu_count()}"}


def extra_runs(args, L, c, info, q_ex, bounds):
    """Runs outside the timed steps: the reference's own method on the same
    resident graph, and one run with the best-move pass on a single stream so
    that every degree bucket is timed alone."""
    # the reference's own method on the same resident graph
    prm_ref = L.params(mode=args.mode, sweep_threshold=args.threshold,
                       level_threshold=args.threshold, gain=L.GAIN_REF_F32, tie=L.TIE_REFERENCE,
                       autotune=0)
    c.reset()
    c.run(prm_ref, best=False, n0=info["n"])
    c.reset()
    t0 = time.perf_counter()
    rr = c.run(prm_ref, best=False, n0=info["n"])
    ref_s = time.perf_counter() - t0
    q_rf = rr["modularity"]
    reference_method = {
        "modularity": q_rf, "n_communities": rr["n_comm"], "levels": rr["levels"],
        "sweeps": rr["total_sweeps"], "s_to_converge": ref_s,
        "relative_q_difference": (q_ex - q_rf) / abs(q_rf),
        "within_half_percent_below": q_ex >= (1 - 0.005) * q_rf,
        "within_half_percent_two_sided": abs(q_ex - q_rf) <= 0.005 * abs(q_rf),
        "note": "gain=ref_f32 (operatorsDelta.h:19-20) with highest-id ties in hash sweeps "
                "(phaseOptimizationHash.cu:23-35); relative_q_difference = (Q_default - Q_ref) "
                "/ Q_ref"}
    log(f"reference method: Q={q_rf:.9f} (default {q_ex:.9f}), {ref_s * 1e3:.1f} ms")
    # each bucket alone: one run with the pass on a single stream, every bucket timed
    prm3 = L.params(mode=args.mode, sweep_threshold=args.threshold,
                    level_threshold=args.threshold, profile=3, autotune=0,
                    bucket_max_deg=bounds)   # the bounds of the timed runs
    c.reset()
    c.run(prm3, best=False, n0=info["n"])
    kb = c.kernel_stats()
    buckets = []
    for b in range(9):
        for f in (1, 0):
            i = 2 * b + f
            if not kb["bucket_launches"][i]:
                continue
            ms = kb["bucket_ms"][i]
            gbs = kb["bucket_bytes"][i] / ms / 1e6 if ms > 0 else 0.0
            buckets.append({"bucket": BUCKET_NAMES[b], "sweep": "fast" if f else args.mode,
                            "launches": kb["bucket_launches"][i], "ms": ms,
                            "entries": kb["bucket_entries"][i], "bytes": kb["bucket_bytes"][i],
                            "achieved_gbs": gbs, "frac": gbs / HBM_PEAK_GBS})
    log("buckets: " + ", ".join(f"{x['bucket']}/{x['sweep']} {x['ms']:.2f} ms {x['achieved_gbs']:.0f} GB/s"
                                 for x in buckets))
    return reference_method, buckets


def c1_result(L):
    """The CPU-path configuration: the bundled fixture (networkx
    powerlaw_cluster_graph(10000, 2, 0.99, seed=1)) and a literal
    Barabasi-Albert graph (n = 10,000, m = 2, seed 1), each run to convergence
    on the oracle at thresholds 1e-6 and 0.001, with the GPU's Q beside it."""
    from oracle import oracle as O
    ea, eb = L.gen_ba(10000, 2, seed=1)
    cases = [("fixture graph-power-law-10000-2-99-1.edge", O.Graph.load_edge(FIXTURE),
              L.Graph.load_edge(FIXTURE)),
             ("Barabasi-Albert n=10000 m=2 seed 1", O.Graph.from_edges(ea, eb),
              L.Graph.from_edges(ea, eb))]
    out = []
    for name, og, g in cases:
        for thr in (1e-6, 1e-3):
            t0 = time.perf_counter()
            r = og.louvain(O.params(sweep_threshold=thr, level_threshold=thr))
            dt = time.perf_counter() - t0
            c = L.Community(g)
            prm = L.params(sweep_threshold=thr, level_threshold=thr)
            c.reset()
            c.run(prm, best=False, n0=g.n)               # warm
            c.reset()
            rg = c.run(prm, best=False, n0=g.n)
            c.close()
            out.append({"graph": name, "threshold": thr, "cpu_s_to_converge": dt,
                        "modularity": r["modularity"], "n_communities": r["n_comm"],
                        "levels": r["levels"], "sweeps": len(r["sweeps"]),
                        "gpu_s_to_converge": rg["seconds"],
                        "gpu_same_modularity": rg["modularity"] == r["modularity"]})
    return out


def modularity_fp64(row, col, w, part):
    """Q of a partition recomputed in FP64 with numpy (community.h:104's definition)."""
    import numpy as np
    row = np.asarray(row, np.int64)
    part = np.asarray(part, np.int64)
    n = len(row) - 1
    wt = np.ones(len(col), np.float64) if w is None else np.asarray(w, np.float64)
    src = np.repeat(np.arange(n), np.diff(row))
    two_m = wt.sum()
    sig = wt[part[src] == part[np.asarray(col, np.int64)]].sum()
    tot = np.bincount(part, weights=np.bincount(src, weights=wt, minlength=n))
    return float(sig / two_m - np.dot(tot, tot) / (two_m * two_m))


def batched_graphs(L, count, seed, n_lo=200, n_hi=200_000, m_lo=2, m_hi=8):
    """Barabasi-Albert graphs, n log-uniform in [n_lo, n_hi], m uniform in
    {m_lo..m_hi}; graph i uses generator seed seed*1e6 + i."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(n_lo), np.log(n_hi), count)).astype(np.int64)
    ms = rng.integers(m_lo, m_hi + 1, count)

    def one(i):
        ea, eb = L.gen_ba(int(ns[i]), int(ms[i]), seed=seed * 1_000_000 + i)
        return L.Graph.from_edges(ea, eb)
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        return list(ex.map(one, range(count)))


def batched_result(args, L):
    """4,096 BA graphs in one batch versus a loop of per-graph runs."""
    import numpy as np
    t = time.perf_counter()
    graphs = batched_graphs(L, args.graphs, 4)
    n_tot = sum(g.n for g in graphs)
    e_tot = sum(g.E for g in graphs)
    log(f"{len(graphs)} BA graphs: n={n_tot} E={e_tot} ({time.perf_counter() - t:.1f}s)")
    prm = L.params(mode=args.mode, sweep_threshold=args.threshold, level_threshold=args.threshold,
                   autotune=args.autotune)
    t = time.perf_counter()
    B = L.Batch(graphs)
    log(f"batch union + upload ({time.perf_counter() - t:.1f}s)")
    for i in range(args.warmup):
        r = B.louvain(prm, best=False)
        log(f"warmup {i}: {r['seconds']:.3f} s")
    times, res = [], None
    for i in range(args.steps):
        t0 = time.perf_counter()
        res = B.louvain(prm, best=(i == args.steps - 1))
        times.append(time.perf_counter() - t0)
        log(f"step {i}: {times[-1]:.3f} s, union levels {res['union_levels']}, "
            f"union sweeps {res['union_sweeps']}, traversed {res['traversed_entries']}")
    batch_s = sum(times) / len(times)
    # the same graphs one after the other, each resident (upload untimed)
    order = list(range(len(graphs)))
    if args.loop_sample > 1:
        order = order[::args.loop_sample]
    loop_s, q_loop, t_lat = 0.0, {}, []
    for k, i in enumerate(order):
        c = L.Community(graphs[i])
        t0 = time.perf_counter()
        c.reset()
        r = c.run(prm, sweeps_cap=1, levels_cap=1, best=False, n0=graphs[i].n)
        dt = time.perf_counter() - t0
        loop_s += dt
        t_lat.append(dt)
        q_loop[i] = r["modularity"]
        c.close()
        if k % 512 == 0:
            log(f"per-graph loop: {k}/{len(order)} graphs, {loop_s:.2f} s")
    loop_gps = len(order) / loop_s
    dq = max(abs(res["modularity"][i] - q_loop[i]) for i in order)
    # FP64 recomputation of Q for a few graphs (the largest and the smallest ones)
    by_size = sorted(range(len(graphs)), key=lambda i: graphs[i].E)
    check = by_size[:2] + by_size[len(by_size) // 2:len(by_size) // 2 + 2] + by_size[-4:]
    dq64 = 0.0
    for i in check:
        row, col, w = graphs[i].csr()[:3]
        dq64 = max(dq64, abs(res["modularity"][i] - modularity_fp64(row, col, None, res["best"][i])))
    gps = len(graphs) / batch_s
    return {
        "graphs_per_s": gps,
        "seconds_per_batch": batch_s,
        "steps": args.steps,
        "config": {"workload": f"batched Louvain: {len(graphs)} Barabasi-Albert graphs, n "
                               f"log-uniform in [200, 200000], m uniform in {{2..8}}, seed 4, "
                               f"unit weights, one disjoint-union launch sequence",
                   "graphs": len(graphs), "n": n_tot, "E": e_tot, "mode": args.mode,
                   "threshold": args.threshold},
        "edges_per_s": res["traversed_entries"] / batch_s,
        "per_graph_loop": {"graphs_per_s": loop_gps, "graphs_timed": len(order),
                           "seconds": loop_s,
                           "note": "each graph resident, reset + run timed, upload not timed"},
        "speedup_vs_loop": gps / loop_gps,
        "max_abs_dq_batch_vs_loop": dq,
        "max_abs_dq_vs_fp64": dq64,
        "fp64_checked_graphs": len(check),
        "mean_q": float(np.mean(res["modularity"])),
        "union_levels": res["union_levels"],
        "union_sweeps": res["union_sweeps"],
    }


def served_result(args, L):
    """Poisson arrivals (rate args.rate for args.window s, seed 5) served with
    continuous batching under an HBM budget; 95% of the requests draw from the
    batched configuration's size/degree distribution (Barabasi-Albert, n
    log-uniform in [200, 200000], m uniform in {2..8}, unit weights; or, with
    --served-mix planted, planted-partition graphs with the planted
    configuration's laws at args.served_n vertices, weights 1..10), 5% are
    R-MAT scale 20.  Requests draw from a pool of distinct seeded graphs; every
    request is its own job (its own upload, device state and result)."""
    import numpy as np
    t = time.perf_counter()
    pool_p = []
    if args.served_mix == "ba":
        rng_p = np.random.default_rng(55)
        ns = np.exp(rng_p.uniform(np.log(200), np.log(200_000), args.pool)).astype(np.int64)
        ms = rng_p.integers(2, 9, args.pool)
        for i in range(args.pool):
            ea, eb = L.gen_ba(int(ns[i]), int(ms[i]), seed=5_000_000 + i)
            pool_p.append(L.Graph.from_edges(ea, eb))
        mix_desc = ("Barabasi-Albert graphs of the batched configuration's distribution (n "
                    "log-uniform in [200, 200000], m uniform in {2..8}, unit weights)")
    else:
        for i in range(args.pool):
            ea, eb, ew, _ = L.gen_planted(n=args.served_n, n_comm=max(2, args.served_n // 500),
                                          seed=5000 + i)
            pool_p.append(L.Graph.from_edges(ea, eb, ew))
        mix_desc = (f"planted-partition graphs (n={args.served_n}, degree exponent 2.5, mean 20, "
                    f"max 5000, community sizes exponent 1.5 in [50, 5000], mixing 0.3, weights "
                    f"1..10)")
    pool_r = []
    for i in range(4):
        ea, eb, _ = L.gen_rmat(20, 16, 0.57, 0.19, 0.19, seed=500 + i)
        pool_r.append(L.Graph.from_edges(ea, eb))
    log(f"served pool: {len(pool_p)} {args.served_mix} (mean E="
        f"{np.mean([g.E for g in pool_p]):.0f}), {len(pool_r)} R-MAT-20 (mean E="
        f"{np.mean([g.E for g in pool_r]):.0f}) ({time.perf_counter() - t:.1f}s)")
    rng = np.random.default_rng(5)
    arr = np.cumsum(rng.exponential(1.0 / args.rate, int(args.rate * args.window * 1.5) + 10))
    arr = arr[arr < args.window]
    is_r = rng.random(len(arr)) < 0.05
    pick = rng.integers(0, 1 << 30, len(arr))
    req = [pool_r[pick[i] % len(pool_r)] if is_r[i] else pool_p[pick[i] % len(pool_p)]
           for i in range(len(arr))]
    rng_b = np.random.default_rng(6)

    def req_pool(r):
        return pool_r[r.integers(len(pool_r))] if r.random() < 0.05 else pool_p[r.integers(len(pool_p))]
    prm = L.params(mode=args.mode, sweep_threshold=args.threshold, level_threshold=args.threshold)
    budget = int(args.hbm_budget_gb * (1 << 30))
    # warm the server's buffers and the code paths with a few jobs
    S = L.Server(prm, hbm_budget=budget)
    for g in pool_p[:2] + pool_r[:1]:
        S.submit(g)
    while S.step()[:2] != (0, 0):
        pass
    for jid in S.take_finished():
        S.result(jid, best=False)
    # open loop in real time
    lat, kind_of, q_of = [], [], {}
    job_req = {}
    start = time.perf_counter()
    i, n_done, steps, max_active = 0, 0, 0, 0
    last_log = start
    while n_done < len(arr):
        now = time.perf_counter() - start
        while i < len(arr) and arr[i] <= now:
            job_req[S.submit(req[i])] = i
            i += 1
        a, q, _ = S.step()
        steps += 1
        max_active = max(max_active, a)
        t_now = time.perf_counter() - start
        for jid in S.take_finished():
            r = S.result(jid, best=False)
            k = job_req.pop(jid)
            lat.append(t_now - arr[k])
            kind_of.append(bool(is_r[k]))
            q_of[k] = r["modularity"]
            n_done += 1
        if a == 0 and q == 0 and i < len(arr):
            time.sleep(max(0.0, arr[i] - (time.perf_counter() - start)))
        if time.perf_counter() - last_log > 5:
            last_log = time.perf_counter()
            log(f"served: t={t_now:.1f}s submitted {i} done {n_done} active {a} queued {q}")
    elapsed = max(time.perf_counter() - start, 1e-9)
    lat = np.array(lat) if lat else np.zeros(1)
    kind_of = np.array(kind_of) if kind_of else np.zeros(1, bool)
    # saturation: a burst of the same request mix, server versus a per-graph loop
    burst = [req_pool(rng_b) for _ in range(args.burst)] if len(req) < args.burst else req[:args.burst]
    st0 = S.stats()
    t0 = time.perf_counter()
    ids = [S.submit(g) for g in burst]
    while S.step()[:2] != (0, 0):
        pass
    srv_s = time.perf_counter() - t0
    st1 = S.stats()
    phases = {k: st1[k] - st0[k] for k in ("seconds_assemble", "seconds_sweeps",
                                           "seconds_aggregate", "uploaded_bytes", "steps")}
    q_srv = {}
    for jid in S.take_finished():
        q_srv[jid] = S.result(jid, best=False)["modularity"]
    q_srv = [q_srv[j] for j in ids]
    t0 = time.perf_counter()
    q_loop = []
    for g in burst:
        r = g.louvain(prm, sweeps_cap=1, levels_cap=1, best=False)
        q_loop.append(r["modularity"])
    loop_s = time.perf_counter() - t0
    dq = max(abs(x - y) for x, y in zip(q_srv, q_loop))
    S.close()
    return {
        "config": {"workload": f"served Louvain: Poisson arrivals at {args.rate:g} req/s for "
                               f"{args.window:g} s (seed 5); 95% {mix_desc}, 5% R-MAT scale 20; "
                               f"HBM budget {args.hbm_budget_gb:g} GB",
                   "requests": int(len(arr)), f"pool_{args.served_mix}": len(pool_p),
                   "pool_rmat20": len(pool_r)},
        "graphs_per_s": len(arr) / elapsed,
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        f"p50_latency_{args.served_mix}_s": float(np.percentile(lat[~kind_of], 50)),
        "p99_latency_rmat20_s": float(np.percentile(lat[kind_of], 99)) if kind_of.any() else None,
        "elapsed_s": elapsed,
        "union_levels": steps,
        "max_active_jobs": max_active,
        "saturation": {"requests": len(burst), "server_graphs_per_s": len(burst) / srv_s,
                       "per_graph_loop_graphs_per_s": len(burst) / loop_s,
                       "speedup": loop_s / srv_s, "max_abs_dq": dq, "server_phases": phases,
                       "note": "the same burst through the server and through lv_louvain per "
                               "graph (upload + run each), both timed end to end"},
    }


def bench_served(args, L):
    b = served_result(args, L)
    out = {
        "metric": METRIC,
        "value": b["p99_latency_s"],
        "unit": "s (p99 latency)",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": b["elapsed_s"] * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": b.pop("config"),
    }
    out.update(b)
    print(json.dumps(out), flush=True)
    return 0


def ingest_result(args, L, ea, eb, prm):
    """End to end from an .edge file: the host reader (graph.cu:17-87 on the
    CPU, then upload) against the GPU ingest, each followed by one Louvain run."""
    import tempfile
    fd, path = tempfile.mkstemp(suffix=".edge", dir=args.tmpdir)
    os.close(fd)
    try:
        t = time.perf_counter()
        L.write_edge(path, ea, eb)
        size = os.path.getsize(path)
        log(f"ingest: wrote {size / 1e9:.2f} GB .edge ({time.perf_counter() - t:.1f}s)")
        t0 = time.perf_counter()
        gh = L.Graph.load_edge(path)
        t1 = time.perf_counter()
        ch = L.Community(gh)
        t2 = time.perf_counter()
        rh = ch.run(prm, sweeps_cap=1, levels_cap=1, best=False, n0=gh.n)
        t3 = time.perf_counter()
        del ch, gh
        log(f"ingest: host read + CSR {t1 - t0:.2f}s, upload {t2 - t1:.2f}s, run {t3 - t2:.3f}s")
        gpu = []
        for _ in range(2):   # the first also pages the file into the OS cache
            t4 = time.perf_counter()
            cd, info = L.Community.load_edge(path)
            t5 = time.perf_counter()
            rd = cd.run(prm, sweeps_cap=1, levels_cap=1, best=False, n0=info["n"])
            t6 = time.perf_counter()
            gpu.append((t5 - t4, t6 - t5, info))
            del cd
        g_ing, g_run, info = gpu[-1]
        log(f"ingest: GPU {g_ing:.3f}s ({info}), run {g_run:.3f}s")
        return {"file_bytes": size, "host_read_csr_s": t1 - t0, "host_upload_s": t2 - t1,
                "gpu_ingest_s": g_ing, "gpu_ingest_phases_s": {
                    "read_upload": info["seconds_read_upload"],
                    "tokenize": info["seconds_tokenize"], "build": info["seconds_build"]},
                "speedup": (t1 - t0 + t2 - t1) / g_ing,
                "end_to_end_host_path_s": t3 - t0, "end_to_end_gpu_path_s": g_ing + g_run,
                "same_modularity": rh["modularity"] == rd["modularity"],
                "note": "file in the OS page cache (second read); host path = GraphHost(path) "
                        "+ upload; GPU path = lv_community_load_edge"}
    finally:
        os.unlink(path)


def bench_batched(args, L):
    """The batched configuration as its own bench line (value = graphs/s)."""
    b = batched_result(args, L)
    out = {
        "metric": METRIC,
        "value": b["graphs_per_s"],
        "unit": "graphs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": b["seconds_per_batch"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": b.pop("config"),
    }
    for k in ("graphs_per_s", "seconds_per_batch", "steps"):
        b.pop(k)
    out.update(b)
    print(json.dumps(out), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--mode", default="hash")
    ap.add_argument("--threshold", type=float, default=0.001)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--probe-bytes", type=int, default=1 << 30,
                    help="size of the plain copy kernel run as the bandwidth reference")
    ap.add_argument("--workload", default="rmat", choices=["rmat", "batched", "served"])
    ap.add_argument("--autotune", type=int, default=1,
                    help="1: tune the best-move tiles during the warm-up run")
    ap.add_argument("--rate", type=float, default=200.0, help="served: arrivals per second")
    ap.add_argument("--window", type=float, default=60.0, help="served: arrival window (s)")
    ap.add_argument("--served-mix", default="ba", choices=["ba", "planted"],
                    help="served: the 95%% requests (ba: the batched configuration's "
                         "distribution; planted: planted-partition graphs of --served-n vertices)")
    ap.add_argument("--served-n", type=int, default=1_000_000,
                    help="served: vertices of the planted-partition requests")
    ap.add_argument("--pool", type=int, default=64, help="served: distinct graphs of the 95%% mix")
    ap.add_argument("--burst", type=int, default=400, help="served: saturation burst size")
    ap.add_argument("--hbm-budget-gb", type=float, default=200.0)
    ap.add_argument("--no-batched", action="store_true",
                    help="rmat: leave out the batched configuration's sub-object")
    ap.add_argument("--no-ingest", action="store_true",
                    help="rmat: leave out the .edge ingest sub-object")
    ap.add_argument("--tmpdir", default=None, help="where the ingest .edge file is written")
    ap.add_argument("--no-served", action="store_true",
                    help="rmat: leave out the served configuration's sub-object")
    ap.add_argument("--served-window", type=float, default=60.0,
                    help="rmat: arrival window of the served sub-object (s)")
    ap.add_argument("--graphs", type=int, default=4096, help="batched: number of graphs")
    ap.add_argument("--no-extras", action="store_true",
                    help="rmat: only the timed runs (no reference-method or per-bucket runs), "
                         "for kernel traces")
    ap.add_argument("--bucket-max-deg", default=None,
                    help="rmat: explicit degree bounds of the 8 best-move buckets, comma "
                         "separated (default: the library's default or tuned bounds)")
    ap.add_argument("--loop-sample", type=int, default=1,
                    help="batched: time the per-graph loop on every k-th graph only")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is described in DESIGN.md and not implemented", flush=True)
        return 2

    import louvain_amd as L
    if L.device_count() < 1:
        raise SystemExit("bench.py: no GPU visible")
    if args.workload == "batched":
        return bench_batched(args, L)
    if args.workload == "served":
        return bench_served(args, L)
    copy_gbs = L.bandwidth_probe(args.probe_bytes, 5) if args.probe_bytes > 0 else None
    log(f"plain copy kernel: {copy_gbs} GB/s")
    t = time.perf_counter()
    ea, eb, n = L.gen_rmat(args.scale, 16, 0.57, 0.19, 0.19, seed=2)
    log(f"rmat{args.scale}: {len(ea)} undirected edges, n={n} ({time.perf_counter() - t:.1f}s)")
    t = time.perf_counter()
    g = L.Graph.from_edges(ea, eb)
    info = g.info()
    log(f"csr: n={info['n']} E={info['E']} ({time.perf_counter() - t:.1f}s)")
    t = time.perf_counter()
    c = L.Community(g)
    log(f"upload + level setup ({time.perf_counter() - t:.2f}s)")
    bounds = {} if not args.bucket_max_deg else {
        "bucket_max_deg": [int(x) for x in args.bucket_max_deg.split(",")]}
    prm = L.params(mode=args.mode, sweep_threshold=args.threshold,
                   level_threshold=args.threshold, profile=2, autotune=args.autotune, **bounds)
    for i in range(args.warmup):
        c.reset()
        r = c.run(prm, best=False, n0=info["n"])
        log(f"warmup {i}: {r['seconds'] * 1e3:.1f} ms, Q={r['modularity']:.6f}")
    times, results, kstats = [], [], []
    for i in range(args.steps):
        t0 = time.perf_counter()
        c.reset()
        r = c.run(prm, best=False, n0=info["n"])
        dt = time.perf_counter() - t0
        times.append(dt)
        results.append(r)
        kstats.append(c.kernel_stats())
        log(f"step {i}: {dt * 1e3:.1f} ms (in-library {r['seconds'] * 1e3:.1f} ms), "
            f"Q={r['modularity']:.6f}, comms={r['n_comm']}, levels={r['levels']}, "
            f"sweeps={r['total_sweeps']}")
    timed_bounds = c.buckets()   # the bucket bounds of the timed runs
    total = sum(times)
    traversed = sum(r["traversed_entries"] for r in results)
    last = results[-1]
    ks = {k: (sum(s[k] for s in kstats) if not isinstance(kstats[0][k], list)
              else [sum(s[k][i] for s in kstats) for i in range(len(kstats[0][k]))])
          for k in kstats[0]}
    # traversed entries per second of the best-move pass (HIP events), per sweep
    # (a level's first sweep decided by the contraction that built the level
    # ran no best-move pass: it is left out of the pass rates)
    all_sw = [s for r in results for s in r["sweeps"] if not s["pre_decided"]]
    sw0 = [s for r in results for s in r["sweeps"][:1]]
    per_sweep = {
        "sweep0_level0_teps": sum(s["e_active"] for s in sw0) / (sum(s["move_ms"] for s in sw0) / 1e3),
        "weighted_mean_teps": sum(s["e_active"] for s in all_sw) / (sum(s["move_ms"] for s in all_sw) / 1e3),
        "sweep0_level0_move_ms": sum(s["move_ms"] for s in sw0) / len(sw0),
        "sweeps": [{"level": s["level"], "sweep": s["sweep"], "kind": s["kind"],
                    "pre_decided": s["pre_decided"],
                    "e_active": s["e_active"], "n_active": s["n_active"],
                    "candidates": s["candidates"], "move_ms": s["move_ms"],
                    "sweep_ms": s["sweep_ms"], "commit_ms": s["commit_ms"],
                    "mover_entries": s["mover_entries"],
                    "move_gbs": s["move_bytes"] / s["move_ms"] / 1e6 if s["move_ms"] > 0 else 0.0}
                   for s in last["sweeps"]],
        "note": "e_active / HIP-event time of the sweep's best-move pass; the weighted mean is "
                "total entries over total pass time of the timed runs.  The first (fast) sweep "
                "of every level above 0 is pre_decided: the contraction that built the level "
                "scored its candidates while summing the rows, so it has no best-move pass "
                "(move_ms 0) and is left out of the weighted mean"}
    # the measured ceiling of the pass's dependent access on this graph: the
    # rate of a kernel that only streams col and gathers arr[col[j]]
    # (4 bytes: tot in the first sweep of a level; 8 bytes: comm | tot in the
    # hash sweeps)
    g4, g8 = c.gather_probe(4), c.gather_probe(8)
    hs = [s for s in all_sw if s["kind"] != 0 and s["level"] == 0]
    hash_teps = (sum(s["e_active"] for s in hs) / (sum(s["move_ms"] for s in hs) / 1e3)
                 if hs else None)
    gather_ceiling = {
        "gathers_per_s_4B": g4, "gathers_per_s_8B": g8,
        "sweep0_level0_frac": per_sweep["sweep0_level0_teps"] / g4,
        "hash_sweeps_level0_teps": hash_teps,
        "hash_sweeps_level0_frac": hash_teps / g8 if hash_teps else None,
        "note": "lv_gather_probe on the level-0 CSR: entries/s of the best-move pass over "
                "gathers/s of the bare stream-col + gather-arr[col] kernel (4 B for the first "
                "sweep's totals, 8 B for the hash sweeps' packed community|total words)"}
    log(f"gather ceiling: {g4 / 1e9:.1f} G/s (4 B), {g8 / 1e9:.1f} G/s (8 B); "
        f"sweep 0 at {gather_ceiling['sweep0_level0_frac']:.2f}, level-0 hash sweeps at "
        f"{gather_ceiling['hash_sweeps_level0_frac']}")
    if args.no_extras:
        reference_method, buckets = None, None
    else:
        reference_method, buckets = extra_runs(args, L, c, info, last["modularity"],
                                               timed_bounds)
    move_s = ks["move_ms"] / 1e3
    pass_gbs = ks["move_bytes"] / move_s / 1e9 if move_s > 0 else 0.0
    # the commit (apply + push-form marking, the incremental change of
    # sigma_in and the next lists) and the contraction, by their per-unit byte
    # figures over their HIP-event time.  The commit's incremental sigma_in
    # replaces the reference's full modularity pass after every sweep
    # (beta*E + 16*n bytes); the second figure counts that pass's bytes too.
    commit_s = ks["commit_ms"] / 1e3
    bq = sum((12 if sw["level"] > 0 else 8) * sw["E"] + 16 * sw["n"]
             for r in results for sw in r["sweeps"])
    commit_pass = {
        "achieved_gbs": ks["commit_bytes"] / commit_s / 1e9 if commit_s > 0 else 0.0,
        "frac": ks["commit_bytes"] / commit_s / 1e9 / HBM_PEAK_GBS if commit_s > 0 else 0.0,
        "achieved_gbs_incl_modularity_pass":
            (ks["commit_bytes"] + bq) / commit_s / 1e9 if commit_s > 0 else 0.0,
        "ms_per_run": ks["commit_ms"] / args.steps,
        "mover_entries_per_run": ks["commit_entries"] / args.steps,
        "note": "bytes 8*n_A + 24*n_moved + 5*mover_entries + n per sweep; time = sweep - "
                "best-move pass (commit kernels and the finalize pass)"}
    contract_s = ks["contract_ms"] / 1e3
    contraction = {
        "achieved_gbs": ks["contract_bytes"] / contract_s / 1e9 if contract_s > 0 else 0.0,
        "frac": ks["contract_bytes"] / contract_s / 1e9 / HBM_PEAK_GBS if contract_s > 0 else 0.0,
        "ms_per_run": ks["contract_ms"] / args.steps,
        "contractions_per_run": ks["contractions"] / args.steps,
        "levels": [{"level": lv["level"], "E": lv["E"], "E_new": lv["E_new"],
                    "contract_ms": lv["contract_ms"],
                    "gbs": lv["contract_bytes"] / lv["contract_ms"] / 1e6
                    if lv["contract_ms"] > 0 else 0.0} for lv in last["level_recs"]],
        "note": "bytes beta*E + 8*E_new + 12*n + 12*n_comm per contraction; time from "
                "renumbering to the contracted graph, host planning included, and the "
                "decisions of the next level's first sweep made on the way (their bytes not "
                "counted)"}
    # dominant best-move kernel: the degree 257-1024 bucket's variant (one
    # kernel launch per sweep) with the most event-timed time over the timed
    # runs.  That bucket runs alone on the device; the others overlap on side
    # streams, so their event times include the neighbours they share it with.
    # variant = ((bucket * 2 + fast) * 2 + weighted) * 3 + tile
    b6 = range(6 * 12, 7 * 12)
    vi = max(b6, key=lambda i: ks["variant_ms"][i])
    v_launch = ks["variant_launches"][vi]
    v_ms = ks["variant_ms"][vi] / max(1, v_launch)
    v_bytes = ks["variant_bytes"][vi] / max(1, v_launch)
    achieved = v_bytes / (v_ms / 1e3) / 1e9 if v_ms > 0 else 0.0
    # hardware traffic of the dominant kernel from the committed counter runs
    traffic = None
    tf = os.path.join(REPO, "profiles", "k_move_table_256_1540_rmat22_traffic.json")
    if os.path.exists(tf) and args.scale == 22:
        t = json.load(open(tf))
        if t.get("kernel") == L.kernel_variant_name(vi):
            traffic = t["traffic_bytes_per_launch"]
    out = {
        "metric": METRIC,
        "value": traversed / total,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": f"plain Louvain, R-MAT scale {args.scale}, edge factor 16, "
                               f"(0.57,0.19,0.19,0.05), seed 2, symmetrised, unit weights",
                   "mode": args.mode, "threshold": args.threshold, "n": info["n"],
                   "E": info["E"], "bucket_max_deg": timed_bounds},
        "s_to_converge": total / args.steps,
        "modularity": last["modularity"],
        "n_communities": last["n_comm"],
        "levels": last["levels"],
        "sweeps": last["total_sweeps"],
        "sweep0_entries": last["sweeps"][0]["e_active"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": L.kernel_variant_name(vi),
                     "launches": v_launch, "avg_launch_ms": v_ms, "avg_bytes": v_bytes,
                     "copy_kernel_gbs": copy_gbs,
                     "gather_ceiling": gather_ceiling,
                     "best_move_pass": {"achieved_gbs": pass_gbs,
                                        "frac": pass_gbs / HBM_PEAK_GBS,
                                        "ms_per_run": ks["move_ms"] / args.steps,
                                        "sweeps_per_run": ks["move_launches"] / args.steps},
                     "commit_pass": commit_pass,
                     "contraction": contraction},
        "per_sweep": per_sweep,
        "buckets": buckets,
        "reference_method": reference_method,
    }
    if not args.no_cpu:
        row, col, _ = g.csr()
        log("cpu baseline (oracle to convergence) ...")
        cb = cpu_baseline(row, col, args.mode, args.threshold)
        out["cpu_baseline"] = cb
        out["speedup_s_to_converge_vs_cpu"] = cb["s_to_converge"] / out["s_to_converge"]
        out["cpu_same_modularity"] = cb["modularity"] == out["modularity"]
        log(f"cpu: {cb['s_to_converge']:.1f} s, Q={cb['modularity']:.9f}")
        del row, col
        log("c1 (CPU-path configuration) ...")
        out["c1"] = c1_result(L)
    del c, g
    if not args.no_ingest:
        log("ingest ...")
        out["ingest"] = ingest_result(args, L, ea, eb, prm)
    del ea, eb
    if not args.no_batched:
        log("batched configuration ...")
        out["batched"] = batched_result(args, L)
    if not args.no_served:
        log("served configuration ...")
        sv = argparse.Namespace(**vars(args))
        sv.window = args.served_window
        sv.served_n = 100_000
        out["served"] = served_result(sv, L)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
