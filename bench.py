"""Benchmark: exact kNN (k=20) + SNN on the count-matrix workload.

Workload (one MI355X): X = 20,000 rows x 100,000 columns CSC, density 0.06,
NB(mu=2, size=0.5) zero-truncated values then per-column log1p(1e4 x /
colsum), float64; self-search sparseKNN with method "binary" (Jaccard) and
"cosine", k = 20, block_size = 8192, each followed by sparseSNN(prune=1/15).
One step = both searches over all 100,000 query columns plus both SNN graphs;
X is resident in device memory before the timed region, and every step
starts from a cold matrix (derived layouts dropped, so densification and
normalisation are timed).

After the timed steps (outside the timed region): full-size checks against
a full-grid search and the oracle's SNN graphs, the plain FP32 path, the
CPU baseline (the oracle on a seeded sample of query columns) and the
serving run on the serving workload's resident 20,000 x 500,000 reference.

Prints one JSON line (rank 0, N = 1).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MI355X_PEAK = {"fp4": 10.1e15, "int8": 5.0e15, "bf16": 2.5e15, "fp32": 157.3e12, "hbm": 8.0e12}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


SERVE_RATE = 200.0        # requests per second (Poisson)
SERVE_MAX_COLS = 50_000   # request sizes log-uniform on [1, 50000] columns


def serve_arrivals(window_s, seed, pool_cols, max_cols=SERVE_MAX_COLS):
    """The request stream: Poisson arrivals, log-uniform sizes on [1, max_cols],
    method uniform over {cosine, binary}, k uniform over {10, 20, 30}, SNN on
    half."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / SERVE_RATE)
        if t >= window_s:
            return out
        n = int(np.exp(rng.uniform(0.0, np.log(max_cols + 1))))
        n = max(1, min(n, max_cols, pool_cols))
        out.append(dict(t=t, n=n, method=("cosine", "binary")[int(rng.integers(2))],
                        k=int((10, 20, 30)[int(rng.integers(3))]), snn=bool(rng.random() < 0.5),
                        start=int(rng.integers(0, pool_cols - n + 1))))


def serve_run(sd, dY, pool, arrivals, precision, slots=4, chunk_cols=4096, max_pending=16,
              graphs=False):
    """Replays the request stream against a server over dY; each request is a
    contiguous slice of the query pool.  Returns latency and throughput."""
    srv = sd.Server(dY, slots=slots, chunk_cols=chunk_cols, max_pending=max_pending,
                    graphs=graphs)
    try:
        # build Y's derived layouts for both methods before the clock starts
        for m in ("binary", "cosine"):
            srv.result(srv.submit_csc(pool.shape[0], pool.indptr[:2] - pool.indptr[0],
                                      pool.indices[: pool.indptr[1]], pool.data[: pool.indptr[1]],
                                      k=10, method=m, precision=precision))
        live, refused = [], 0
        t0 = time.perf_counter()
        for a in arrivals:
            now = time.perf_counter() - t0
            if now < a["t"]:
                time.sleep(a["t"] - now)
            c0, c1 = a["start"], a["start"] + a["n"]
            p0, p1 = int(pool.indptr[c0]), int(pool.indptr[c1])
            try:
                tk = srv.submit_csc(pool.shape[0], pool.indptr[c0:c1 + 1] - p0,
                                    pool.indices[p0:p1], pool.data[p0:p1], k=a["k"],
                                    method=a["method"], snn=a["snn"], prune=1 / 15,
                                    precision=precision, borrow=True)
                live.append((tk, a))
            except sd.SparseDistError as e:
                if e.code != sd._lib.SD_EBUDGET:
                    raise
                refused += 1
        infos = []
        for tk, a in live:
            info = srv.wait(tk)
            if info["status"] != 0:
                raise RuntimeError(f"request failed: {srv.result(tk)}")
            infos.append((info, a))
            srv.release(tk)
        t_end = time.perf_counter() - t0
        gst = srv.stats()
    finally:
        srv.close()
    lat = np.array([i["total_ms"] for i, _ in infos])
    cols = sum(a["n"] for _, a in infos)
    small = np.array([i["total_ms"] for i, a in infos if a["n"] <= 1000])
    return dict(precision=precision, offered=len(arrivals), served=len(infos), refused=refused,
                p50_ms=float(np.percentile(lat, 50)) if lat.size else None,
                p99_ms=float(np.percentile(lat, 99)) if lat.size else None,
                p99_small_ms=float(np.percentile(small, 99)) if small.size else None,
                served_cols_per_s=cols / t_end, wall_s=t_end,
                snn_requests=sum(1 for _, a in infos if a["snn"]), graphs=graphs,
                graph_items=gst["graph_items"], graph_captures=gst["graph_captures"],
                graph_fallbacks=gst["graph_fallbacks"], packed_items=gst["packed_items"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ncol", type=int, default=100_000)
    ap.add_argument("--nrow", type=int, default=20_000)
    ap.add_argument("--density", type=float, default=0.06)
    ap.add_argument("--k", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=48,
                    help="query columns of the CPU baseline's sample (a seeded draw over all "
                         "columns)")
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--serve-window", type=float, default=4.0,
                    help="seconds of request arrivals in the serving run (0: skip it)")
    ap.add_argument("--serve-window-plain", type=float, default=1.0,
                    help="seconds of arrivals for the plain FP32 serving run (its requests "
                         "take far longer to drain)")
    ap.add_argument("--serve", action="store_true",
                    help="serving benchmark only, on the 20000 x 500000 resident reference")
    ap.add_argument("--serve-ref", default="20000x500000",
                    help="with --serve: rows x columns of the resident reference")
    ap.add_argument("--serve-small", action="store_true",
                    help="with --serve: a stream of small requests (1-256 columns) served with "
                         "and without the captured graphs of small work items")
    ap.add_argument("--sweep", action="store_true",
                    help="path-chooser sweep: densified vs merge walk vs auto over density "
                         "and batch size (20000-row references of 20000 columns)")
    ap.add_argument("--exact", action="store_true",
                    help="FP64 merge-walk path (the one the cost model picks for sparse "
                         "references): 2048 queries against a 20000 x 20000 reference at "
                         "density 0.002, binary and cosine")
    ap.add_argument("--large", action="store_true",
                    help="large sparse self-search only (30000 x 1000000, density 0.03)")
    ap.add_argument("--cross", action="store_true",
                    help="dense-embedding cross kNN only (E 256 x 200000, Y 256 x 1000000)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one plain 1 GiB device copy (calibrates rocprofv3's "
                         "FETCH_SIZE / WRITE_SIZE counters against known bytes)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the query-sharded multi-GPU design is described but not run).", file=sys.stderr)
        sys.exit(2)

    import sparsedist_amd as sd
    from sparsedist_amd import synth

    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    if args.serve:
        return serve_main(args, sd, synth, workers)
    if args.cross:
        return cross_main(args, sd, synth)
    if args.large:
        return large_main(args, sd, synth, workers)
    if args.sweep:
        return sweep_main(args, sd, synth, workers)
    if args.exact:
        return exact_main(args, sd, synth, workers)
    t0 = time.time()
    X = synth.nb_log1p(args.nrow, args.ncol, args.density, seed=20)
    log(f"generated X {X.shape} nnz={X.nnz} in {time.time() - t0:.1f}s")
    pool = Y5 = None
    if args.serve_window > 0:
        # the serving workload's resident reference and query pool, generated
        # before the GPU is touched
        t0 = time.time()
        Y5 = synth.nb_log1p_parallel(20_000, 500_000, 0.05, 50, workers)
        pool = synth.nb_log1p_parallel(20_000, SERVE_MAX_COLS, 0.05, 51, workers)
        log(f"generated the serving reference {Y5.shape} nnz={Y5.nnz} and the query pool "
            f"in {time.time() - t0:.1f}s")
    ctx = sd.default_context()
    dX = sd.DeviceMatrix(X, ctx)
    nq = args.ncol
    k = args.k

    def step(precision="auto", timing=False):
        res = {}
        for method in ("binary", "cosine"):
            dX.drop_cache()
            ctx.stats(reset=True)
            ctx.set_timing(timing)
            t = time.perf_counter()
            nn = sd.sparseKNN(dX, k=k, method=method, block_size=8192, precision=precision,
                              verbose=False, context=ctx)
            t_knn = time.perf_counter() - t
            st = ctx.stats(reset=True)
            t = time.perf_counter()
            g = sd.sparseSNN(nn, prune=1 / 15, context=ctx)
            t_snn = time.perf_counter() - t
            st2 = ctx.stats(reset=True)
            ctx.set_timing(False)
            res[method] = dict(t_knn=t_knn, t_snn=t_snn, stats=st, snn_stats=st2, nnz=g.nnz,
                               idx=nn.idx, dist=nn.dist, graph=g)
        return res

    # setup: the threshold-sample tuner times each of its four candidates
    # once per class of large searches, one per search (a search that had to
    # grow a device buffer is not timed), and keeps the fastest; a deployment
    # tunes once, so the workload runs until both classes are tuned (at most
    # 10 times) before the warm-up; results are the same whichever candidate
    # runs
    sd.tune_samples(reset=True)
    for t_ in range(10):
        step()
        tab = sd.tune_samples()
        if len(tab) >= 2 and all(min(e["runs"]) >= 1 for e in tab):
            break
    log(f"sample tuning took {t_ + 1} runs of the workload")
    for w in range(args.warmup):
        step()
        log(f"warmup {w + 1}/{args.warmup} done")
    times, per = [], []
    for s in range(args.steps):
        t = time.perf_counter()
        r = step(timing=True)
        times.append(time.perf_counter() - t)
        per.append(r)
        log(f"step {s + 1}/{args.steps}: {times[-1] * 1e3:.1f} ms "
            f"(binary {r['binary']['t_knn'] * 1e3:.1f} ms, cosine {r['cosine']['t_knn'] * 1e3:.1f} ms, "
            f"snn {r['binary']['t_snn'] * 1e3 + r['cosine']['t_snn'] * 1e3:.1f} ms, "
            f"fallback {r['cosine']['stats']['fallback_queries']})")
    ms = float(np.median(times)) * 1e3
    value = 2 * nq / (ms / 1e3)
    # the tuner's table after the timed steps: per class, the candidates'
    # measured times, the fixed rule's choice and the candidate now running
    tuned = []
    for e in sd.tune_samples():
        timed = [c for c in range(4) if e["runs"][c] > 0]
        cand = [f"{e['candidates'][c][0]}/{e['candidates'][c][1]}" for c in range(4)]
        tuned.append({"method": {0: "binary", 1: "cosine"}.get(e["method"], e["method"]),
                      "symmetric": e["symmetric"],
                      "ns_per_mpair": {cand[c]: round(e["ns_per_mpair"][c], 1) for c in timed},
                      "rule": cand[e["rule"]], "running": cand[e["last_run"]]})
    log(f"sample tuning (stride/rank): {tuned}")

    def med(method, key):
        return float(np.median([p[method][key] for p in per]))

    def kern(method):
        # average launch of the filtered full pass (the kernel that visits every
        # (query, reference) pair); list-mode launches are timed separately
        def one(st):
            if st["filter_launches"] > 0:
                return st["filter_ms"] / st["filter_launches"]
            return st["slab_ms"] / max(st["slab_launches"], 1)  # small search: list mode only
        return float(np.median([one(p[method]["stats"]) for p in per]))

    def pairs(method):
        # (query, reference) pairs one filtered launch scores: a self-search
        # over all columns at once scores each unordered pair once
        st = per[0][method]["stats"]
        if st["filter_launches"] > 0:
            return st["filter_pairs"] / st["filter_launches"]
        return float(nq) * nq

    # 2 r per scored pair: dense-equivalent operations of one filtered launch
    flop_bin = 2.0 * args.nrow * pairs("binary")
    flop = 2.0 * args.nrow * pairs("cosine")
    bin_ms = kern("binary")
    cos_ms = kern("cosine")
    roof_bin = dict(bound="mfma", achieved=flop_bin / (bin_ms / 1e3) / 1e12, peak=10100.0,
                    unit="TFLOP/s", frac=(flop_bin / (bin_ms / 1e3)) / MI355X_PEAK["fp4"],
                    traffic=None,
                    kernel="k_filter_wide<OpF4, KeyJ> (FP4-pattern Jaccard filtered pass)",
                    avg_launch_ms=bin_ms)
    roof_cos = dict(bound="mfma", achieved=flop / (cos_ms / 1e3) / 1e12, peak=2500.0,
                    unit="TFLOP/s", frac=(flop / (cos_ms / 1e3)) / MI355X_PEAK["bf16"],
                    traffic=None, kernel="k_filter_wide<OpBf16, float> (BF16 cosine filtered pass)",
                    avg_launch_ms=cos_ms)
    # HBM bytes per launch (reads + writes) from the committed rocprofv3 PMC
    # runs of this workload (FETCH_SIZE / WRITE_SIZE, copy-calibrated)
    tpath = os.path.join(ROOT, "profiles", "traffic_count_matrix_20000x100000.json")
    if (args.nrow, args.ncol, args.density, k) == (20_000, 100_000, 0.06, 20) and \
            os.path.exists(tpath):
        tr = json.load(open(tpath))
        # the timed searches run the symmetric pass (SYM = true)
        for roof, name in ((roof_bin, "void k_filter_wide<OpF4, KeyJ, 128, 0, false, true>"),
                           (roof_cos, "void k_filter_wide<OpBf16, float, 128, 0, false, true>")):
            if name in tr:
                t = tr[name]
                roof["traffic"] = t["fetch_bytes"] + t["write_bytes"]
                roof["traffic_unit"] = "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE)"
    dominant = roof_bin if bin_ms >= cos_ms else roof_cos
    edges = 2 * nq * k
    snn_s = med("binary", "t_snn") + med("cosine", "t_snn")

    # full-size check of the symmetric pass (outside the timed region): a
    # cross search against a copy of X runs the full grid of tiles; with the
    # self rules applied to its k + 1 lists it must equal every timed row
    full_grid = {}
    dXc = sd.DeviceMatrix(X, ctx)
    for method in ("binary", "cosine"):
        c = sd.sparseKNN(dX, dXc, k=k + 1, method=method, block_size=8192, verbose=False,
                         context=ctx)
        n = c.idx.shape[0]
        hit = c.idx == np.arange(1, n + 1, dtype=np.int32)[:, None]
        pos = np.where(hit.any(axis=1), hit.argmax(axis=1), k)  # the own column, else the last
        keep = np.ones(c.idx.shape, bool)
        keep[np.arange(n), pos] = False
        wi, wd = c.idx[keep].reshape(n, k), c.dist[keep].reshape(n, k)
        gi, gd = per[-1][method]["idx"], per[-1][method]["dist"]
        full_grid[method] = bool(np.array_equal(gi, wi) and np.array_equal(
            np.ascontiguousarray(gd).view(np.int64), np.ascontiguousarray(wd).view(np.int64)))
        del c
    del dXc
    log(f"symmetric pass vs full grid, all {nq} rows: {full_grid}")
    plain = None
    if not args.no_plain:
        r = step(precision="fp32", timing=True)
        same = all(np.array_equal(r[m]["idx"], per[-1][m]["idx"]) for m in ("binary", "cosine"))
        plain = dict(binary_ms=r["binary"]["t_knn"] * 1e3, cosine_ms=r["cosine"]["t_knn"] * 1e3,
                     cols_per_s=2 * nq / (r["binary"]["t_knn"] + r["cosine"]["t_knn"]),
                     idx_identical=bool(same))
        log(f"plain fp32 path: binary {plain['binary_ms']:.1f} ms cosine {plain['cosine_ms']:.1f} ms")

    # full-size SNN check (outside the timed region): both graphs of the last
    # timed step equal the oracle's CSC (structure and weight bits)
    from oracle import oracle as O
    O.build()
    nt = min(16, os.cpu_count() or 1)
    snn_parity = {}
    for method in ("binary", "cosine"):
        g = per[-1][method]["graph"]
        p_, ri_, x_ = O.snn_graph(per[-1][method]["idx"], include_self=True, prune=1 / 15,
                                  symmetrise=True, nthreads=nt)
        snn_parity[method] = bool(np.array_equal(g.indptr, p_) and np.array_equal(g.indices, ri_)
                                  and np.array_equal(g.data.view(np.int64), x_.view(np.int64)))
    log(f"SNN graphs of all {nq} columns equal the oracle's: {snn_parity}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args, O, X, per[-1], nt)
        log(f"cpu baseline: {cpu['value']:.2f} cols/s on {nt} threads (median of "
            f"{args.cpu_reps}), {cpu['one_thread']['value']:.2f} cols/s on 1 thread; GPU rows of "
            f"the sample equal the oracle's: {cpu['sample_parity']}")

    serving = None
    if pool is not None:
        dX.drop_cache()
        dY5 = sd.DeviceMatrix(Y5, ctx)
        arr = serve_arrivals(args.serve_window, 22, pool.shape[1])
        runs = [serve_run(sd, dY5, pool, arr, "auto")]
        if not args.no_plain:
            arr_p = serve_arrivals(args.serve_window_plain, 22, pool.shape[1])
            runs.append(serve_run(sd, dY5, pool, arr_p, "fp32"))
        dY5.free()
        serving = dict(reference=f"Y {Y5.shape[0]} x {Y5.shape[1]} d=0.05 NB/log1p "
                                 f"(nnz {Y5.nnz}) resident",
                       stream=f"Poisson {SERVE_RATE:.0f} req/s, sizes log-uniform on [1, 50000] "
                              "columns, cosine/binary, k in {10, 20, 30}, SNN on half; 4 slots, "
                              "4096-column work items, 16 waiting requests at most; "
                              f"{args.serve_window:.0f} s of arrivals for the fast path, "
                              f"{args.serve_window_plain:.0f} s (the same stream's first "
                              "arrivals) for plain FP32",
                       runs=runs)
        for r in runs:
            log(f"serving {r['precision']}: served {r['served']}/{r['offered']} p50 "
                f"{r['p50_ms']:.0f} ms p99 {r['p99_ms']:.0f} ms, {r['served_cols_per_s']:.0f} cols/s")

    out = {
        "metric": "exact-kNN query cols/s (k=20) + SNN edges/s",
        "value": value,
        "unit": "query cols/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp4 0/1 patterns (Jaccard, exact counts) + bf16 (cosine) slab, fp64 re-rank",
        "data": "synthetic (seeded NB/log1p count matrix)",
        "config": {"workload": f"count-matrix kNN {args.nrow}x{nq} d={args.density} "
                               f"binary+cosine k={k} + SNN prune=1/15",
                   "nrow": args.nrow, "ncol": nq, "nnz": int(X.nnz), "k": k,
                   "block_size": 8192, "parallelism": "single GPU"},
        "binary_cols_per_s": nq / med("binary", "t_knn"),
        "cosine_cols_per_s": nq / med("cosine", "t_knn"),
        "snn_edges_per_s": edges / snn_s,
        "snn_nnz": {m: per[-1][m]["nnz"] for m in ("binary", "cosine")},
        "cosine_fallback_queries": int(per[-1]["cosine"]["stats"]["fallback_queries"]),
        "plain_fp32": plain,
        "vs_plain_fp32": (value / plain["cols_per_s"]) if plain else None,
        "roofline": dominant,
        "roofline_other": roof_cos if dominant is roof_bin else roof_bin,
        "cpu_baseline": cpu,
        "symmetric_vs_full_grid": full_grid,
        "sample_tuning": tuned,
        "snn_vs_oracle": snn_parity,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


def self_rows(oi, od, cols, k):
    """Self-search rows (k neighbours) of query columns cols from an oracle
    search of X[:, cols] against X with k + 1 neighbours: the query's own
    column leaves its list."""
    wi = np.empty((len(cols), k), np.int32)
    wd = np.empty((len(cols), k))
    for r, c in enumerate(cols):
        keep = oi[r] != c + 1
        if keep.all():
            keep[-1] = False
        wi[r] = oi[r][keep][:k]
        wd[r] = od[r][keep][:k]
    return wi, wd


def cpu_baseline(args, O, X, last, nt):
    """The oracle (C + OpenMP restatement of the reference's slab-then-top-k
    loop) on a seeded sample of query columns spread over all of X, against
    every column of X, both methods: median and MAD over args.cpu_reps
    repetitions on nt threads, one repetition on one thread (a smaller
    sample), and the slab-kernel / top-k / end-to-end split.  The timed
    step's GPU rows of the sample must equal the oracle's."""
    nq, k, ns = X.shape[1], args.k, args.cpu_sample
    cols = np.sort(np.random.default_rng(23).choice(nq, ns, replace=False))
    Xs = X[:, cols]
    reps, phases, parity = [], [], {}
    for rep in range(args.cpu_reps):
        t = time.perf_counter()
        slab = topk = 0.0
        for method in ("binary", "cosine"):
            oi, od, _ = O.knn(Xs, X, k=k + 1, method=method, block_size=8192, nthreads=nt)
            a, b = O.last_phases()
            slab, topk = slab + a, topk + b
            if rep == 0:
                wi, wd = self_rows(oi, od, cols, k)
                gi, gd = last[method]["idx"][cols], last[method]["dist"][cols]
                parity[method] = bool(np.array_equal(gi, wi) and np.array_equal(
                    np.ascontiguousarray(gd).view(np.int64), wd.view(np.int64)))
        reps.append(time.perf_counter() - t)
        phases.append((slab, topk))
        log(f"cpu baseline rep {rep + 1}/{args.cpu_reps}: {reps[-1]:.1f} s")
    med = float(np.median(reps))
    mad = float(np.median(np.abs(np.array(reps) - med)))
    n1 = max(1, ns // 6)
    t = time.perf_counter()
    for method in ("binary", "cosine"):
        O.knn(Xs[:, :n1], X, k=k + 1, method=method, block_size=8192, nthreads=1)
    t1 = time.perf_counter() - t
    slab_m = float(np.median([p[0] for p in phases]))
    topk_m = float(np.median([p[1] for p in phases]))
    return dict(value=2 * ns / med, unit="query cols/s", cores=nt, kind="port",
                sample=f"{ns} query columns (seeded draw over all {nq}) x all {nq} reference "
                       f"columns, binary and cosine k={k} (+1 for the query's own column), "
                       f"oracle/sd_oracle.c with OpenMP; median of {args.cpu_reps} "
                       f"repetitions ({med:.1f} s, MAD {mad:.2f} s)",
                seconds_median=med, seconds_mad=mad, seconds_reps=reps,
                phases={"slab_kernel_s": slab_m, "topk_s": topk_m,
                        "end_to_end_s": med},
                one_thread=dict(value=2 * n1 / t1, unit="query cols/s", cores=1,
                                sample=f"{n1} of the sampled columns, both methods ({t1:.1f} s)"),
                sample_parity=parity)


def serve_main(args, sd, synth, workers):
    """Serving on the 20000 x 500000 reference (density 0.05, NB/log1p)."""
    t0 = time.time()
    nr, nc = (int(v) for v in args.serve_ref.lower().split("x"))
    Y = synth.nb_log1p_parallel(nr, nc, 0.05, 50, workers)
    pool = synth.nb_log1p_parallel(nr, SERVE_MAX_COLS, 0.05, 51, workers)
    log(f"generated Y {Y.shape} nnz={Y.nnz} and the query pool in {time.time() - t0:.1f}s")
    ctx = sd.default_context()
    dY = sd.DeviceMatrix(Y, ctx)
    del Y
    window = args.serve_window if args.serve_window > 0 else 10.0
    if args.serve_small:
        # small requests only: what the captured graphs of small items change
        arr = serve_arrivals(window, 53, pool.shape[1], max_cols=256)
        runs = [serve_run(sd, dY, pool, arr, "auto", graphs=True),
                serve_run(sd, dY, pool, arr, "auto", graphs=False)]
        for r in runs:
            log(f"small requests, graphs {r['graphs']}: served {r['served']}/{r['offered']} p50 "
                f"{r['p50_ms']:.1f} ms p99 {r['p99_ms']:.1f} ms, graph items {r['graph_items']}, "
                f"fallbacks {r['graph_fallbacks']}")
        print(json.dumps({"metric": "serving p50/p99 latency (ms), small requests",
                          "config": {"workload": f"Y {nr} x {nc} d=0.05 resident; Poisson "
                                     f"{SERVE_RATE:.0f} req/s for {window:.0f} s; sizes "
                                     "log-uniform [1, 256]; cosine/binary; k in {10, 20, 30}; "
                                     "SNN on half", "slots": 4, "chunk_cols": 4096},
                          "runs": runs}), flush=True)
        return
    arr = serve_arrivals(window, 52, pool.shape[1])
    runs = [serve_run(sd, dY, pool, arr, "auto")]
    if not args.no_plain:
        runs.append(serve_run(sd, dY, pool, arr, "fp32"))
    for r in runs:
        log(f"serving {r['precision']}: served {r['served']}/{r['offered']} p50 "
            f"{r['p50_ms']:.0f} ms p99 {r['p99_ms']:.0f} ms, {r['served_cols_per_s']:.0f} cols/s")
    a = runs[0]
    out = {
        "metric": "serving p50/p99 latency (ms) and served query cols/s",
        "value": a["served_cols_per_s"], "unit": "served query cols/s", "n_gpus": 1,
        "higher_is_better": True, "p50_ms": a["p50_ms"], "p99_ms": a["p99_ms"],
        "dtype": "fp4 0/1 patterns (Jaccard, exact counts) + bf16 (cosine) slab, fp64 re-rank", "data": "synthetic (seeded NB/log1p)",
        "config": {"workload": "serving: Y 20000 x 500000 d=0.05 NB/log1p resident; Poisson "
                               f"{SERVE_RATE:.0f} req/s for {window:.0f} s; sizes log-uniform "
                               "[1, 50000]; cosine/binary; k in {10, 20, 30}; SNN on half",
                   "slots": 4, "chunk_cols": 4096, "max_pending": 16},
        "runs": runs,
        "vs_plain_fp32": (a["served_cols_per_s"] / runs[1]["served_cols_per_s"])
        if len(runs) > 1 else None,
    }
    print(json.dumps(out), flush=True)


def cross_main(args, sd, synth):
    """sparseKNNCross(E, Y, k=10, method="cosine", dist=False) on dense embeddings:
    E 256 x 200000 and Y 256 x 1000000, N(0,1) FP32 columns unit-normalised."""
    t0 = time.time()
    E = synth.unit_normal_dense(256, 200_000, 60)
    Y = synth.unit_normal_dense(256, 1_000_000, 61)
    log(f"generated E {E.shape} and Y {Y.shape} in {time.time() - t0:.1f}s")
    ctx = sd.default_context()
    dE = sd.DeviceMatrix(E.astype(np.float64), ctx)
    dY = sd.DeviceMatrix(Y.astype(np.float64), ctx)
    k = 10

    def run(precision):
        dE.drop_cache()
        dY.drop_cache()
        ctx.stats(reset=True)
        ctx.set_timing(True)
        t = time.perf_counter()
        cr = sd.sparseKNNCross(dE, dY, k=k, method="cosine", dist=False, precision=precision,
                               verbose=False, context=ctx)
        dt = time.perf_counter() - t
        st = ctx.stats(reset=True)
        ctx.set_timing(False)
        return cr, dt, st

    run("auto")  # warm-up (code objects, allocations)
    cr, dt, st = run("auto")
    nq = 200_000 + 1_000_000
    log(f"cross auto: {dt * 1e3:.0f} ms for both directions ({nq / dt:.0f} query cols/s); "
        f"certified {st['certified_queries']}, escalated {st['escalated_queries']}, "
        f"fallback {st['fallback_queries']}")
    # full-size check of the one-pass structure: two independent searches
    # (each with its own filtered pass) give the same results on every row
    fw = sd.sparseKNN(dE, dY, k=k, method="cosine", dist=False, verbose=False, context=ctx)
    rv = sd.sparseKNN(dY, dE, k=k, method="cosine", dist=False, verbose=False, context=ctx)
    one_pass = dict(
        forward_identical=bool(np.array_equal(fw.idx, cr["forward"].idx) and np.array_equal(
            fw.dist.view(np.int64), cr["forward"].dist.view(np.int64))),
        reverse_identical=bool(np.array_equal(rv.idx, cr["reverse"].idx) and np.array_equal(
            rv.dist.view(np.int64), cr["reverse"].dist.view(np.int64))))
    del fw, rv
    log(f"one filtered pass vs two independent searches: {one_pass}")
    plain = None
    if not args.no_plain:
        cp, dtp, _ = run("fp32")
        same = (np.array_equal(cp["forward"].idx, cr["forward"].idx)
                and np.array_equal(cp["reverse"].idx, cr["reverse"].idx))
        plain = dict(seconds=dtp, cols_per_s=nq / dtp, idx_identical=bool(same))
        log(f"cross plain fp32: {dtp * 1e3:.0f} ms")
    # 2 r per scored pair: one filtered launch scores the 2e11 (x, y) pairs
    # once for both directions
    fms = st["filter_ms"] / max(st["filter_launches"], 1)
    flop = 2.0 * 256 * st["filter_pairs"] / max(st["filter_launches"], 1)
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        import scipy.sparse as sp
        O.build()
        nt = min(16, os.cpu_count() or 1)
        ns = 16
        cols = np.sort(np.random.default_rng(62).choice(E.shape[1], ns, replace=False))
        t = time.perf_counter()
        oi, od, _ = O.knn(sp.csc_matrix(E[:, cols].astype(np.float64)),
                          sp.csc_matrix(Y.astype(np.float64)), k=k, method="cosine", dist=False,
                          nthreads=nt)
        tc = time.perf_counter() - t
        # the timed search's forward rows of the sample equal the oracle's
        parity = bool(np.array_equal(cr["forward"].idx[cols], oi) and np.array_equal(
            np.ascontiguousarray(cr["forward"].dist[cols]).view(np.int64), od.view(np.int64)))
        cpu = dict(value=ns / tc, unit="query cols/s", cores=nt, kind="port",
                   sample=f"{ns} E columns (seeded draw) x all 1,000,000 Y columns (forward "
                          f"direction), oracle/sd_oracle.c with OpenMP ({tc:.1f} s)",
                   sample_parity=parity)
        log(f"cpu baseline {ns / tc:.2f} cols/s; GPU forward rows of the sample equal the "
            f"oracle's: {parity}")
    out = {
        "metric": "exact cross-kNN query cols/s (k=10, both directions)",
        "value": nq / dt, "unit": "query cols/s", "n_gpus": 1, "higher_is_better": True,
        "ms": dt * 1e3, "dtype": "bf16 slab, fp64 re-rank", "data": "synthetic N(0,1) unit columns",
        "config": {"workload": "dense-embedding cross kNN: E 256 x 200000, Y 256 x 1000000, "
                               "cosine, dist=FALSE, k=10, forward + reverse"},
        "roofline": dict(bound="mfma", achieved=flop / (fms / 1e3) / 1e12, peak=2500.0,
                         unit="TFLOP/s", frac=flop / (fms / 1e3) / MI355X_PEAK["bf16"],
                         traffic=None, kernel="k_filter_wide<OpBf16, float> (both directions)",
                         avg_launch_ms=fms, launches=st["filter_launches"]),
        "certified": st["certified_queries"], "escalated": st["escalated_queries"],
        "fallback": st["fallback_queries"], "one_pass_vs_two_searches": one_pass,
        "plain_fp32": plain, "vs_plain_fp32": (dtp / dt) if plain else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def large_main(args, sd, synth, workers):
    """Self-search kNN (k=20) on X = 30000 x 1000000 CSC, density 0.03, NB/log1p:
    binary, cosine and euclidean, each over every column in one call (the
    filtered pass scores each unordered pair once)."""
    t0 = time.time()
    X = synth.nb_log1p_parallel(30_000, 1_000_000, 0.03, 70, workers)
    log(f"generated X {X.shape} nnz={X.nnz} in {time.time() - t0:.1f}s")
    ctx = sd.default_context()
    t0 = time.time()
    dX = sd.DeviceMatrix(X, ctx)
    log(f"uploaded in {time.time() - t0:.1f}s")
    k, nq = 20, X.shape[1]
    # full-size parity on a seeded sample of query columns: the exact FP64
    # merge walk over all 10^6 reference columns must give the same idx and
    # bit-identical dist as the fused search
    sample = np.sort(np.random.default_rng(77).choice(nq, 512, replace=False))
    Xs = X[:, sample]
    res = {}
    for method in ("binary", "cosine", "euclidean"):
        dX.drop_cache()
        ctx.stats(reset=True)
        ctx.set_timing(True)
        t = time.perf_counter()
        nn = sd.sparseKNN(dX, k=k, method=method, block_size=8192, verbose=False, context=ctx)
        got_i, got_d = nn.idx[sample].copy(), nn.dist[sample].copy()
        del nn
        dt = time.perf_counter() - t
        st = ctx.stats(reset=True)
        ctx.set_timing(False)
        fms = st["filter_ms"] / max(st["filter_launches"], 1)
        peak = MI355X_PEAK["fp4" if method == "binary" else "bf16"]
        # dense-equivalent operations of one filtered pass (the split BF16
        # layout runs three products per row)
        rows = X.shape[0] * max(st["slab_products"], 1)
        pairs = st["filter_pairs"] / max(st["filter_launches"], 1)  # scored per launch
        flop = 2.0 * rows * pairs
        res[method] = dict(seconds=dt, cols_per_s=nq / dt, filter_ms=fms,
                           slab_products_per_row=st["slab_products"],
                           filter_frac_of_peak=flop / (fms / 1e3) / peak,
                           certified=st["certified_queries"], escalated=st["escalated_queries"],
                           fallback=st["fallback_queries"], overflow=st["overflow_queries"])
        log(f"{method}: {dt:.1f} s ({nq / dt:.0f} cols/s), filtered pass {fms:.0f} ms "
            f"({res[method]['filter_frac_of_peak'] * 100:.0f}% of peak), escalated "
            f"{st['escalated_queries']}, fallback {st['fallback_queries']}")
        t = time.perf_counter()
        # self-search without self: drop each column's own entry from k + 1
        ex = sd.sparseKNN(Xs, dX, k=k + 1, method=method, precision="exact", verbose=False,
                          context=ctx)
        want_i = np.empty((sample.size, k), np.int32)
        want_d = np.empty((sample.size, k))
        for r, c in enumerate(sample):
            keep = ex.idx[r] != c + 1
            if keep.all():
                keep[-1] = False
            want_i[r] = ex.idx[r][keep][:k]
            want_d[r] = ex.dist[r][keep][:k]
        same_i = bool(np.array_equal(got_i, want_i))
        same_d = bool(np.array_equal(got_d.view(np.int64), want_d.view(np.int64)))
        res[method]["sample_parity"] = dict(queries=int(sample.size), idx_identical=same_i,
                                            dist_bit_identical=same_d,
                                            exact_seconds=time.perf_counter() - t)
        log(f"{method}: exact FP64 check of {sample.size} sampled queries against all "
            f"{nq} references: idx identical {same_i}, dist identical {same_d} "
            f"({time.perf_counter() - t:.1f} s)")
    secs = sum(res[m]["seconds"] for m in res)
    out = {
        "metric": "exact-kNN query cols/s (k=20), large sparse self-search",
        "value": 3 * nq / secs,
        "unit": "query cols/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "fp4 0/1 patterns (binary), bf16 (cosine, euclidean on non-negative data) slab, "
                 "fp64 re-rank", "data": "synthetic (seeded NB/log1p)",
        "config": {"workload": "X 30000 x 1000000 CSC d=0.03 NB/log1p, k=20: binary, cosine "
                               "and euclidean over all 1000000 query columns"},
        "methods": res,
    }
    print(json.dumps(out), flush=True)


def exact_main(args, sd, synth, workers):
    """The FP64 merge walk (k_exact_slab + k_topk_dense) on the inputs where the
    cost model prefers it: every pair walks the reference column against the
    query staged in LDS.  Algorithmic bytes per query: the reference CSC once
    (12 B per stored value), the slab column written and read back by the
    top-k (16 B per reference)."""
    Y = synth.nb_log1p_parallel(20_000, 20_000, 0.002, 81, workers)
    X = synth.nb_log1p_parallel(20_000, 2048, 0.002, 91, workers)
    ctx = sd.default_context()
    dY = sd.DeviceMatrix(Y, ctx)
    dX = sd.DeviceMatrix(X, ctx)
    res = {}
    for method in ("binary", "cosine"):
        sd.sparseKNN(dX, dY, k=10, method=method, precision="exact", verbose=False, context=ctx)
        best = 1e30
        for _ in range(3):
            t = time.perf_counter()
            sd.sparseKNN(dX, dY, k=10, method=method, precision="exact", verbose=False,
                         context=ctx)
            best = min(best, time.perf_counter() - t)
        alg = X.shape[1] * (12.0 * Y.nnz + 16.0 * Y.shape[1])
        res[method] = dict(ms=best * 1e3, cols_per_s=X.shape[1] / best,
                           algorithmic_bytes=alg, algorithmic_gb_per_s=alg / best / 1e9)
        log(f"{method}: {best * 1e3:.2f} ms, {alg / best / 1e9:.0f} GB/s algorithmic")
    print(json.dumps({"metric": "FP64 merge-walk kNN (k=10) query cols/s", "unit": "query cols/s",
                      "value": 2 * X.shape[1] / sum(r["ms"] / 1e3 for r in res.values()),
                      "config": {"workload": "2048 queries vs Y 20000 x 20000, d=0.002, "
                                             "NB/log1p, precision=exact"},
                      "methods": res}), flush=True)


def sweep_main(args, sd, synth, workers):
    """Times each search three ways (forced densified path, forced FP64 merge
    walk, precision="auto") over reference density and query batch size, and
    reports which path the cost model chose and whether it was the faster."""
    dens = (0.0005, 0.002, 0.01, 0.06)
    sizes = (1, 32, 2048)
    refs = {d: synth.nb_log1p_parallel(20_000, 20_000, d, 80 + i, workers)
            for i, d in enumerate(dens)}
    qrys = {d: synth.nb_log1p_parallel(20_000, max(sizes), d, 90 + i, workers)
            for i, d in enumerate(dens)}
    ctx = sd.default_context()
    rates = ctx.autotune()
    log(f"calibrated rates: {rates}")
    rows = []
    for d in dens:
        dY = sd.DeviceMatrix(refs[d], ctx)
        for n in sizes:
            X = qrys[d][:, :n]
            for method in ("binary", "cosine"):
                t = {}
                for prec in ("bf16", "exact", "auto"):
                    best = 1e30
                    for _ in range(3):
                        ctx.stats(reset=True)
                        t0 = time.perf_counter()
                        sd.sparseKNN(X, dY, k=10, method=method, precision=prec, verbose=False,
                                     context=ctx)
                        best = min(best, time.perf_counter() - t0)
                        st = ctx.stats(reset=True)
                    t[prec] = best
                    if prec == "auto":
                        chose = "exact" if st["exact_calls"] else "densified"
                faster = "exact" if t["exact"] < t["bf16"] else "densified"
                rows.append(dict(density=d, query_cols=n, method=method,
                                 densified_ms=t["bf16"] * 1e3, merge_walk_ms=t["exact"] * 1e3,
                                 auto_ms=t["auto"] * 1e3, auto_chose=chose,
                                 faster=faster, right=(chose == faster)))
                log(f"d={d} n={n} {method}: densified {t['bf16'] * 1e3:.2f} ms, merge walk "
                    f"{t['exact'] * 1e3:.2f} ms, auto {t['auto'] * 1e3:.2f} ms ({chose})")
        dY.free()
    print(json.dumps({"metric": "path chooser sweep (ms per search)", "rates": rates,
                      "config": {"workload": "references 20000 x 20000 NB/log1p at densities "
                                             f"{list(dens)}; query batches {list(sizes)}; k=10"},
                      "rows": rows,
                      "right_choices": sum(r["right"] for r in rows), "cases": len(rows)}),
          flush=True)


if __name__ == "__main__":
    main()
