#!/usr/bin/env python3
"""Search-throughput benchmark of the MI355X COTHER hot path.

Workload: 64 query profiles (lengths lognormal, median 250, sigma 0.35,
clipped 40-1000, predicted distance maps) against an HBM-resident database of
60,000 template profiles (lengths lognormal, median 220, sigma 0.35, clipped
30-1500, C-alpha chain distances; about 14.0M positions), all seeded
synthetic profiles in the reference's value conventions.

A step is one query searched against the whole resident database: score
matrix (log-odds + SS + CVS + DDMS), phase-1 DP with traceback and
finalisation, statistics and E-values of the passed profiles, E-value filter,
MAP realignment (forward, backward, MAP DP, traceback) of the significant
hits and hit ordering.  Step s uses query s mod 64.  Cells = Lq x database
positions; value = cells of the timed steps / wall time of the timed steps.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "GCUPS (score+DP+MAP) over resident DB; p50/p99 query latency vs plain FP32"

# MI355X published peaks
PEAK_VALU_FP32_TFLOPS = 157.3   # FP32 vector
PEAK_HBM_GBS = 8000.0

# algorithmic FP32 operations per score cell outside DDMS: log-odds over 20
# amino acids (5 ops each + log), SS (9 products, 9 table terms, H term),
# CVS (19 multiply-adds, norms, determinant, log, piecewise-linear map)
OPS_CELL_BASE = 216
# algorithmic FP32 operations per DDMS segment-cell update: add, clamp at 0,
# compare with the step maximum, plus one per update for the interpolated cells
OPS_DDMS_UPDATE = 4
DSTSEGM = 6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def ddms_updates_table(nmax=128):
    """U[n1][n2]: segment-cell updates of one DDMS evaluation for list lengths
    n1 (query) and n2 (template): upper diagonals j = n2>>2 .. 0 step 2 with
    ceil(min(n1, n2-j)/4) steps, lower diagonals i = n1>>2 .. 1 step 2 with
    ceil(min(n2, n1-i)/4) steps, 4 updates per step; 0 if either list is
    shorter than 6 (CalcDDMS2ScoreSMEMh2, cusco/CuBatchScoreMatrix_ddms2s_h2.cuh:435-500)."""
    U = np.zeros((nmax, nmax), np.int64)
    n = np.arange(nmax)
    for n1 in range(DSTSEGM, nmax):
        for n2 in range(DSTSEGM, nmax):
            s = 0
            for j in range(n2 >> 2, -1, -2):
                s += (min(n1, n2 - j) + 3) // 4
            for i in range(n1 >> 2, 0, -2):
                s += (min(n2, n1 - i) + 3) // 4
            U[n1, n2] = 4 * s
    return U


def algorithmic_ops(qlens_hist, db_hist, lq, npos, U):
    """Algorithmic FP32 ops of one query's score matrix against the database."""
    ddms = float(qlens_hist @ U @ db_hist)
    return OPS_DDMS_UPDATE * ddms + OPS_CELL_BASE * float(lq) * float(npos)


def read_traffic():
    """HBM bytes per score cell of the score kernel from the committed rocprofv3
    --pmc summary (line 'bytes_per_cell: X'), or None."""
    p = os.path.join(ROOT, "profiles", "k_score_db6000_pmc.txt")
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("bytes_per_cell:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(q, db, target_s, threads):
    """The CPU oracle (the C restatement of the same pipeline, OpenMP over
    templates) on a bounded sample: query 0 against a contiguous block of
    templates from the middle of the length-sorted database, grown until the
    run takes about target_s seconds."""
    sys.path.insert(0, ROOT)
    from tests import _oracle
    orc = _oracle.load()
    lens = db.lens()
    lq = int(q.lens()[0])
    p0 = db.nprof // 2
    nblk = 8
    while True:
        p1 = min(db.nprof, p0 + nblk)
        t0 = time.perf_counter()
        orc.search_query(q, 0, db, p0, p1, 10.0, 2, db.npos, db.nprof, threads)
        dt = time.perf_counter() - t0
        cells = float(lq) * float(lens[p0:p1].sum())
        if dt >= target_s or p1 == db.nprof:
            break
        nblk = int(nblk * min(8.0, max(1.5, 1.2 * target_s / max(dt, 1e-3))))
    return {"value": cells / dt / 1e9, "unit": "GCUPS", "cores": threads, "kind": "port",
            "host_cpus": os.cpu_count(),
            "sample": f"query 0 (Lq={lq}) vs templates [{p0},{p1}) ({int(lens[p0:p1].sum())} positions), "
                      f"{cells:.3g} cells in {dt:.1f} s, full pipeline, {threads} OpenMP threads "
                      f"(the process's share of the host's {os.cpu_count()} CPUs)"}


def serve_phase(ctx, qs, db, args, seq_qps, rate=None, n=None, contraction=None, first=None):
    """Requests arriving as a seeded Poisson process at serve_load x the sequential query
    rate (or `rate` per second), run by the serving loop (slots overlap on the device,
    admission by buffer footprint); latency is submit -> done, throughput is cells over
    first submit -> last done."""
    import threading
    import cother_amd
    n = args.serve_requests if n is None else n
    ql = qs.lens()
    first = args.warmup if first is None else first
    order = [(first + s) % qs.nprof for s in range(n)]
    if rate is None:
        rate = args.serve_rate if args.serve_rate > 0 else seq_qps * args.serve_load
    contraction = args.contraction if contraction is None else contraction
    rng = np.random.default_rng(args.seed + 7)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n))
    srv = cother_amd.Server(ctx, nslots=args.serve_slots, mem_budget=64 << 30, chunk_cols=4 << 20, eval=args.eval,
                            contraction=contraction)
    tickets = [None] * n
    ready = [threading.Event() for _ in range(n)]

    def submitter():
        t0 = time.perf_counter()
        for k, qi in enumerate(order):
            dt = arrivals[k] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            tickets[k] = srv.submit(qi)
            ready[k].set()

    log(f"serving: {n} requests, Poisson at {rate:.3f}/s ({rate / seq_qps:.2f}x the sequential rate), "
        f"{args.serve_slots} slots")
    t_start = time.perf_counter()
    th = threading.Thread(target=submitter)
    th.start()
    lat, queued, cells = [], [], 0.0
    for k in range(n):
        ready[k].wait()
        _, _, info = srv.wait(tickets[k])
        lat.append(info["ms_latency"])
        queued.append(info["ms_queued"])
        cells += float(ql[order[k]]) * float(db.npos)
        log(f"request {k}: query {order[k]} Lq={ql[order[k]]} latency {info['ms_latency']:.0f} ms "
            f"(queued {info['ms_queued']:.0f} ms) slot {info['slot']}")
    th.join()
    makespan = time.perf_counter() - t_start
    srv.close()
    lat = np.array(lat)
    return {"requests": n, "slots": args.serve_slots, "offered_qps": round(rate, 4),
            "contraction": ("fp32", "bf16-mfma", "fp16-mfma")[contraction],
            "offered_load": round(rate / seq_qps, 3),
            "gcups": round(cells / makespan / 1e9, 4), "p50_ms": round(float(np.percentile(lat, 50)), 2),
            "p99_ms": round(float(np.percentile(lat, 99)), 2), "mean_queued_ms": round(float(np.mean(queued)), 2),
            "mem_budget_gib": 64, "chunk_cols": 4 << 20}


def served_burst(ctx, q, db, eval_, batch, nslots=2, bursts=4, wait_us=0):
    """All of q's queries submitted at once to a serving loop over the resident
    database (budget for the whole database per request, so that a tick can
    batch them), `bursts` times on one server (its slots keep their buffers,
    tuned grid sizes and graphs, as a long-running server's do); the last
    burst is measured: GCUPS over first submit -> last done, latency
    percentiles, the largest tick."""
    import cother_amd
    srv = cother_amd.Server(ctx, nslots=nslots, mem_budget=64 << 30, chunk_cols=0, eval=eval_, batch=batch,
                            batch_wait_us=wait_us)
    for _ in range(bursts):
        t0 = time.perf_counter()
        tickets = [srv.submit(qi) for qi in range(q.nprof)]
        lat, groups = [], []
        for t in tickets:
            _, _, info = srv.wait(t)
            lat.append(info["ms_latency"])
            groups.append(info["group"])
        wall = time.perf_counter() - t0
    srv.close()
    lat = np.array(lat)
    return {"batch": batch, "slots": nslots, "batch_wait_us": wait_us,
            "gcups": round(float(q.lens().sum()) * db.npos / wall / 1e9, 4),
            "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
            "largest_tick": int(max(groups)), "bursts": bursts}


def batch_phase(args, threads, tdist):
    """Many short queries against a small resident database, searched one
    query at a time (cth_search) and as one batch (cth_search_batch: one
    launch per stage over all pairs, device-side selection, graph replay):
    32 queries (Lq lognormal median 90, clipped 40-150, predicted maps)
    against 4000 templates (lognormal median 220, clipped 30-400; C-alpha chain
    distance lists, or none with tdist 2).  Both paths' hits are compared."""
    import cother_amd
    q = cother_amd.synth(32, 1, 90.0, 0.35, 40, 150, dist_mode=0, sort_desc=0, seed=args.seed + 21, threads=threads)
    db = cother_amd.synth(4000, 1, 220.0, 0.35, 30, 400, dist_mode=tdist, sort_desc=1, seed=args.seed + 22,
                          threads=threads)
    ctx = cother_amd.Context(0)
    ctx.upload_queries(q)
    ctx.upload_db(db)
    cells = float(q.lens().sum()) * db.npos
    ctx.search(0, 2, eval=args.eval)                    # warmup
    ctx.search(eval=args.eval, batch=True)              # autotune + graph capture
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        for qi in range(q.nprof):
            a, na, _ = ctx.search(qi, qi + 1, eval=args.eval)
    per_query = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(reps):
        b, nb, sb = ctx.search(eval=args.eval, batch=True)
    batched = time.perf_counter() - t0
    ref, nref, _ = ctx.search(eval=args.eval)
    same = len(ref) == len(b) and all((ref[f] == b[f]).all() for f in ("query", "profile", "score", "logE", "alnlen"))
    caps, reps_g = ctx.graph_counts()
    # the serving loop on the same burst of requests: one query per tick,
    # then ticks of up to all of them, started at once or after a 2 ms
    # batching window (the last of four bursts on a server)
    served = [served_burst(ctx, q, db, args.eval, nb) for nb in (1, q.nprof)]
    served.append(served_burst(ctx, q, db, args.eval, q.nprof, wait_us=2000))
    ctx.close()
    g1, g2 = cells * reps / per_query / 1e9, cells * reps / batched / 1e9
    log(f"batch phase (template distances {'none' if tdist == 2 else 'C-alpha'}): per-query {g1:.3f} GCUPS, "
        f"batched {g2:.3f} GCUPS ({g2 / g1:.2f}x), hits equal: {same}; served burst, one query per tick "
        f"{served[0]['gcups']:.3f} GCUPS, batched ticks {served[1]['gcups']:.3f} GCUPS "
        f"(largest tick {served[1]['largest_tick']}), with a 2 ms batching window {served[2]['gcups']:.3f} GCUPS "
        f"(largest tick {served[2]['largest_tick']})")
    return {"workload": "32 queries (Lq lognormal median 90, 40-150) vs 4000 templates (lognormal median 220, "
                        f"30-400), template distance lists: {'none' if tdist == 2 else 'C-alpha chains'}",
            "cells": cells, "per_query_gcups": round(g1, 4), "batched_gcups": round(g2, 4),
            "speedup": round(g2 / g1, 3), "hits": int(len(b)), "hits_equal": bool(same),
            "batches_per_search": sb.nchunks, "graph_captures": caps, "graph_replays": reps_g,
            "served_burst": served}


def long_profile_run(args, threads):
    """One query of 2000 positions with a full predicted distance map against 5000
    templates of 1000-3000 positions (C-alpha chain distances): long DP/MAP rows, a
    database processed in several chunks; one JSON line of its own."""
    import cother_amd
    q = cother_amd.synth(1, 0, 0.0, 0.0, 2000, 2000, dist_mode=0, sort_desc=0, seed=args.seed + 11, threads=threads)
    db = cother_amd.synth(5000, 0, 0.0, 0.0, 1000, 3000, dist_mode=1, sort_desc=1, seed=args.seed + 12,
                          threads=threads)
    ctx = cother_amd.Context(0)
    ctx.upload_queries(q)
    ctx.upload_db(db)
    log(f"long profile: Lq={int(q.lens()[0])} vs {db.nprof} templates ({db.npos} positions)")
    ctx.search(0, 1, eval=args.eval, contraction=args.contraction)  # warmup
    t1 = time.perf_counter()
    hits, nh, st = ctx.search(0, 1, eval=args.eval, timing=True, contraction=args.contraction)
    wall = time.perf_counter() - t1
    out = {"metric": METRIC, "value": round(st.cells / wall / 1e9, 4), "unit": "GCUPS", "n_gpus": 1, "steps": 1,
           "warmup": 1, "ms_per_step": round(wall * 1e3, 2), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "long-profile stress: one query of 2000 positions (predicted distance map) vs "
                                  "5000 templates of 1000-3000 positions",
                      "db_positions": int(db.npos), "cells": st.cells, "chunks": st.nchunks,
                      "contraction": ("fp32", "bf16-mfma", "fp16-mfma")[args.contraction]},
           "stage_ms": {"score": round(st.ms_k_score, 1), "dp": round(st.ms_k_dp, 1), "stats": round(st.ms_k_stats, 1),
                        "map": round(st.ms_k_map, 1)},
           "hits": int(nh[0]), "passed": st.npassed, "map_cells": st.map_cells,
           "longest_alignment": int(hits["alnlen"].max()) if len(hits) else 0}
    print(json.dumps(out), flush=True)
    ctx.close()


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--queries", type=int, default=64)
    ap.add_argument("--templates", type=int, default=60000)
    ap.add_argument("--eval", type=float, default=10.0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline sample size in seconds (0: skip)")
    ap.add_argument("--seed", type=int, default=20260)
    ap.add_argument("--serve-requests", type=int, default=16,
                    help="serving phase: requests (the timed queries again) arriving as a Poisson process (0: skip)")
    ap.add_argument("--serve-load", type=float, default=1.5,
                    help="serving phase: offered load as a multiple of the sequential query rate "
                         "(above 1 the queue never drains, so the phase's GCUPS is the sustained rate)")
    ap.add_argument("--serve-rate", type=float, default=0.0,
                    help="serving phase: absolute arrival rate in queries/s (overrides --serve-load)")
    ap.add_argument("--serve-slots", type=int, default=2)
    ap.add_argument("--serve-sweep", type=int, default=6, metavar="N",
                    help="Poisson arrivals at 20, 50 and 100 queries/s, FP32 and BF16 contraction, N requests each "
                         "(0: skip)")
    ap.add_argument("--batch-phase", type=int, default=1, choices=(0, 1),
                    help="per-query vs batched search of 32 short queries against 4000 templates, with and "
                         "without template distance lists, reported in 'batched' (0: skip)")
    ap.add_argument("--map-heavy", type=int, default=4, metavar="N",
                    help="N queries at E-value 1e4 (many hits realigned by MAP), reported in 'map_heavy' (0: skip)")
    ap.add_argument("--contraction", type=int, default=0, choices=(0, 1, 2),
                    help="log-odds term: 0 FP32 (the reference's arithmetic), 1 BF16 / 2 FP16 matrix-core variant")
    ap.add_argument("--long-profile", action="store_true",
                    help="long-profile stress instead: one query of 2000 positions (predicted distance map) "
                         "against 5000 templates of 1000-3000 positions")
    ap.add_argument("--workload", nargs=8, type=float, default=None,
                    metavar=("QMED", "QMIN", "QMAX", "TMED", "TMIN", "TMAX", "QDIST", "TDIST"),
                    help="other query/template length distributions (lognormal median, clip range, sigma 0.35) "
                         "and distance modes (0 predicted maps, 1 C-alpha chains, 2 none) in place of the "
                         "benchmark's 250/40/1000, 220/30/1500, 0, 1")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel over BYTES (traffic-counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the search path shards over GPUs by database chunk "
              "(one process per GPU), and only the single-GPU figure is produced here; run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import cother_amd
    if args.copy_calibration:
        ms = ctypes.c_float()
        rc = cother_amd.lib().cth_bench_copy(0, args.copy_calibration, 3, ctypes.byref(ms))
        if rc:
            raise RuntimeError(cother_amd.lib().cth_last_error().decode())
        print(json.dumps({"copy_bytes": args.copy_calibration, "launches": 4, "ms_per_launch": ms.value,
                          "GBps": 2 * args.copy_calibration / (ms.value * 1e-3) / 1e9}), flush=True)
        return
    # the host threads this process may use: the OpenMP setting where the
    # environment fixes it (the GPU host's share of cores), else the CPUs the
    # process is allowed to run on
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    if args.long_profile:
        long_profile_run(args, threads)
        return
    t0 = time.perf_counter()
    wl = args.workload or (250, 40, 1000, 220, 30, 1500, 0, 1)
    qs = cother_amd.synth(args.queries, 1, wl[0], 0.35, int(wl[1]), int(wl[2]), dist_mode=int(wl[6]), sort_desc=0,
                          seed=args.seed, threads=threads)
    db = cother_amd.synth(args.templates, 1, wl[3], 0.35, int(wl[4]), int(wl[5]), dist_mode=int(wl[7]), sort_desc=1,
                          seed=args.seed + 1, threads=threads)
    log(f"workload: {qs.nprof} queries ({qs.npos} positions), {db.nprof} templates ({db.npos} positions); "
        f"generated in {time.perf_counter() - t0:.1f} s")
    ctx = cother_amd.Context(0)
    ctx.upload_queries(qs)
    ctx.upload_db(db)
    log("uploaded")

    U = ddms_updates_table()
    db_hist = np.bincount(db.field("dd")[:, 0], minlength=128)[:128].astype(np.float64)
    qdd = qs.field("dd")[:, 0]
    qd, ql = qs.dists(), qs.lens()

    for s in range(args.warmup):
        qi = s % qs.nprof
        ctx.search(qi, qi + 1, eval=args.eval, contraction=args.contraction)
        log(f"warmup step {s} (query {qi}, Lq={ql[qi]}) done")

    lat, cells, ops, ms_score, nlaunch, ms_dp, ms_st, npass = [], 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0
    ms_map, map_cells, nhits = 0.0, 0.0, 0
    t_all = time.perf_counter()
    for s in range(args.steps):
        qi = (args.warmup + s) % qs.nprof
        t1 = time.perf_counter()
        hits, nh, st = ctx.search(qi, qi + 1, eval=args.eval, timing=True, contraction=args.contraction)
        lat.append((time.perf_counter() - t1) * 1e3)
        cells += st.cells
        ms_score += st.ms_k_score
        ms_dp += st.ms_k_dp
        ms_st += st.ms_k_stats
        ms_map += st.ms_k_map
        map_cells += st.map_cells
        nhits += int(nh[0])
        nlaunch += st.nchunks
        npass += st.npassed
        qh = np.bincount(qdd[qd[qi]:qd[qi] + ql[qi]], minlength=128)[:128].astype(np.float64)
        ops += algorithmic_ops(qh, db_hist, ql[qi], db.npos, U)
        log(f"step {s}: query {qi} Lq={ql[qi]} {lat[-1]:.0f} ms, {int(nh[0])} hits, {int(st.npassed)} passed, "
            f"score {st.ms_k_score:.0f} ms dp {st.ms_k_dp:.0f} ms stats {st.ms_k_stats:.0f} ms "
            f"map {st.ms_k_map:.0f} ms")
    wall = time.perf_counter() - t_all
    gcups = cells / wall / 1e9

    achieved = ops / (ms_score * 1e-3) / 1e12 if ms_score > 0 else 0.0
    bpc = read_traffic()
    cells_per_launch = cells / max(nlaunch, 1.0)
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_VALU_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_VALU_FP32_TFLOPS, 4),
                "traffic": (bpc * cells_per_launch) if bpc is not None else None,
                "kernel": "k_score", "avg_launch_ms": round(ms_score / max(nlaunch, 1.0), 3),
                "ops_per_launch": ops / max(nlaunch, 1.0)}
    serving = None
    if args.serve_requests > 0:
        serving = serve_phase(ctx, qs, db, args, args.steps / wall)
    sweep = None
    if args.serve_sweep > 0:
        sweep = []
        for prec in (0, 1):
            for rate in (20.0, 50.0, 100.0):
                r = serve_phase(ctx, qs, db, args, args.steps / wall, rate=rate, n=args.serve_sweep,
                                contraction=prec, first=args.warmup + args.steps)
                log(f"sweep: {r}")
                sweep.append(r)
    batched = None
    if args.batch_phase:
        batched = [batch_phase(args, threads, 1), batch_phase(args, threads, 2)]
    map_heavy = None
    if args.map_heavy > 0:
        mh_cells, mh_map_ms, mh_map_cells, mh_hits = 0.0, 0.0, 0.0, 0
        t1 = time.perf_counter()
        for s in range(args.map_heavy):
            qi = (args.warmup + s) % qs.nprof
            hits, nh, st = ctx.search(qi, qi + 1, eval=1e4, timing=True, contraction=args.contraction)
            mh_cells += st.cells
            mh_map_ms += st.ms_k_map
            mh_map_cells += st.map_cells
            mh_hits += int(nh[0])
            log(f"map-heavy {s}: query {qi}, {int(nh[0])} hits, map {st.ms_k_map:.0f} ms")
        mh_wall = time.perf_counter() - t1
        map_heavy = {"queries": args.map_heavy, "eval": 1e4, "gcups": round(mh_cells / mh_wall / 1e9, 4),
                     "hits_per_query": round(mh_hits / args.map_heavy, 1), "map_ms": round(mh_map_ms, 1),
                     "map_cells": mh_map_cells,
                     "map_gcups": round(mh_map_cells / (mh_map_ms * 1e-3) / 1e9, 4) if mh_map_ms > 0 else None,
                     "ms_per_query": round(mh_wall * 1e3 / args.map_heavy, 1)}
    cpu = None
    if args.cpu_seconds > 0:
        log("cpu baseline ...")
        cpu = cpu_baseline(qs, db, args.cpu_seconds, threads)
    lat_a = np.array(lat)
    out = {
        "metric": METRIC, "value": round(gcups, 4), "unit": "GCUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": (f"{qs.nprof} queries (Lq lognormal median {wl[0]:g}, {wl[1]:g}-{wl[2]:g}) vs resident "
                                f"DB of {db.nprof} templates (lognormal median {wl[3]:g}, {wl[4]:g}-{wl[5]:g}); one "
                                f"query per step"),
                   "queries": qs.nprof, "templates": db.nprof, "db_positions": int(db.npos),
                   "query_positions_timed": int(sum(ql[(args.warmup + s) % qs.nprof] for s in range(args.steps))),
                   "eval": args.eval, "map": True,
                   "contraction": ("fp32", "bf16-mfma", "fp16-mfma")[args.contraction]},
        "p50_ms": round(float(np.percentile(lat_a, 50)), 2), "p99_ms": round(float(np.percentile(lat_a, 99)), 2),
        "stage_ms": {"score": round(ms_score, 1), "dp": round(ms_dp, 1), "stats": round(ms_st, 1),
                     "map": round(ms_map, 1)},
        "passed_per_query": round(npass / args.steps, 1), "hits_per_query": round(nhits / args.steps, 1),
        "map_cells": map_cells,
        "roofline": roofline, "cpu_baseline": cpu, "serving": serving, "serving_sweep": sweep,
        "map_heavy": map_heavy, "batched": batched,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
