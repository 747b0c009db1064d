"""Benchmark: symbolic factorisation of the 659,033-row 9-point ND workload on one MI355X.

A step is one full run of the engine on the resident pattern: ELL packing, transpose, both
reachability directions, supernode detection and the explicit L/U structure. Prints one JSON line.

  value        W-GTEPS = W / step time, W = schedule-independent edge checks of the row
               traversals (a unit of work fixed by the input, as for the reference's TEPS). The
               engine reuses reached sets between rows and executes far fewer inspections than W:
               "executed" reports what the kernels really did, per second.
  roofline     the dominant kernel (workgroup bitmap path), priced on the work it executed:
               8 B per neighbour inspection + 12 B per expanded row + 4 B per F entry written,
               over its event-timed duration; "step_roofline" prices the whole step the same way
  batched      the 256-matrix batch in one captured graph against one plan per member
  served       a Poisson stream (200 requests/s) into the continuous-batching server against
               plain sequential plans under the same arrivals, and both saturated
  reference_dataset  the reference's own 11,948-row input, beside its published V100 time
  cpu_baseline the oracle (host C++ restatement, OpenMP over sources) on this host's cores:
               one warm-up, then the median of three timed passes over a bounded seeded sample of
               the sources (--cpu-full: all 659,033 sources)
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GTEPS + fill-nnz/s at 659k-row shape; batched matrices/s & p99 vs plain"
HBM_PEAK_GBS = 8000.0
# The reference's one published timing: its 11,948-row input (BC) on one V100, space-optimised
# build, 11 runs (SpaceOptimization/with_supernodes/space_optimization_log_1_264_GPU_warp_centric.dat:4-124)
REFERENCE_DS_MS = {"min": 35.634, "median": 37.807, "max": 42.123, "hardware": "NVIDIA V100"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="grid2d_659033", choices=["grid2d_659033", "grid3d_110592", "grid2d_4096"])
    ap.add_argument("--chunk", type=int, default=128)
    ap.add_argument("--no-structure", action="store_true")
    ap.add_argument("--early-top", action="store_true",
                    help="start the top separator's items beside the wave path (gsofa_options.early_top = 1)")
    ap.add_argument("--cpu-sample", type=int, default=150000,
                    help="sources in the CPU baseline's seeded sample (each of 3 timed passes)")
    ap.add_argument("--cpu-full", action="store_true", help="CPU baseline over all sources (minutes)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic_heavy_grid2d_659033.json"),
                    help="rocprofv3 FETCH_SIZE/WRITE_SIZE summary for the dominant kernel (optional)")
    ap.add_argument("--batch-members", type=int, default=256,
                    help="members of the 256-matrix batch to run batched vs plain (0 = skip)")
    ap.add_argument("--batch-config", default="",
                    help="heavy_threads,wgs_per_cu,chain_block for the batch instead of autotuning it")
    ap.add_argument("--batch-only", action="store_true",
                    help="only replay the batch (--steps replays after one warm-up), for kernel traces")
    ap.add_argument("--served-requests", type=int, default=1000,
                    help="requests of the Poisson stream to serve (0 = skip)")
    ap.add_argument("--served-rate", type=float, default=200.0)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--cu-share", type=int, default=0, help="CUs per served request (0 = all)")
    ap.add_argument("--no-reference-dataset", action="store_true")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def fnv(a):
    import hashlib
    return hashlib.sha256(a.tobytes()).hexdigest()[:16]


def generate_many(g, workload, members, threads=8):
    """Patterns of a workload's members; the generator runs in native code, so threads overlap."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(lambda k: g.generate_workload(workload, k), range(members)))


def batched_vs_plain(g, members, config=None, only_replays=0):
    """The batch workload: batched (one captured graph for all members) vs one plan per member
    run back to back. Per-member results (totals, nzU and supernode flags) must be identical."""
    pats = generate_many(g, g.WL_BATCH256, members)
    log(f"[batched] {members} members, n total {sum(p.n for p in pats)}")
    t0 = time.perf_counter()
    if config:
        b = g.Batch(pats, heavy_threads=config[0], heavy_wgs_per_cu=config[1], chain_block=config[2])
    else:
        b = g.Batch(pats, autotune=1)  # shape and row-block size timed on this set (not in the timed region)
    t_tune = time.perf_counter() - t0
    cfg = b.config()
    log(f"[batched] {'set created' if config else 'autotuned'} in {t_tune:.1f} s: {cfg}")
    b.run().sync()  # capture + warm-up
    if only_replays:
        t0 = time.perf_counter()
        for _ in range(only_replays):
            b.run().sync()
        t_batch = (time.perf_counter() - t0) / only_replays
        b.close()
        log(f"[batched] {only_replays} replays, {t_batch * 1e3:.1f} ms each")
        return {"members": members, "batched_ms": t_batch * 1e3, "batched_matrices_per_s": members / t_batch,
                "config": cfg}
    reps = 2
    t0 = time.perf_counter()
    for _ in range(reps):
        b.run().sync()
    t_batch = (time.perf_counter() - t0) / reps
    keys = ("fill_count", "nnz_F", "edge_checks", "reach", "supernodes")
    batch_res = [tuple(getattr(b.result(i), k) for k in keys) + (fnv(b.nzu(i)), fnv(b.supernode_flags(i)))
                 for i in range(members)]
    b.close()
    t_plain = 0.0
    mismatches = 0
    for i, p in enumerate(pats):
        with g.Plan(p) as plan:
            plan.run().sync()  # warm-up (allocation and first launch excluded, as for the batch)
            t0 = time.perf_counter()
            plan.run().sync()
            t_plain += time.perf_counter() - t0
            r = plan.result()
            mismatches += tuple(getattr(r, k) for k in keys) + (fnv(plan.nzu()), fnv(plan.supernode_flags())) \
                != batch_res[i]
    log(f"[batched] batch {t_batch * 1e3:.1f} ms, plain {t_plain * 1e3:.1f} ms, mismatches {mismatches}")
    return {"members": members, "batched_matrices_per_s": members / t_batch, "plain_matrices_per_s": members / t_plain,
            "speedup": t_plain / t_batch, "batched_ms": t_batch * 1e3, "plain_ms": t_plain * 1e3,
            "bit_exact_vs_plain": mismatches == 0,
            "compared": "fill, nnz_F, W, |R|, supernodes, nzU and supernode flags of every member",
            "tuned_config": cfg, "autotune_s": t_tune}


def served_vs_plain(g, nreq, rate, slots, cu_share=0):
    """The served stream: Poisson arrivals (seeded, `rate` per second) of stream requests.

    served   each request submitted at its arrival time to the continuous-batching server;
             latency = submit to result
    plain    the same arrivals handled one at a time by plain plans (create, run, read the
             result, destroy): latency = arrival to result, including waiting for the previous
             request (what a one-matrix-per-launch service would give)
    saturated  all requests queued at once: matrices/s of the server against the plain form's
             sum of per-request times
    Results must be identical served and plain."""
    import numpy as np
    rng = np.random.default_rng(0x5EED0005)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
    t0 = time.perf_counter()
    pats = generate_many(g, g.WL_STREAM, nreq)
    ns = np.array([p.n for p in pats])
    log(f"[served] {nreq} requests generated in {time.perf_counter() - t0:.1f} s, n from {ns.min()} to {ns.max()}, "
        f"arrivals over {arrivals[-1]:.2f} s")
    keys = ("fill_count", "nnz_F", "edge_checks", "reach", "supernodes")

    def serve(at):
        srv = g.Server(slots=slots, cu_share=cu_share)
        ids = []
        start = time.perf_counter()
        for p, a in zip(pats, at):
            delay = start + a - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            ids.append(srv.submit(p))
        res, lat = [], []
        for rid in ids:
            r = srv.wait(rid, 600.0)
            info = srv.info(rid)
            res.append(tuple(getattr(r, k) for k in keys) + (info["nzu_digest"], info["supernode_digest"]))
            lat.append(info["latency_ms"])
            srv.release(rid)
        wall = time.perf_counter() - start
        st = srv.stats()
        kept = srv.retained()
        srv.stop()
        return res, np.array(lat), wall, st, kept

    res_s, lat_s, wall_s, st_s, kept = serve(arrivals)
    log(f"[served] Poisson: wall {wall_s:.2f} s, p50 {np.percentile(lat_s, 50):.1f} ms, "
        f"p99 {np.percentile(lat_s, 99):.1f} ms, peak concurrent {st_s['peak_running']}, retained {kept}")
    # plain sequential under the same arrivals
    res_p, lat_p, service, run_only = [], [], [], []
    start = time.perf_counter()
    for p, a in zip(pats, arrivals):
        delay = start + a - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        t1 = time.perf_counter()
        with g.Plan(p) as plan:
            t2 = time.perf_counter()
            plan.run().sync()
            run_only.append(time.perf_counter() - t2)
            r = plan.result()
            res_p.append(tuple(getattr(r, k) for k in keys) +
                         (g.fnv1a_array(plan.nzu()), g.fnv1a_array(plan.supernode_flags())))
        t3 = time.perf_counter()
        service.append(t3 - t1)
        lat_p.append((t3 - (start + a)) * 1e3)
    wall_p = time.perf_counter() - start
    lat_p = np.array(lat_p)
    log(f"[served] plain: wall {wall_p:.2f} s, p50 {np.percentile(lat_p, 50):.1f} ms, "
        f"p99 {np.percentile(lat_p, 99):.1f} ms, service {sum(service):.2f} s (runs {sum(run_only):.2f} s)")
    # saturated: everything queued at t = 0
    res_t, lat_t, wall_t, st_t, _ = serve(np.zeros(nreq))
    log(f"[served] saturated: wall {wall_t:.2f} s ({nreq / wall_t:.0f}/s) vs plain {nreq / sum(service):.0f}/s")
    same = res_s == res_p and res_t == res_p
    return {"requests": nreq, "rate_per_s": rate, "slots": slots, "cu_share": cu_share,
            "n_range": [int(ns.min()), int(ns.max())], "arrival_span_s": float(arrivals[-1]),
            "matrices_per_s": nreq / wall_s,
            "p50_ms": float(np.percentile(lat_s, 50)), "p99_ms": float(np.percentile(lat_s, 99)),
            "peak_concurrent": st_s["peak_running"],
            "plain_sequential": {"matrices_per_s": nreq / wall_p, "p50_ms": float(np.percentile(lat_p, 50)),
                                 "p99_ms": float(np.percentile(lat_p, 99)),
                                 "per_request_ms_mean": 1e3 * float(np.mean(service)),
                                 "run_only_ms_mean": 1e3 * float(np.mean(run_only)),
                                 "what": "one plan per request: create (prepare, allocate, upload), run, read "
                                         "the result, destroy; requests handled in arrival order"},
            "saturated": {"served_matrices_per_s": nreq / wall_t, "plain_matrices_per_s": nreq / sum(service),
                          "speedup": (nreq / wall_t) / (nreq / sum(service)),
                          "served_p99_ms": float(np.percentile(lat_t, 99)), "peak_concurrent": st_t["peak_running"]},
            "retained_after_release": kept,
            "identical_alone_and_in_mix": bool(same)}


def reference_dataset(g, runs=11):
    """The reference's 11,948-row input (BC), S = 128: median of `runs` timed runs."""
    import numpy as np
    z = np.load(os.path.join(ROOT, "tests", "golden", "ds_11948.npz"))
    p = g.Pattern(int(len(z["beg"])), z["beg"], z["end"], z["adj"])
    out = {}
    for structure in (0, 1):
        with g.Plan(p, chunk_size=128, structure=structure) as plan:
            plan.run().sync()
            times = []
            for _ in range(runs):
                plan.run().sync()
                times.append(plan.result().time_ms)
            r = plan.result()
        out["structure" if structure else "counts_only"] = {"median_ms": float(np.median(times)),
                                                            "min_ms": float(np.min(times))}
    out.update({"n": p.n, "nnz": p.nnz, "fill_count": r.fill_count, "supernodes": r.supernodes,
                "runs": runs, "reference_v100_ms": REFERENCE_DS_MS,
                "note": "context only: other hardware, and the reference times its traversal and supernodes "
                        "(...cuh:2944-3491) while this times the whole device sequence"})
    log(f"[reference dataset] {out['counts_only']['median_ms']:.3f} ms counts only, "
        f"{out['structure']['median_ms']:.3f} ms with structure; fill {r.fill_count}")
    return out


def host_facts():
    model = platform.processor() or ""
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True).stdout.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model, "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(pattern, sample, full=False):
    """The oracle over a fixed seeded uniform sample of sources (or all of them): one warm-up
    pass, then the median of three timed passes, all host threads."""
    import numpy as np
    import oracle
    threads = oracle.max_threads()
    rng = np.random.default_rng(0x5EED0003)
    srcs = np.arange(pattern.n, dtype=np.int32) if full else \
        np.sort(rng.choice(pattern.n, size=min(sample, pattern.n), replace=False)).astype(np.int32)
    w, _ = oracle.sources_work(pattern, srcs, threads=threads)  # untimed: the metric's numerator
    oracle.sources(pattern, srcs[: max(1, len(srcs) // 4)], threads=threads)  # warm-up
    times, checks, fills = [], 0, 0
    for _ in range(3):
        t0 = time.perf_counter()
        fills, checks = oracle.sources(pattern, srcs, threads=threads)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    return {"value": w / t / 1e9, "unit": "GTEPS", "cores": threads, "kind": "port",
            "sample": (f"all {pattern.n} sources" if full else
                       f"{len(srcs)} of {pattern.n} sources (seeded uniform sample)") +
                      f"; one warm-up pass over a quarter of them, then the median of 3 passes: "
                      f"{', '.join(f'{x:.2f}' for x in times)} s; label-correcting edge inspections "
                      f"{checks} = {checks / max(w, 1):.2f} x W",
            "fill_nnz_per_s": fills / t, "seconds_per_pass": t, "host": host_facts()}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs one GPU: every box here has one, and the engine's multi-GPU split (whole "
              f"supernode chunks cyclic over devices, one 32-byte reduction) is designed but not built; "
              f"--gpus {args.gpus} is not supported.", file=sys.stderr)
        sys.exit(2)
    import gsofatpds_amd as g

    batch_config = tuple(int(x) for x in args.batch_config.split(",")) if args.batch_config else None
    if args.batch_only:
        print(json.dumps({"batched": batched_vs_plain(g, args.batch_members, batch_config, max(args.steps, 1))}),
              flush=True)
        return

    wl = {"grid2d_659033": g.WL_GRID2D_659033, "grid3d_110592": g.WL_GRID3D_110592,
          "grid2d_4096": g.WL_GRID2D_4096}[args.workload]
    t0 = time.perf_counter()
    p = g.generate_workload(wl)
    log(f"[bench] generated {args.workload}: n={p.n} nnz={p.nnz} in {time.perf_counter() - t0:.1f}s")
    info = g.device_query(0)
    log(f"[bench] device {info}")
    structure = not args.no_structure
    plan = g.Plan(p, chunk_size=args.chunk, structure=int(structure), early_top=1 if args.early_top else 0)
    log("[bench] plan created")
    for i in range(args.warmup):
        plan.run().sync()
        log(f"[bench] warmup {i}: {plan.result().time_ms:.3f} ms")
    g.device_synchronize()
    # the timed steps: a run and its completion with the totals read back, nothing else (no
    # stage events between the kernels)
    plan.stage_timing(False)
    t0 = time.perf_counter()
    for i in range(args.steps):
        plan.run().sync()
    g.device_synchronize()
    wall = time.perf_counter() - t0
    r = plan.result()
    plan.stage_timing(True)
    # the per-stage breakdown (HIP events on the run's stream) from as many further runs
    stage_acc = {"prep_ms": 0.0, "light_ms": 0.0, "heavy_ms": 0.0, "supernodes_ms": 0.0}
    event_ms = early_ms = 0.0
    for i in range(args.steps):
        plan.run().sync()
        st = plan.stage_times()
        for k in stage_acc:
            stage_acc[k] += st[k]
        event_ms += plan.result().time_ms
        early_ms += plan.early_top()["ms"]
    g.device_synchronize()
    xw = plan.executed_work()
    et = plan.early_top()
    et["ms"] = early_ms / args.steps
    ms_per_step = wall * 1e3 / args.steps
    stages = {k: v / args.steps for k, v in stage_acc.items()}
    log(f"[bench] {args.steps} steps: {ms_per_step:.3f} ms/step wall, {event_ms / args.steps:.3f} ms events, "
        f"stages {stages}")

    W, R = r.edge_checks, r.reach
    gteps = W / (ms_per_step / 1e3) / 1e9
    fill_rate = r.fill_count / (ms_per_step / 1e3)
    # Per-unit bytes of the reference's loop: 8 B per neighbour inspection (index + label,
    # gSoFa_SOAComparison/symbfact_warp_centric_new_scheduling.cuh:952, 959), 12 B per expanded row
    # (row extent + label write, same file :923, 948), 4 B per F entry written. Priced on the work
    # each kernel executed (both directions, after reuse of reached sets): the schedule-independent
    # W would count inspections the kernels never do.
    # k_heavy: the main launch (its stage); the early launch of the top separator's items runs on
    # its own stream beside the wave path and is priced apart (early_launch below).
    kern = {"k_heavy": (xw["heavy_edges"] - et["edges"], xw["heavy_rows"] - et["rows"], xw["heavy_out"] - et["out"],
                        stages["heavy_ms"]),
            "k_light": (xw["light_edges"], xw["light_rows"], xw["light_out"], stages["light_ms"])}
    dom = max(kern, key=lambda k: kern[k][3])
    e_x, v_x, o_x, k_ms = kern[dom]
    kbytes = 8 * e_x + 12 * v_x + 4 * o_x
    achieved = kbytes / (k_ms / 1e3) / 1e9 if k_ms > 0 else 0.0
    traffic = None
    if os.path.exists(args.traffic):
        with open(args.traffic) as f:
            tr = json.load(f)
        if tr.get("kernel") == dom:
            traffic = tr.get("bytes_per_launch")
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom,
                "kernel_ms": k_ms, "algorithmic_bytes_per_launch": kbytes,
                "units_per_launch": {"edge_inspections": e_x, "rows_expanded": v_x, "entries_written": o_x}}
    x_edges = xw["heavy_edges"] + xw["light_edges"]
    x_rows = xw["heavy_rows"] + xw["light_rows"]
    x_out = xw["heavy_out"] + xw["light_out"]
    step_bytes_x = 8 * x_edges + 12 * x_rows + 4 * x_out
    step_bytes_w = 8 * W + 12 * R + 4 * r.nnz_F
    out = {
        "metric": METRIC, "value": gteps, "unit": "GTEPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"{args.workload}: " + {
            "grid2d_659033": "9-point stencil, 812-wide grid truncated to 659,033 rows, ND order, 10% lower drop, "
                             "1 extra entry per 8 rows in its ND leaf",
            "grid3d_110592": "7-point 48^3 stencil, ND order, 15% upper drop",
            "grid2d_4096": "5-point 64x64, natural order, 20% drop"}[args.workload],
                   "n": p.n, "nnz": p.nnz, "chunk_size": args.chunk, "structure": structure},
        "value_definition": "W-GTEPS: W (schedule-independent edge checks, fixed by the input) / step time; "
                            "an effective rate, since the kernels reuse reached sets and execute fewer "
                            "inspections (executed)",
        "fill_nnz_per_s": fill_rate,
        "nnz_F_per_s": r.nnz_F / (ms_per_step / 1e3),
        "result": {"fill_count": r.fill_count, "nnz_F": r.nnz_F, "edge_checks_W": W, "reach_R": R,
                   "supernodes": r.supernodes, "heavy_sources": r.heavy_sources},
        "stages_ms": stages,
        "stages_note": "HIP events on the run's stream, averaged over as many runs after the timed ones "
                       "(the timed loop reads nothing but each run's totals)",
        "executed_work": xw,
        "executed": {"inspections_per_step": x_edges, "rows_expanded_per_step": x_rows,
                     "gteps": x_edges / (ms_per_step / 1e3) / 1e9,
                     "fraction_of_W": x_edges / max(W, 1)},
        "step_roofline": {"achieved": step_bytes_x / (ms_per_step / 1e3) / 1e9, "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": step_bytes_x / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS,
                          "bytes_per_step": step_bytes_x, "what": "executed work of both kernels over the step"},
        "schedule_independent_GBs": step_bytes_w / (ms_per_step / 1e3) / 1e9,
        "roofline": roofline,
    }
    if et["items"] > 0:
        eb = 8 * et["edges"] + 12 * et["rows"] + 4 * et["out"]
        out["early_launch"] = {
            "what": "k_heavy over the top separator's items, on its own stream from the start of the wave path "
                    "(joined before supernode detection; its time overlaps the wave path and the main launch)",
            "items": et["items"], "ms": et["ms"], "algorithmic_bytes_per_launch": eb,
            "achieved_GBs": eb / (et["ms"] / 1e3) / 1e9 if et["ms"] > 0 else 0.0,
            "units_per_launch": {"edge_inspections": et["edges"], "rows_expanded": et["rows"],
                                 "entries_written": et["out"]}}
    plan.close()
    if not args.no_reference_dataset:
        out["reference_dataset"] = reference_dataset(g)
    if args.batch_members > 0:
        out["batched"] = batched_vs_plain(g, args.batch_members, batch_config)
    if args.served_requests > 0:
        out["served"] = served_vs_plain(g, args.served_requests, args.served_rate, args.slots, args.cu_share)
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(p, args.cpu_sample, args.cpu_full)
        log(f"[bench] cpu baseline {out['cpu_baseline']['value']:.3f} GTEPS ({out['cpu_baseline']['sample']})")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
