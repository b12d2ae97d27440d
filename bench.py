"""Benchmark: block-based triangle count of RMAT scale 26 on one MI355X.

Workload (the headline configuration): RMAT scale 26, edge factor 16,
quadrants a=.57 b=.19 c=.19 d=.05, seed 3, symmetrised without loops or
duplicates, degree-ordered upper triangle; pal partitioner, 32 cuts
(5,984 tasks); the whole block CSR resident in HBM before timing
(--copy-first scope).

A step is one count of every task. The timed region is exactly --steps
launches bracketed by device synchronisation; per-launch HIP events on the
launch stream give the kernel time behind the roofline figure.

Prints one JSON line:
  value         edges/s = m / (time per step), m = upper-triangle entries
  roofline      algorithmic bytes per step / mean event time vs 8 TB/s, and
                vs the measured streaming-read bandwidth (achievable); HBM
                traffic from the committed rocprofv3 counters; the
                compulsory bytes (each task's three tiles read once)
                (forms: --kernel, default auto, the measured-cost rule
                without per-graph autotuning)
  end_to_end_s  generate, partition, block layout built on the device from
                the CSR (with the task work statistics), autotuning
                (--kernel tuned only), median count; beside it the host
                layout build plus upload it replaces
  cpu_baseline  (taken after every GPU workload) the CPU oracle (restated reference CPU path) on a
                systematic 1/s sample of every task's rows, scaled by s
                (--cpu-full: every task once, measured)
  c2            G(n=2^22, m=2^26) seed 2, pal 16 (816 tasks), AUTO forms
                (the row form: one count_rows launch over whole rows):
                edges/s, roofline fraction by the tiled algorithmic bytes
                and by the row form's own bytes, HBM traffic from the
                committed counters, and the oracle over all of its tasks
  batched       4,096 uneven graphs (70% RMAT s10-18, 30% G(n,m)) resident,
                one captured persistent launch vs the plain per-graph launch,
                both with the AUTO forms (and the batch with autotuned forms
                beside it): graphs/s, per-graph p50/p99, bit-exact per-graph
                counts, and the batch's algorithmic bytes / launch time vs
                8 TB/s
  serving       requests (50% RMAT s12-24 ef16, 50% G(n,m) n 2^12-2^22 d 32)
                through the continuous-batching server vs the plain
                sequential launch: capacity, and p50/p99 latency at Poisson
                arrivals of 80% of the plain capacity; bit-exact counts
  shards        the RMAT s26 task list split for 2, 4 and 8 GPUs (LPT on the
                AUTO rule's cost estimates, and the reference's task q to
                GPU q mod G): each share counted alone on this GPU with its
                per-task counts checked against the whole run; the slowest
                share, the imbalance, and the tile bytes a GPU would hold
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "count_probe_rmat-s26-ef16_pal32_traffic.json")


def measured_traffic(workload: str, forms: str):
    """HBM bytes per count from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    runs of this workload and kernel-form choice (scaled by the copy-kernel
    calibration), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None, None
    if t.get("workload") != workload or t.get("forms") != forms:
        return None, None
    return int(t["traffic_bytes"]), os.path.relpath(TRAFFIC_FILE, ROOT)


C2_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "count_rows_gnm-n22-m26_pal16_traffic.json")


def measured_c2_traffic(kernel: str):
    """HBM bytes per C2 count from the committed rocprofv3 counters of this
    kernel (scaled by the copy-kernel calibration), or None."""
    try:
        with open(C2_TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None, None
    if t.get("kernel") != kernel:
        return None, None
    return int(t["traffic_bytes"]), os.path.relpath(C2_TRAFFIC_FILE, ROOT)


CPU_FULL_FILE = os.path.join(ROOT, "profiles", "cpu_baseline_rmat-s26-ef16_pal32.json")


def measured_cpu_full(m: int):
    """The committed measurement of the oracle over every task of this graph
    (m edges; bench.py --cpu-full), or None."""
    try:
        with open(CPU_FULL_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("m") != m:
        return None
    return {"seconds_per_count": t["seconds_per_count"], "value": m / t["seconds_per_count"],
            "cores": t["cores"], "source": os.path.relpath(CPU_FULL_FILE, ROOT)}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def compulsory_bytes(L, tasks, offset_bytes):
    """B_min: each tile of each counted task read once, as the reference's
    device tile holds it ((h+1) row offsets and 4-byte column ids); tasks with
    an empty tile are skipped as in bb_tc.cpp:86-91."""
    P = len(L.cuts) - 1
    tile_bytes = np.zeros((P, P), dtype=np.int64)
    empty = np.zeros((P, P), dtype=bool)
    for i in range(P):
        for k in range(i, P):
            t = L.tile(i, k)
            tile_bytes[i, k] = (t.height + 1) * offset_bytes + 4 * t.nnz
            empty[i, k] = t.nnz == 0
    ti, tj, tk = (tasks[c].astype(np.int64) for c in ("i", "j", "k"))
    keep = ~(empty[ti, tj] | empty[tj, tk] | empty[ti, tk])
    return int((tile_bytes[ti, tj] + tile_bytes[tj, tk] + tile_bytes[ti, tk])[keep].sum())


def host_info(nthreads):
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return {"nproc": usable, "omp_threads": nthreads, "cpu_model": model}


def cpu_baseline(g, cuts, tasks, per_task_gpu, m, budget_s, full=False):
    """Time the oracle (the reference CPU path restated) on the box's host
    cores. full: every task once, measured. Otherwise a systematic sample of
    every task's rows, the 64-row blocks b with b % s == 0 of band i (1/s of
    each task's rows, s sized from a 1/64 calibration run so the sample takes
    about budget_s), its time scaled by s; the committed measurement over
    every task (--cpu-full on a box like this one) is reported beside it.
    Either way, the whole-task counts of every 97th task (lexicographic) are
    checked bit-exact against the GPU."""
    import oracle as O

    t0 = time.time()
    ref = O.RefLayout(g.offsets(), g.cols(), cuts)
    log(f"cpu baseline: reference layout built in {time.time() - t0:.1f}s")
    lex = np.lexsort((tasks["k"], tasks["j"], tasks["i"]))
    nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    tt = np.stack([tasks["i"][lex], tasks["j"][lex], tasks["k"][lex]], 1).astype(np.uint32)
    chk = np.arange(0, len(lex), 97)
    per, _ = ref.count(tt[chk], nthreads)
    if not np.array_equal(per, per_task_gpu[lex[chk]]):
        raise RuntimeError("GPU per-task counts differ from the CPU oracle")
    if full:
        t0 = time.time()
        per, tot = ref.count(tt, nthreads)
        dt = time.time() - t0
        if not np.array_equal(per, per_task_gpu[lex]):
            raise RuntimeError("GPU per-task counts differ from the CPU oracle")
        est, how = dt, (f"oracle (C restatement of the reference CPU path) on all {len(tasks)} tasks once, "
                        f"measured: {dt:.1f}s; every task's count bit-exact vs GPU")
    else:
        t0 = time.time()
        ref.count_rows(tt, 64, 0, nthreads)
        cal = time.time() - t0
        stride = int(min(64, max(1, np.ceil(cal * 64 / budget_s))))
        t0 = time.time()
        ref.count_rows(tt, stride, 0, nthreads)
        dt = time.time() - t0
        est = dt * stride
        how = (f"oracle (C restatement of the reference CPU path) on every one of the {len(tasks)} tasks, the "
               f"64-row blocks b % {stride} == 0 of each (1/{stride} of the rows, a systematic sample): {dt:.1f}s, "
               f"scaled by {stride} to {est:.1f}s per count; whole-task counts of every 97th task bit-exact vs GPU")
    out = {
        "value": m / est,
        "unit": "edges/s",
        "cores": nthreads,
        "kind": "port",
        **host_info(nthreads),
        "seconds_per_count": round(est, 2),
        "sample": how,
    }
    committed = None if full else measured_cpu_full(m)
    if committed:
        out["all_tasks_measured"] = committed
    return out


def count_c2(B, dev, steps, cpu=True, defer_cpu=False):
    """G(n=2^22, m=2^26) seed 2, pal 16, AUTO forms: one resident count, and
    the oracle on every task beside it. With defer_cpu, returns (result, the
    oracle's arguments) for a CPU baseline taken later."""
    t0 = time.time()
    g = B.Graph.gnm_device(dev, 1 << 22, 1 << 26, 2)
    cuts = g.partition("pal", 16)
    L = B.Layout(g, cuts)
    tasks = L.tasks("desc")
    _, total_bytes, _ = L.task_bytes(tasks, 4)
    dg = dev.upload(L)
    plan = dg.plan(tasks, "auto")
    prep = time.time() - t0
    plan.time(1)
    ms = plan.time(steps)
    per, total = plan.results()
    kinds = plan.kinds()
    plan.destroy()
    dg.release()
    mean_s = float(np.mean(ms)) * 1e-3
    log(f"c2: G(n=2^22, m=2^26) pal 16: {total} triangles, events ms {np.round(ms, 3).tolist()} "
        f"(prepared in {prep:.1f}s)")
    rows_form = bool((kinds == 6).any())
    out = {"workload": "gnm-n22-m26 pal ncut16", "metric": "edges/s", "value": g.m / mean_s,
           "ms_per_count": mean_s * 1e3, "triangles": int(total), "tasks": int(len(tasks)),
           "kernel": "count_rows" if rows_form else "count_probe_lanes",
           "algorithmic_bytes": int(total_bytes), "offset_bytes": 4,
           "roofline_frac": total_bytes / mean_s / 1e9 / HBM_PEAK_GBS,
           "forms_merge_bitmap_dense_rdense_lane_rows": np.bincount(kinds, minlength=7)[1:].tolist()}
    if rows_form:
        # what the row form itself reads: each row's length and entries once,
        # and N+(v) once per wedge (u, v) (4-byte ids; the tiled figure above
        # reads bc(v) once per task)
        off = g.offsets().astype(np.int64)
        d = np.diff(off)
        own = 4 * g.n + 4 * g.m + 4 * int(d[g.cols()].sum())
        out["row_form_bytes"] = own
        out["row_form_frac"] = own / mean_s / 1e9 / HBM_PEAK_GBS
    traffic, src = measured_c2_traffic(out["kernel"])
    if traffic:
        out["hbm_traffic_bytes"] = traffic
        out["hbm_frac"] = traffic / mean_s / 1e9 / HBM_PEAK_GBS
        out["traffic_over_algorithmic"] = traffic / total_bytes
        out["traffic_source"] = src
    if defer_cpu:
        return out, ((g, cuts, tasks, per, g.m, 0) if cpu else None)
    if cpu:
        out["cpu_baseline"] = cpu_baseline(g, cuts, tasks, per, g.m, 0, full=True)
        log(f"c2 cpu baseline: {out['cpu_baseline']['seconds_per_count']} s per count")
    return out


def shard_shares(dg, tasks, per, whole_ms, shards=(2, 4, 8), reps=2):
    """Each GPU's part of a G-GPU run, counted alone on this GPU at its real
    shape (all tiles resident, only the share's tasks in the launch): the
    LPT shares of bbtc_tasks_shard against the reference's dealing of task q
    to GPU q mod G over the same ordered list. A G-GPU count takes about the
    slowest share; per-task counts of every share equal the whole run's."""
    L = dg.layout
    tile_bytes = {}

    def tbytes(i, k):
        if (i, k) not in tile_bytes:
            t = L.tile(i, k)
            tile_bytes[(i, k)] = 4 * t.nnz + 8 * t.nz_rows + 4
        return tile_bytes[(i, k)]

    tl = tasks.tolist()
    whole_tiles = {x for i, j, k in tl for x in ((i, j), (j, k), (i, k))}
    res = {"whole_ms": whole_ms, "whole_tile_bytes": sum(tbytes(*x) for x in whole_tiles)}
    for G in shards:
        lpt, _ = dg.shard(tasks, G)
        row = {}
        for name, assign in (("lpt", lpt), ("mod_g", np.arange(len(tasks)) % G)):
            ms, foot = [], []
            for s in range(G):
                sel = assign == s
                p = dg.plan(tasks[sel], "auto")
                p.time(1)
                t = p.time(reps)
                pp, _ = p.results()
                p.destroy()
                if not np.array_equal(pp, per[sel]):
                    raise RuntimeError(f"share {s} of {G} ({name}) differs from the whole run per task")
                ms.append(float(np.median(t)))
                tiles = {x for (i, j, k), y in zip(tl, sel) if y for x in ((i, j), (j, k), (i, k))}
                foot.append(sum(tbytes(*x) for x in tiles))
            row[name] = {"share_ms": [round(x, 2) for x in ms], "max_ms": max(ms),
                         "imbalance": max(ms) / float(np.mean(ms)), "speedup": whole_ms / max(ms),
                         "max_tile_bytes": max(foot)}
        res[f"g{G}"] = row
        log(f"shards of {G}: LPT slowest share {row['lpt']['max_ms']:.1f} ms (imbalance "
            f"{row['lpt']['imbalance']:.3f}, {row['lpt']['speedup']:.2f}x), task q to GPU q mod {G}: "
            f"{row['mod_g']['max_ms']:.1f} ms ({row['mod_g']['speedup']:.2f}x)")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=26)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--ncut", type=int, default=32)
    ap.add_argument("--partitioner", default="pal")
    ap.add_argument("--kernel", default="auto", help="auto | tuned | merge | bitmap | dense | rdense | lane")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-full", action="store_true",
                    help="CPU baseline over every task once (about 2.5 minutes on 16 host threads)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c2", action="store_true", help="skip the G(n=2^22, m=2^26) count")
    ap.add_argument("--c2-only", action="store_true",
                    help="only the G(n=2^22, m=2^26) count (--steps launches); prints its object as the line")
    ap.add_argument("--batched-only", action="store_true",
                    help="only the batched workload (--batched-graphs graphs); prints its object as the line")
    ap.add_argument("--batched-graphs", type=int, default=4096, help="0: skip the batched workload")
    ap.add_argument("--no-batch-tune", action="store_true",
                    help="batched workload with the AUTO forms only (no autotuned batch beside it)")
    ap.add_argument("--serving-requests", type=int, default=64, help="0: skip the serving workload")
    ap.add_argument("--serving-only", action="store_true",
                    help="only the serving workload (--serving-requests requests); prints its object as the line")
    ap.add_argument("--no-shards", action="store_true",
                    help="skip the per-GPU shares of 2, 4 and 8 GPUs counted on this GPU")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: it counts one graph on one MI355X; "
              "the multi-GPU task sharding is a design, not a measured path.", file=sys.stderr)
        sys.exit(2)

    import bbtc_amd as B

    offset_bytes = 8 if args.scale >= 26 else 4
    dev = B.Device(0)
    if args.c2_only:
        print(json.dumps(count_c2(B, dev, steps=args.steps, cpu=not args.no_cpu)), flush=True)
        return
    if args.batched_only:
        from bbtc_amd import workloads as W
        print(json.dumps(W.run_batched(dev, W.batched_specs(args.batched_graphs), log=log,
                                       tune=not args.no_batch_tune)), flush=True)
        return
    if args.serving_only:
        from bbtc_amd import workloads as W
        print(json.dumps(W.run_serving(dev, W.serving_specs(args.serving_requests), log=log)), flush=True)
        return
    read_gbs = dev.read_bandwidth(8 << 30, 5)
    stream_gbs = float(np.median(read_gbs))
    log(f"streaming read of 8 GiB: {np.round(read_gbs, 1).tolist()} GB/s (median {stream_gbs:.1f})")
    t_e2e = time.time()
    t0 = time.time()
    g = B.Graph.rmat_device(dev, args.scale, args.edge_factor, 0.57, 0.19, 0.19, args.seed)
    n, _, m = g.info
    log(f"graph: RMAT s{args.scale} ef{args.edge_factor} n={n} m={m} (generated on the GPU in "
        f"{time.time() - t0:.1f}s)")
    gen_s = time.time() - t0
    t0 = time.time()
    cuts = g.partition(args.partitioner, args.ncut)
    part_s = time.time() - t0
    # the block layout built on the device from the CSR, resident from then on
    t0 = time.time()
    dg = dev.build(g, cuts, host_arrays=False)
    build_s = time.time() - t0
    tasks = dg.layout.tasks("desc")
    prep_s = time.time() - t_e2e  # generate, partition, layout
    log(f"partition: {part_s:.2f}s; layout built on the device: {build_s:.2f}s ({dg.bytes / 1e9:.2f} GB "
        f"resident), {len(tasks)} tasks")
    # the host path for comparison (and the host arrays the byte accounting reads)
    t0 = time.time()
    L = B.Layout(g, cuts)
    host_layout_s = time.time() - t0
    t0 = time.time()
    up = dev.upload(L)
    upload_s = time.time() - t0
    if up.bytes != dg.bytes or not np.array_equal(L.tasks("desc"), tasks):
        raise RuntimeError("the device-built layout differs from the host-built one")
    up.release()
    del up
    log(f"host path: layout {host_layout_s:.2f}s + upload {upload_s:.2f}s")
    t0 = time.time()
    bytes_per_task, total_bytes, wedges = L.task_bytes(tasks, offset_bytes)
    log(f"algorithmic bytes: {total_bytes / 1e9:.1f} GB (o={offset_bytes}), wedges {wedges} ({time.time() - t0:.1f}s)")
    min_bytes = compulsory_bytes(L, tasks, offset_bytes)
    log(f"compulsory bytes (each task's three tiles once): {min_bytes / 1e9:.1f} GB")
    tune_s = 0.0
    if args.kernel == "tuned":
        t0 = time.time()
        choice, _ = dg.autotune(tasks)
        tune_s = time.time() - t0
        plan = dg.plan(tasks, "auto", choice)
        log(f"autotune: {tune_s:.1f}s, forms merge/bitmap/dense/rdense/lane = "
            f"{np.bincount(choice & 7, minlength=6)[1:].tolist()}, rows halved 0/1/2 times = "
            f"{np.bincount(choice >> 4, minlength=3).tolist()}")
    else:
        plan = dg.plan(tasks, args.kernel)
    if args.warmup:
        plan.time(args.warmup)
    dev.synchronize()
    t0 = time.perf_counter()
    ms = plan.time(args.steps)  # launches, event pairs, stream sync
    dev.synchronize()
    wall = time.perf_counter() - t0
    per, total = plan.results()
    forms = np.bincount(plan.kinds(), minlength=7)[1:].tolist()
    ms_per_step = wall * 1e3 / args.steps
    ev_mean_s = float(np.mean(ms)) * 1e-3
    ev_median_ms = float(np.median(ms))
    achieved = total_bytes / ev_mean_s / 1e9
    log(f"count: {total} triangles; events ms {np.round(ms, 3).tolist()}; wall/step {ms_per_step:.3f} ms; "
        f"bitmap tasks {plan.bitmap_tasks}")
    if int(per.sum()) != total:
        raise RuntimeError("per-task counts do not sum to the total")

    # the CPU baselines run after every GPU workload: their host threads do
    # not share the host with the timed GPU paths' own host threads
    cpu = None

    workload = f"rmat-s{args.scale}-ef{args.edge_factor} {args.partitioner} ncut{args.ncut}"
    traffic, traffic_src = measured_traffic(workload, args.kernel)
    out = {
        "metric": "edges/s (RMAT s26 block-based triangle count)",
        "value": m / (ms_per_step * 1e-3),
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": f"synthetic (seeded RMAT generator, seed {args.seed})",
        "config": {
            "workload": workload,
            "n": n, "m": m, "ncut": args.ncut, "tasks": int(len(tasks)),
            "partitioner": args.partitioner, "kernel": args.kernel,
            "triangles": int(total), "autotune_s": round(tune_s, 2),
            "forms_merge_bitmap_dense_rdense_lane_rows": forms,
        },
        "end_to_end_s": {
            "generate": round(gen_s, 3), "partition": round(part_s, 3),
            "layout_on_device": round(build_s, 3),
            "autotune": round(tune_s, 3), "count_median": round(ev_median_ms * 1e-3, 4),
            "total": round(prep_s + tune_s + ev_median_ms * 1e-3, 3),
            # the same layout built on the host and uploaded, in place of layout_on_device
            "host_layout_and_upload": round(host_layout_s + upload_s, 3),
        },
        "triangles_per_s": total / (ms_per_step * 1e-3),
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
            "algorithmic_bytes": total_bytes, "offset_bytes": offset_bytes,
            "kernel": "count_probe", "event_ms_mean": ev_mean_s * 1e3, "event_ms_median": ev_median_ms,
            "achievable": stream_gbs, "frac_achievable": achieved / stream_gbs,
            "compulsory_bytes": min_bytes,
            # the HBM bytes rocprofv3 counted for this workload (traffic), per
            # second of this run's count, against the spec peak
            "hbm_frac": traffic / ev_mean_s / 1e9 / HBM_PEAK_GBS if traffic else None,
            "traffic_over_compulsory": traffic / min_bytes if traffic else None,
            "traffic_over_algorithmic": traffic / total_bytes if traffic else None,
        },
        "cpu_baseline": cpu,
    }
    plan.destroy()
    if not args.no_shards:
        out["shards"] = shard_shares(dg, tasks, per, ev_median_ms)
    dg.release()
    del plan, dg, L
    s26_cpu_args = None if args.no_cpu else (g, cuts, tasks, per, m, args.cpu_seconds)

    from bbtc_amd import workloads as W

    def rounded(d):
        return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in d.items()}

    c2_cpu_args = None
    if not args.no_c2:
        c2, c2_cpu_args = count_c2(B, dev, steps=3, cpu=not args.no_cpu, defer_cpu=True)
        out["c2"] = rounded(c2)

    if args.batched_graphs:
        res = W.run_batched(dev, W.batched_specs(args.batched_graphs), log=log, tune=not args.no_batch_tune)
        out["batched"] = rounded(dict(res, metric="graphs/s", unit="graphs/s",
                                      value=res["batched_graphs_per_s"], vs_plain=res["speedup"]))
    if args.serving_requests:
        res = W.run_serving(dev, W.serving_specs(args.serving_requests), log=log)
        out["serving"] = rounded(dict(res, metric="graphs/s at capacity", unit="graphs/s",
                                      value=res["server_capacity_gps"], vs_plain=res["capacity_ratio"]))
    if s26_cpu_args:
        out["cpu_baseline"] = cpu_baseline(*s26_cpu_args, full=args.cpu_full)
    if c2_cpu_args:
        out["c2"]["cpu_baseline"] = cpu_baseline(*c2_cpu_args, full=True)
        log(f"c2 cpu baseline: {out['c2']['cpu_baseline']['seconds_per_count']} s per count")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
