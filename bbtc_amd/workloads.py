"""Seeded multi-graph workloads and their runners: batched, serving, and the
plain per-graph launch each is compared with.

Batched workload (4,096 independent graphs, graph g seeded 1000 + g):
70% RMAT (scale uniform in [10, 18], edge factor uniform in [8, 32],
a=.57 b=.19 c=.19), 30% G(n, m) (n uniform in [2^10, 2^18], average degree d
uniform in [4, 64], m = round(n d / 2)).

Serving workload (request r seeded 7_000_000 + r): 50% RMAT (scale uniform in
[12, 24], edge factor 16), 50% G(n, m) (n uniform in [2^12, 2^22], average
degree 32, m = 16 n); Poisson arrivals (exponential gaps from a generator
seeded 7) at a fixed rate.

Each graph's parameters come from a SplitMix64 draw of its seed, so the lists
are the same on every machine.
"""
from __future__ import annotations

import time

import numpy as np

from . import Batch, Device, Graph, Layout, Server, layouts_build, rule_ncut, tune_ncut

M64 = (1 << 64) - 1


def _mix(z: int) -> int:
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def batched_specs(count: int = 4096, seed0: int = 1000):
    out = []
    for gi in range(count):
        seed = seed0 + gi
        h1, h2 = _mix(seed), _mix(seed ^ 0xA5A5A5A5)
        if h1 % 10 < 7:
            out.append(dict(kind="rmat", scale=10 + (h1 >> 8) % 9, ef=8 + (h1 >> 20) % 25, seed=seed))
        else:
            n = (1 << 10) + (h2 >> 8) % ((1 << 18) - (1 << 10) + 1)
            d = 4 + (h2 >> 40) % 61
            out.append(dict(kind="gnm", n=n, m=int(round(n * d / 2)), seed=seed))
    return out


def make_graph(dev: Device, spec) -> Graph:
    if spec["kind"] == "rmat":
        return Graph.rmat_device(dev, spec["scale"], spec["ef"], 0.57, 0.19, 0.19, spec["seed"])
    return Graph.gnm_device(dev, spec["n"], spec["m"], spec["seed"])


def percentile(x, q):
    return float(np.percentile(np.asarray(x, dtype=np.float64), q)) if len(x) else 0.0


def run_batched(dev: Device, specs, log=print, sample_every: int = 64, cands=(1, 2, 4, 8),
                order: str = "small_first", tune: bool = True):
    """Counts every graph of `specs` with the plain per-graph launch and with
    one batched persistent launch; returns a dict of measurements and checks
    the per-graph counts agree bit for bit. Latencies are completion times
    since the whole set was started (both forms), plus the plain launch's
    per-graph service time."""
    t0 = time.time()
    graphs = [make_graph(dev, s) for s in specs]
    ms = [g.m for g in graphs]
    log(f"batched: {len(graphs)} graphs generated on the GPU in {time.time() - t0:.1f}s, "
        f"{sum(ms):,} edges in total")

    # cut counts: rule tuned on a sample spread over the size order
    t0 = time.time()
    by_size = np.argsort(ms)
    sample = [graphs[i] for i in by_size[:: max(1, len(graphs) // max(1, len(graphs) // sample_every))]]
    rule, _ = tune_ncut(dev, sample, cands)
    ncut = [rule_ncut(rule, m) for m in ms]
    tune_s = time.time() - t0
    log(f"batched: cut rule from {len(sample)} sample graphs in {tune_s:.1f}s; "
        f"ncut histogram {dict(zip(*np.unique(ncut, return_counts=True)))}")

    t0 = time.time()
    layouts = layouts_build(graphs, ncut, "pal")
    log(f"batched: {len(layouts)} layouts in {time.time() - t0:.1f}s")
    t0 = time.time()
    alg_bytes = sum(L.task_bytes(L.tasks("desc"), 4)[1] for L in layouts)
    log(f"batched: algorithmic bytes {alg_bytes / 1e9:.1f} GB (o=4) in {time.time() - t0:.1f}s")

    # plain per-graph launch: graphs resident, one plan per graph (auto forms),
    # counted one after another
    t0 = time.time()
    dgs, plans = [], []
    for L in layouts:
        dg = dev.upload(L)
        dgs.append(dg)
        plans.append(dg.plan(L.tasks("desc"), "auto"))
    log(f"plain: {len(plans)} graphs uploaded and planned in {time.time() - t0:.1f}s")
    for pl in plans[:8]:  # warm
        pl.launch()
        pl.results()
    dev.synchronize()
    plain_counts = np.zeros(len(plans), dtype=np.uint64)
    plain_lat = np.zeros(len(plans))   # one graph's launch and read-back
    plain_done = np.zeros(len(plans))  # completion since the first launch
    t0 = time.perf_counter()
    for x, pl in enumerate(plans):
        s = time.perf_counter()
        pl.launch()
        plain_counts[x] = pl.results()[1]
        e = time.perf_counter()
        plain_lat[x] = (e - s) * 1e3
        plain_done[x] = (e - t0) * 1e3
    plain_s = time.perf_counter() - t0
    for pl in plans:
        pl.destroy()
    for dg in dgs:
        dg.release()

    # batched: one arena, one captured graph of launches; first with the
    # AUTO forms (the plain path's rule: like for like), then with forms
    # autotuned over the whole batch
    def timed_batch(b):
        b.capture()
        b.launch()
        b.results()
        dev.synchronize()
        t0 = time.perf_counter()
        b.launch()
        counts, _, lat = b.results()
        el = time.perf_counter() - t0
        if not np.array_equal(counts, plain_counts):
            bad = int(np.sum(counts != plain_counts))
            raise RuntimeError(f"batched counts differ from the plain launch on {bad} graphs")
        return el, counts, lat

    t0 = time.time()
    b = Batch(dev, layouts, "auto", order=order)
    setup_s = time.time() - t0
    batched_s, counts, lat = timed_batch(b)
    batch_tune_s, tuned_s, tuned_lat = 0.0, batched_s, lat
    if tune:
        t0 = time.time()
        b.autotune()
        batch_tune_s = time.time() - t0
        tuned_s, _, tuned_lat = timed_batch(b)
    res = dict(
        graphs=len(graphs), edges=int(sum(ms)), triangles=int(counts.sum()), order=order,
        plain_s=plain_s, plain_graphs_per_s=len(graphs) / plain_s,
        plain_p50_ms=percentile(plain_done, 50), plain_p99_ms=percentile(plain_done, 99),
        plain_service_p50_ms=percentile(plain_lat, 50), plain_service_p99_ms=percentile(plain_lat, 99),
        batched_s=batched_s, batched_graphs_per_s=len(graphs) / batched_s,
        batched_p50_ms=percentile(lat, 50), batched_p99_ms=percentile(lat, 99),
        speedup=plain_s / batched_s, forms="auto (batch and plain alike)",
        tuned_batched_s=tuned_s, tuned_batched_graphs_per_s=len(graphs) / tuned_s,
        tuned_batched_p50_ms=percentile(tuned_lat, 50), tuned_batched_p99_ms=percentile(tuned_lat, 99),
        tuned_speedup=plain_s / tuned_s, batch_autotune_s=batch_tune_s,
        ncut_tune_s=tune_s, batch_setup_s=setup_s,
        algorithmic_bytes=int(alg_bytes), roofline_frac=alg_bytes / batched_s / 8e12,
        device_bytes=b.bytes, ntasks=b.ntasks,
    )
    b.destroy()
    log(f"batched ({order}, auto forms): {res['batched_graphs_per_s']:.0f} graphs/s vs plain "
        f"{res['plain_graphs_per_s']:.0f} graphs/s ({res['speedup']:.2f}x); completion p50/p99 "
        f"{res['batched_p50_ms']:.2f}/{res['batched_p99_ms']:.2f} ms vs plain {res['plain_p50_ms']:.1f}/"
        f"{res['plain_p99_ms']:.1f} ms (plain per-graph service {res['plain_service_p50_ms']:.3f}/"
        f"{res['plain_service_p99_ms']:.3f} ms); with forms autotuned over the batch ({batch_tune_s:.1f}s): "
        f"{res['tuned_batched_graphs_per_s']:.0f} graphs/s ({res['tuned_speedup']:.2f}x)")
    return res


# ---- serving -------------------------------------------------------------------
def serving_specs(count: int, seed0: int = 7_000_000):
    out = []
    for r in range(count):
        seed = seed0 + r
        h1, h2 = _mix(seed), _mix(seed ^ 0x5A5A5A5A)
        if h1 % 2 == 0:
            out.append(dict(kind="rmat", scale=12 + (h1 >> 8) % 13, ef=16, seed=seed))
        else:
            n = (1 << 12) + (h2 >> 8) % ((1 << 22) - (1 << 12) + 1)
            out.append(dict(kind="gnm", n=n, m=16 * n, seed=seed))
    return out


def poisson_arrivals(count: int, rate: float, seed: int = 7):
    gaps = np.random.default_rng(seed).exponential(1.0 / rate, size=count)
    return np.cumsum(gaps) - gaps[0]  # the first request arrives at 0


def prepare_requests(dev: Device, specs, log=print, cands=(1, 2, 4, 8, 16), nsample: int = 8,
                     pin: bool = True):
    """Graphs, per-graph cut counts from a measured rule, layouts and task
    lists (descending order) of every request; with pin, the layouts'
    arrays in page-locked host memory (Device.pin), as the reference's
    BlockCSR holds its tiles."""
    t0 = time.time()
    graphs = [make_graph(dev, s) for s in specs]
    ms = [g.m for g in graphs]
    log(f"serving: {len(graphs)} graphs generated in {time.time() - t0:.1f}s, {sum(ms):,} edges")
    t0 = time.time()
    order = np.argsort(ms)
    pick = np.unique(np.linspace(0, len(graphs) - 1, min(nsample, len(graphs))).round().astype(int))
    rule, _ = tune_ncut(dev, [graphs[order[x]] for x in pick], cands, tune_forms=False, by_time=True)
    ncut = [rule_ncut(rule, m) for m in ms]
    log(f"serving: cut rule from {len(pick)} sample graphs in {time.time() - t0:.1f}s; "
        f"ncut histogram {dict(zip(*np.unique(ncut, return_counts=True)))}")
    t0 = time.time()
    layouts = [None] * len(graphs)
    big = [x for x in range(len(graphs)) if ms[x] > (8 << 20)]
    small = [x for x in range(len(graphs)) if ms[x] <= (8 << 20)]
    for x in big:  # parallel inside
        layouts[x] = Layout(graphs[x], graphs[x].partition("pal", ncut[x]))
    for x, L in zip(small, layouts_build([graphs[x] for x in small], [ncut[x] for x in small], "pal")):
        layouts[x] = L
    tasks = [L.tasks("desc") for L in layouts]
    log(f"serving: layouts in {time.time() - t0:.1f}s, host block CSR "
        f"{sum(L.device_bytes for L in layouts) / 1e9:.1f} GB")
    if pin:
        t0 = time.time()
        for L in layouts:
            dev.pin(L)
        log(f"serving: layouts pinned in {time.time() - t0:.1f}s")
    return graphs, layouts, tasks, ncut


def run_plain(dev: Device, layouts, tasks, arrivals=None, split=None):
    """The plain per-graph launch, one request after another: upload, count,
    release. arrivals (seconds, or None: all at once) delay each request until
    its arrival. Returns (per-request totals, latencies s, makespan s,
    per-task counts); split (a dict, optional) receives the seconds spent in
    uploads ("upload_s") and in counts with their release ("count_s")."""
    n = len(layouts)
    tot = np.zeros(n, dtype=np.uint64)
    lat = np.zeros(n)
    per = [None] * n
    up_s = cnt_s = 0.0
    dev.synchronize()
    t0 = time.perf_counter()
    for x in range(n):
        a = 0.0 if arrivals is None else float(arrivals[x])
        wait = a - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        t1 = time.perf_counter()
        dg = dev.upload(layouts[x])
        t2 = time.perf_counter()
        per[x], tot[x] = dg.count(tasks[x], "auto")
        dg.release()
        t3 = time.perf_counter()
        up_s += t2 - t1
        cnt_s += t3 - t2
        lat[x] = t3 - t0 - a
    if split is not None:
        split.update(upload_s=up_s, count_s=cnt_s)
    return tot, lat, time.perf_counter() - t0, per


def run_server(dev: Device, layouts, tasks, arrivals=None, heap_bytes: int = 200_000_000_000, srv=None, **opts):
    """Every request through one continuous-batching server (`srv`, or a new
    one destroyed afterwards). Returns (totals, latencies s, makespan s,
    per-task counts, server statistics of this run)."""
    n = len(layouts)
    own = srv is None
    if own:
        srv = Server(dev, heap_bytes=heap_bytes, **opts)
    try:
        srv.reset_clock()
        t_sub = time.perf_counter()
        ids = [srv.submit(layouts[x], tasks[x], 0.0 if arrivals is None else float(arrivals[x]))
               for x in range(n)]
        t_sub = time.perf_counter() - t_sub
        res = {r.id: r for r in srv.drain()}
        stats = srv.stats()
        stats["submit_s"] = t_sub  # host time of the submit calls (request set-up)
    finally:
        if own:
            srv.destroy()
    rs = [res[i] for i in ids]
    bad = [r.status for r in rs if r.status != 0]
    if bad:
        raise RuntimeError(f"{len(bad)} requests failed in the server (status {bad[0]})")
    tot = np.array([r.total for r in rs], dtype=np.uint64)
    lat = np.array([r.latency_s for r in rs])
    return tot, lat, max(r.done_s for r in rs), [r.per_task for r in rs], stats


def run_serving(dev: Device, specs, log=print, rate=None, load: float = 0.8, heap_bytes: int = 200_000_000_000,
                pin: bool = True, **server_opts):
    """Serving workload: the plain per-graph launch and the server, first
    saturated (every request at time 0: capacity), then at Poisson arrivals of
    `rate` requests/s (default: `load` times the plain capacity). Per-task
    counts must agree bit for bit between the two. server_opts: Server
    options (launch_us, oldest_pct, part_chunks, ...)."""
    graphs, layouts, tasks, ncut = prepare_requests(dev, specs, log, pin=pin)
    del graphs
    n = len(layouts)
    # warm code paths on the smallest request
    small = int(np.argmin([L.device_bytes for L in layouts]))
    run_plain(dev, [layouts[small]], [tasks[small]])

    split = {}
    p_tot, _, p_make, p_per = run_plain(dev, layouts, tasks, split=split)
    p_cap = n / p_make
    log(f"plain saturated: {p_cap:.2f} graphs/s ({p_make:.2f}s for {n}: uploads {split['upload_s']:.2f}s, "
        f"counts {split['count_s']:.2f}s)")
    # one server (one heap reservation) for both server runs
    srv = Server(dev, heap_bytes=heap_bytes, **server_opts)
    try:
        res = _serving_runs(dev, layouts, tasks, srv, n, p_tot, p_per, p_cap, rate, load, log)
    finally:
        srv.destroy()
    # A server that overlapped every upload with counting, and counted and
    # uploaded each request as fast as the plain path does, would take at
    # least max(U, C) for the mix: at most (U + C) / max(U, C) of the plain
    # capacity, U and C the plain path's upload and count seconds
    u, c = split["upload_s"], split["count_s"]
    res.update(plain_upload_s=u, plain_count_s=c, overlap_bound=(u + c) / max(u, c, 1e-9))
    return res


def _serving_runs(dev, layouts, tasks, srv, n, p_tot, p_per, p_cap, rate, load, log):
    s_tot, _, s_make, s_per, st = run_server(dev, layouts, tasks, srv=srv)
    s_cap = n / s_make
    log(f"server saturated: {s_cap:.2f} graphs/s ({s_make:.2f}s); {st}")
    if not np.array_equal(p_tot, s_tot) or not all(np.array_equal(a, b) for a, b in zip(p_per, s_per)):
        raise RuntimeError("server counts differ from the plain per-graph launch")

    lam = rate if rate else load * p_cap
    arr = poisson_arrivals(n, lam)
    p_tot2, p_lat, p_make2, _ = run_plain(dev, layouts, tasks, arr)
    s_tot2, s_lat, s_make2, _, st2 = run_server(dev, layouts, tasks, arr, srv=srv)
    if not (np.array_equal(p_tot2, p_tot) and np.array_equal(s_tot2, p_tot)):
        raise RuntimeError("counts at the Poisson rate differ from the saturated runs")
    res = dict(
        requests=n, edges=int(sum(L.graph.m for L in layouts)), triangles=int(p_tot.sum()),
        plain_capacity_gps=p_cap, server_capacity_gps=s_cap, capacity_ratio=s_cap / p_cap,
        rate=lam, load_vs_plain=lam / p_cap,
        plain_gps=n / p_make2, server_gps=n / s_make2,
        plain_p50_ms=percentile(p_lat, 50) * 1e3, plain_p99_ms=percentile(p_lat, 99) * 1e3,
        server_p50_ms=percentile(s_lat, 50) * 1e3, server_p99_ms=percentile(s_lat, 99) * 1e3,
        server_launches=st2["launches"], server_peak_active=st2["peak_active"],
        server_peak_resident_gb=st2["peak_resident_bytes"] / 1e9,
        server_busy_s=st["busy_s"], server_upload_s=st["upload_s"], server_submit_s=st["submit_s"],
        server_makespan_s=s_make,
        pinned_layouts=all(L.pinned for L in layouts),
    )
    log(f"serving at {lam:.2f} req/s: server p50/p99 {res['server_p50_ms']:.1f}/{res['server_p99_ms']:.1f} ms, "
        f"plain p50/p99 {res['plain_p50_ms']:.1f}/{res['plain_p99_ms']:.1f} ms; capacity "
        f"{s_cap:.2f} vs {p_cap:.2f} graphs/s ({res['capacity_ratio']:.2f}x)")
    return res
