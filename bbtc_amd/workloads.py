"""Seeded multi-graph workloads and the batched / plain-launch runners.

Batched workload (4,096 independent graphs, graph g seeded 1000 + g):
70% RMAT (scale uniform in [10, 18], edge factor uniform in [8, 32],
a=.57 b=.19 c=.19), 30% G(n, m) (n uniform in [2^10, 2^18], average degree d
uniform in [4, 64], m = round(n d / 2)). Each graph's parameters come from a
SplitMix64 draw of its seed, so the list is the same on every machine.
"""
from __future__ import annotations

import time

import numpy as np

from . import Batch, Device, Graph, layouts_build, rule_ncut, tune_ncut

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


def run_batched(dev: Device, specs, log=print, sample_every: int = 64, cands=(1, 2, 4, 8)):
    """Counts every graph of `specs` with the plain per-graph launch and with
    one batched persistent launch; returns a dict of measurements and checks
    the per-graph counts agree bit for bit."""
    t0 = time.time()
    graphs = [make_graph(dev, s) for s in specs]
    ms = [g.m for g in graphs]
    log(f"batched: {len(graphs)} graphs generated on the GPU in {time.time() - t0:.1f}s, "
        f"{sum(ms):,} edges in total")

    # cut counts: rule tuned on a sample spread over the size order
    t0 = time.time()
    order = np.argsort(ms)
    sample = [graphs[i] for i in order[:: max(1, len(graphs) // max(1, len(graphs) // sample_every))]]
    rule, _ = tune_ncut(dev, sample, cands)
    ncut = [rule_ncut(rule, m) for m in ms]
    tune_s = time.time() - t0
    log(f"batched: cut rule from {len(sample)} sample graphs in {tune_s:.1f}s; "
        f"ncut histogram {dict(zip(*np.unique(ncut, return_counts=True)))}")

    t0 = time.time()
    layouts = layouts_build(graphs, ncut, "pal")
    log(f"batched: {len(layouts)} layouts in {time.time() - t0:.1f}s")

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
    plain_lat = np.zeros(len(plans))
    t0 = time.perf_counter()
    for x, pl in enumerate(plans):
        s = time.perf_counter()
        pl.launch()
        plain_counts[x] = pl.results()[1]
        plain_lat[x] = (time.perf_counter() - s) * 1e3
    plain_s = time.perf_counter() - t0
    for pl in plans:
        pl.destroy()
    for dg in dgs:
        dg.release()

    # batched: one arena, forms autotuned over the whole batch, one captured
    # graph of launches
    t0 = time.time()
    b = Batch(dev, layouts, "auto")
    b.autotune()
    b.capture()
    setup_s = time.time() - t0
    b.launch()
    b.results()
    dev.synchronize()
    t0 = time.perf_counter()
    b.launch()
    counts, _, lat = b.results()
    batched_s = time.perf_counter() - t0
    if not np.array_equal(counts, plain_counts):
        bad = int(np.sum(counts != plain_counts))
        raise RuntimeError(f"batched counts differ from the plain launch on {bad} graphs")
    res = dict(
        graphs=len(graphs), edges=int(sum(ms)), triangles=int(counts.sum()),
        plain_s=plain_s, plain_graphs_per_s=len(graphs) / plain_s,
        plain_p50_ms=percentile(plain_lat, 50), plain_p99_ms=percentile(plain_lat, 99),
        batched_s=batched_s, batched_graphs_per_s=len(graphs) / batched_s,
        batched_p50_ms=percentile(lat, 50), batched_p99_ms=percentile(lat, 99),
        speedup=plain_s / batched_s, ncut_tune_s=tune_s, batch_setup_s=setup_s,
        device_bytes=b.bytes, ntasks=b.ntasks,
    )
    b.destroy()
    log(f"batched: {res['batched_graphs_per_s']:.0f} graphs/s vs plain {res['plain_graphs_per_s']:.0f} "
        f"graphs/s ({res['speedup']:.2f}x); p50/p99 {res['batched_p50_ms']:.2f}/{res['batched_p99_ms']:.2f} ms "
        f"(plain {res['plain_p50_ms']:.3f}/{res['plain_p99_ms']:.3f} ms)")
    return res
