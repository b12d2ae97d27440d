"""Faithfulness of the CPU baseline: the oracle's count (the C restatement of
the reference CPU path that bench.py times as `cpu_baseline`) against the
reference's own CPU binary built from its sources (oracle/_ref/bbtc_cpu,
`make -C oracle ref`), on the same .bbtc file, cuts and cores.

Test infrastructure, run in the container that holds the reference sources:

    python oracle/check_cpu_baseline.py [--repeat 3]

For each workload it prints both count times (the reference's "Triangle
Counting Execution Time", the oracle's wall time around orc_count over the
task list in the reference CPU build's ascending order), their ratio and the
two triangle counts, which must be equal.
"""
import argparse
import os
import re
import statistics
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

import bbtc_amd as B  # noqa: E402  (graph generation and .bbtc writing)
import oracle as O  # noqa: E402

REF = os.path.join(HERE, "_ref", "bbtc_cpu")

WORKLOADS = [
    ("rmat-s16-ef8 uni ncut4", lambda: B.Graph.rmat(16, 8, 0.57, 0.19, 0.19, 1), "uni", 4),
    ("gnm-n22-m26 pal ncut16", lambda: B.Graph.gnm(1 << 22, 1 << 26, 2), "pal", 16),
]


def ref_run(path, part, ncut, repeat):
    # the CPU build takes no --ngpu (it has only CPU workers)
    out = subprocess.run([REF, "--graph", path, "--partitioner", part, "--ncut", str(ncut),
                          "--repeat", str(repeat)], capture_output=True, text=True, timeout=3600, check=True).stdout
    times = [float(x) for x in re.findall(r"Triangle Counting Execution Time \(s\)\s*:\s*([0-9.eE+-]+)", out)]
    tri = int(re.search(r"Number of Triangles:\s*(\d+)", out).group(1))
    return times, tri


def oracle_run(g, part, ncut, repeat, threads):
    cuts = g.partition(part, ncut)
    R = O.RefLayout(g.offsets(), g.cols(), cuts)
    tasks = R.tasks("asc")
    times, tot = [], None
    for _ in range(repeat):
        t0 = time.perf_counter()
        _, tot = R.count(tasks, threads)
        times.append(time.perf_counter() - t0)
    return times, int(tot)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--repeat", type=int, default=3)
    args = ap.parse_args()
    if not os.path.exists(REF):
        sys.exit(f"{REF} is missing: build it with `make -C oracle ref` where the reference sources exist")
    threads = len(os.sched_getaffinity(0))
    print(f"host threads: {threads}")
    with tempfile.TemporaryDirectory() as d:
        for name, make, part, ncut in WORKLOADS:
            g = make()
            path = os.path.join(d, "g.bbtc")
            g.save(path)
            rt, rtri = ref_run(path, part, ncut, args.repeat)
            ot, otri = oracle_run(g, part, ncut, args.repeat, threads)
            r, o = statistics.median(rt), statistics.median(ot)
            print(f"{name}: reference {r:.4f} s, oracle {o:.4f} s (ratio {o / r:.3f}); "
                  f"triangles {rtri} / {otri} {'equal' if rtri == otri else 'DIFFERENT'}", flush=True)
            if rtri != otri:
                sys.exit(1)


if __name__ == "__main__":
    main()
