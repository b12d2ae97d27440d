"""Write golden vectors from the reference's own code (run once, offline).

Each fixture pins one graph: it is generated here with this repository's
generator, written as .bbtc, and handed to oracle/_ref/ref_harness, which is
the reference's CPU code built from its sources by oracle/Makefile. The
harness prints the cut vector (sparsify + partitioner + cuts[p] = N+1), the
task order of tasks::SortTasks (CPU build, ascending) and the count of every
(i,j,k) task through BBTc. Only those numbers are committed.

Usage: python tests/golden/make_golden.py [--only NAME] [--with-s26]
Needs /root/reference (for `make -C oracle ref`) and the built libbbtc.so.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)

import bbtc_amd as B  # noqa: E402
from tests.golden.fingerprint import fingerprint  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")

RMAT = dict(a=0.57, b=0.19, c=0.19)

# (name, graph spec, [(partitioner, ncut, cuts_only)])
FIXTURES = [
    ("c1_rmat_s16_ef8", dict(kind="rmat", scale=16, ef=8, seed=1, **RMAT),
     [("uni", 4, False), ("n1d", 4, False), ("pal", 4, False), ("rac", 4, False)]),
    ("c2_gnm_n22_m26", dict(kind="gnm", n=1 << 22, m=1 << 26, seed=2),
     [("pal", 16, False)]),
    ("small_rmat_s8", dict(kind="rmat", scale=8, ef=8, seed=11, **RMAT),
     [("uni", 2, False), ("n1d", 3, False), ("pal", 5, False), ("pal", 8, False), ("rac", 3, False)]),
    ("small_rmat_s10", dict(kind="rmat", scale=10, ef=16, seed=12, **RMAT),
     [("uni", 3, False), ("n1d", 4, False), ("pal", 4, False), ("pal", 8, False), ("rac", 8, False)]),
    ("small_rmat_s12", dict(kind="rmat", scale=12, ef=24, seed=13, **RMAT),
     [("uni", 8, False), ("n1d", 8, False), ("pal", 2, False), ("pal", 8, False), ("rac", 5, False)]),
    ("small_gnm_2k", dict(kind="gnm", n=2000, m=16000, seed=14),
     [("uni", 5, False), ("n1d", 2, False), ("pal", 6, False), ("rac", 6, False)]),
    ("small_gnm_dense", dict(kind="gnm", n=600, m=60000, seed=15),
     [("uni", 4, False), ("n1d", 4, False), ("pal", 4, False), ("rac", 4, False)]),
    ("batch_rmat_s11", dict(kind="rmat", scale=11, ef=20, seed=1003, **RMAT),
     [("pal", 2, False), ("pal", 4, False), ("pal", 8, False)]),
    ("batch_gnm_4k", dict(kind="gnm", n=4096, m=4096 * 20, seed=1007),
     [("pal", 2, False), ("pal", 4, False), ("pal", 8, False)]),
    ("rmat_s18_ef16", dict(kind="rmat", scale=18, ef=16, seed=7, **RMAT),
     [("pal", 16, False), ("n1d", 16, True), ("uni", 16, True), ("rac", 16, True)]),
]
# The headline graph: its per-task counts come from the reference's one-task
# BBTc objects run concurrently (ref_harness tasks-parallel), about an hour of
# the reference's CPU code on 8 cores and 30 GB of memory.
S26 = ("c3_rmat_s26_ef16", dict(kind="rmat", scale=26, ef=16, seed=3, **RMAT),
       [("pal", 32, "tasks-parallel")])


def make_graph(spec):
    if spec["kind"] == "rmat":
        return B.Graph.rmat(spec["scale"], spec["ef"], spec["a"], spec["b"], spec["c"], spec["seed"])
    if spec["kind"] == "gnm":
        return B.Graph.gnm(spec["n"], spec["m"], spec["seed"])
    raise ValueError(spec)


def run_harness(path, part, p, mode):
    """mode: False (cuts, task order, per-task counts and the whole-list
    total), True ("cuts-only") or "tasks-parallel"."""
    flag = {False: [], True: ["cuts-only"]}.get(mode, [mode])
    out = subprocess.run([HARNESS, path, part, str(p)] + flag, check=True, capture_output=True, text=True).stdout
    return json.loads(out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default=None)
    ap.add_argument("--with-s26", action="store_true")
    args = ap.parse_args()
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    fixtures = FIXTURES + ([S26] if args.with_s26 else [])
    for name, spec, runs in fixtures:
        if args.only and args.only != name:
            continue
        g = make_graph(spec)
        fp = fingerprint(g.offsets(), g.cols())
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, name + ".bbtc")
            g.save(path)
            del g
            out_runs = []
            for part, p, mode in runs:
                r = run_harness(path, part, p, mode)
                rec = {"partitioner": part, "ncut": p, "cuts": r["cuts"]}
                if mode is not True:
                    rec["order_asc"] = r["order_asc"]
                    rec["tasks"] = r["tasks"]
                    rec["total"] = r["total"]
                    if mode is False:
                        assert r["total"] == r["total_all"], (name, part, p)
                out_runs.append(rec)
                print(name, part, p, r["cuts"], r.get("total"), flush=True)
        rec = {"name": name, "graph": spec, "fingerprint": fp,
               "source": "oracle/_ref/ref_harness (reference CPU code built from its sources)",
               "runs": out_runs}
        if name == S26[0]:
            rec["full_size"] = True
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(rec, f, separators=(",", ":"))


if __name__ == "__main__":
    main()
