"""Host code under OpenMP teams smaller than omp_get_max_threads().

OMP_THREAD_LIMIT and OMP_DYNAMIC can hand a parallel region fewer threads
than it asked for; every part of the work must still be done. Each case runs
in a child process with the limit set and must give the same result as the
golden fixtures the reference wrote (text ingest hashes, rac cuts) and as
an unlimited run (generator fingerprints)."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))

LIMITS = [
    {"OMP_NUM_THREADS": "8", "OMP_THREAD_LIMIT": "2"},
    {"OMP_NUM_THREADS": "8", "OMP_DYNAMIC": "true"},
]

CHILD = r"""
import hashlib, json, os, sys, tempfile
sys.path.insert(0, %(root)r)
import bbtc_amd as B
from tests.conftest import graph_from_spec, load_golden
from tests.golden import textgen
from tests.golden.fingerprint import fingerprint
out = {}
g = B.Graph.rmat(14, 8, seed=3)
h = B.Graph.gnm(5000, 40000, 22)
out["gen"] = [fingerprint(g.offsets(), g.cols()), fingerprint(h.offsets(), h.cols())]
gold = json.load(open(os.path.join(%(root)r, "tests", "golden", "ingest.json")))["cases"]
td = tempfile.TemporaryDirectory()
d = td.name
out["ingest"] = {}
for rec in gold:
    src = os.path.join(d, textgen.filename(rec["case"]))
    with open(src, "wb") as f:
        f.write(textgen.make(rec["case"]))
    dst = os.path.join(d, "o.bbtc")
    B.convert_text(src, dst, dedup=False)
    out["ingest"][json.dumps(rec["case"], sort_keys=True)] = hashlib.sha256(open(dst, "rb").read()).hexdigest()
out["rac"] = []
for fx in load_golden():
    if fx["graph"].get("scale", 0) > 14 or fx["graph"].get("n", 0) > 100000:
        continue
    gg = graph_from_spec(fx["graph"])
    for r in fx["runs"]:
        if r["partitioner"] == "rac":
            out["rac"].append([fx["name"], r["ncut"], gg.partition("rac", r["ncut"]).tolist()])
td.cleanup()
print(json.dumps(out))
"""


def run_child(env_extra):
    env = dict(os.environ)
    for k in ("OMP_THREAD_LIMIT", "OMP_DYNAMIC"):
        env.pop(k, None)
    env.update(env_extra)
    p = subprocess.run([sys.executable, "-c", CHILD % {"root": ROOT}], env=env, capture_output=True,
                       text=True, timeout=600)
    assert p.returncode == 0, p.stderr[-2000:]
    return json.loads(p.stdout.strip().splitlines()[-1])


@pytest.fixture(scope="module")
def unlimited():
    return run_child({"OMP_NUM_THREADS": "8"})


@pytest.mark.parametrize("limit", LIMITS, ids=["thread_limit_2", "dynamic"])
def test_smaller_team_same_results(unlimited, limit):
    got = run_child(limit)
    assert got["gen"] == unlimited["gen"]
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "ingest.json")))["cases"]
    want = {json.dumps(r["case"], sort_keys=True): r["sha256"] for r in gold}
    assert got["ingest"] == want
    assert unlimited["ingest"] == want
    from tests.conftest import load_golden
    cuts = {(fx["name"], r["ncut"]): r["cuts"] for fx in load_golden() for r in fx["runs"]
            if r["partitioner"] == "rac"}
    assert got["rac"], "no rac fixtures ran"
    for name, p, c in got["rac"]:
        assert c == cuts[(name, p)], (name, p)
