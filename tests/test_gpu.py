"""GPU parity tests: every kernel form, per task, bit-exact.

Per-task uint64 counts from the device are compared with
  - the golden vectors written by the reference's own code (tests/golden),
  - the CPU oracle (restated reference CPU path) on random graphs,
  - closed forms;
and, at the full RMAT scale-26 size, with size-independent properties
(sum of tasks = total, forms agree, total invariant across cuts) plus the
oracle on a sample of tasks.
"""
import os
import subprocess

import numpy as np
import pytest

import bbtc_amd as B
import oracle as O
from tests.conftest import graph_from_spec, load_golden
from tests.test_count_cpu import CLOSED

pytestmark = pytest.mark.gpu

FORMS = ["merge", "bitmap", "dense", "rdense", "lane", "auto", "rows"]
GOLDEN = load_golden()
COUNTED = [(d, r) for d in GOLDEN for r in d["runs"] if "tasks" in r]


def log(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def dev():
    assert B.device_count() >= 1, "no HIP device"
    d = B.Device(0)
    yield d
    d.close()


def all_forms(dev, L, tasks):
    dg = dev.upload(L)
    out = {}
    for f in FORMS:
        out[f] = dg.count(tasks, f)
    choice, _ = dg.autotune(tasks)
    pl = dg.plan(tasks, "auto", choice)
    pl.launch()
    out["tuned"] = pl.results()
    pl.destroy()
    # every form with its rows per work item halved twice (choice bits 4-5)
    for form in (1, 2, 3, 4, 5):
        pl = dg.plan(tasks, "auto", np.full(len(tasks), form | (2 << 4), dtype=np.uint8))
        pl.launch()
        out[f"form{form}-rows/4"] = pl.results()
        pl.destroy()
    dg.release()
    return out


@pytest.mark.parametrize("d,r", COUNTED, ids=[f"{d['name']}-{r['partitioner']}{r['ncut']}" for d, r in COUNTED])
def test_golden_per_task_all_forms(dev, d, r):
    log("case", d["name"], r["partitioner"], r["ncut"])
    g = graph_from_spec(d["graph"])
    L = B.Layout(g, np.array(r["cuts"], dtype=np.uint32))
    ref = np.array(r["tasks"], dtype=np.uint64)
    tasks = ref[:, :3].astype(np.uint32)
    for form, (per, tot) in all_forms(dev, L, tasks).items():
        assert tot == r["total"], form
        assert np.array_equal(per, ref[:, 3]), form


@pytest.mark.parametrize("name,g,expect", CLOSED, ids=[c[0] for c in CLOSED])
def test_closed_forms_gpu(dev, name, g, expect):
    for p in (1, 3):
        L = B.Layout(g, g.partition("uni", p))
        t = L.tasks("desc")
        for form, (per, tot) in all_forms(dev, L, t).items():
            assert tot == expect, (form, p)


@pytest.mark.parametrize("seed", range(6))
def test_random_graphs_vs_oracle(dev, seed):
    rng = np.random.default_rng(seed + 50)
    if seed % 2:
        g = B.Graph.rmat(int(rng.integers(8, 15)), int(rng.integers(4, 32)), seed=seed)
    else:
        n = int(rng.integers(50, 20000))
        g = B.Graph.gnm(n, int(rng.integers(n, 20 * n)), seed)
    part = ["uni", "n1d", "pal"][seed % 3]
    p = int(rng.integers(1, 17))
    log("case seed", seed, "n", g.n, "m", g.m, part, p)
    cuts = g.partition(part, p)
    L = B.Layout(g, cuts)
    t = L.tasks("desc")
    R = O.RefLayout(g.offsets(), g.cols(), cuts)
    rper, rtot = R.count(np.array(t.tolist(), dtype=np.uint32))
    for form, (per, tot) in all_forms(dev, L, t).items():
        assert tot == rtot and np.array_equal(per, rper), form


@pytest.mark.parametrize("part,p", [("uni", 1), ("pal", 2), ("n1d", 3)])
def test_heavy_items_split(dev, part, p):
    # RMAT s16 with one to three bands: hub-core items far above the per-item
    # step budget (their wedges are spilled to the second phase in chunks) and
    # ac(u) lists of 64 rows larger than the LDS slice (row windows)
    g = B.Graph.rmat(16, 18, seed=1492)
    cuts = g.partition(part, p)
    L = B.Layout(g, cuts)
    t = L.tasks("desc")
    log("case heavy", part, p, "tasks", len(t))
    R = O.RefLayout(g.offsets(), g.cols(), cuts)
    rper, rtot = R.count(np.array(t.tolist(), dtype=np.uint32))
    for form, (per, tot) in all_forms(dev, L, t).items():
        assert tot == rtot and np.array_equal(per, rper), form


@pytest.mark.parametrize("p", [1, 2, 3])
def test_lane_slots_wide_bands(dev, p):
    # lane-only plans read short bc(v) rows from 16-byte row slots; bands of
    # 2^20 + 8, 2^19 + 5 and about 2^18.4 columns, rows of at most 8 entries
    n = (1 << 20) + 7
    g = B.Graph.gnm(n, 3 << 20, 77)
    cuts = g.partition("uni", p)
    L = B.Layout(g, cuts)
    t = L.tasks("desc")
    log("case lane slots", p, "widest band", int(np.max(np.diff(cuts.astype(np.int64)))))
    R = O.RefLayout(g.offsets(), g.cols(), cuts)
    rper, rtot = R.count(np.array(t.tolist(), dtype=np.uint32))
    dg = dev.upload(L)
    for form in ("lane", "auto"):
        pl = dg.plan(t, form)
        # every task in the lane form; AUTO then counts them by whole rows
        assert (pl.kinds() == (5 if form == "lane" else 6)).all(), form
        pl.launch()
        per, tot = pl.results()
        pl.destroy()
        assert tot == rtot and np.array_equal(per, rper), form
    dg.release()


def test_plan_own_order(dev):
    # an AUTO plan over several forms lists its tasks in an order of its own
    # (the forms interleaved); every per-task output comes back in the
    # caller's order: counts against the oracle, forms and the tasks the
    # clock profile finds empty against a plan of the reversed list
    g = B.Graph.rmat(17, 16, seed=7)
    cuts = g.partition("pal", 8)
    L = B.Layout(g, cuts)
    t = L.tasks("desc")
    rper, _ = O.RefLayout(g.offsets(), g.cols(), cuts).count(np.array(t.tolist(), dtype=np.uint32))
    dg = dev.upload(L)
    out = []
    for tl in (t, t[::-1].copy()):
        p = dg.plan(tl, "auto")
        p.launch()
        per, tot = p.results()
        out.append((per, tot, p.kinds(), p.profile(), p.time_tasks()))
        p.destroy()
    dg.release()
    (per, tot, kinds, cyc, ms), (per2, tot2, kinds2, cyc2, ms2) = out
    assert len(set(kinds.tolist())) >= 2, np.bincount(kinds)
    assert np.array_equal(per, rper) and tot == int(rper.sum())
    assert np.array_equal(per2, rper[::-1]) and tot2 == tot
    assert np.array_equal(kinds2, kinds[::-1])
    assert np.array_equal(cyc > 0, cyc2[::-1] > 0) and (cyc[rper > 0] > 0).all()
    assert ms.shape == (len(t),) and (ms >= 0).all() and (ms2 >= 0).all()


@pytest.mark.parametrize("case", range(8))
def test_rows_form_vs_oracle(dev, case):
    # the row form: every task of a lane-only plan counted by one launch over
    # whole rows (N+(u) merged with N+(v) across all bands), each common
    # entry added to its own task; sparse G(n, m) and RMAT graphs with 1-24
    # cuts (uni bands leave some tiles empty), per task against the oracle
    rng = np.random.default_rng(900 + case)
    if case % 4 == 3:
        g = B.Graph.rmat(int(rng.integers(8, 13)), 2, seed=case)
    else:
        n = int(rng.integers(2000, 200000))
        g = B.Graph.gnm(n, int(rng.integers(n, 12 * n)), 300 + case)
    part = ["uni", "pal", "n1d"][case % 3]
    p = int(rng.integers(1, 25))
    cuts = g.partition(part, p)
    L = B.Layout(g, cuts)
    t = L.tasks("desc")
    log("case rows", case, "n", g.n, "m", g.m, part, p)
    R = O.RefLayout(g.offsets(), g.cols(), cuts)
    rper, rtot = R.count(np.array(t.tolist(), dtype=np.uint32))
    dg = dev.upload(L)
    pl = dg.plan(t, "rows")
    kinds = pl.kinds()
    for _ in range(2):  # relaunch: counters cleared each time
        pl.launch()
        per, tot = pl.results()
        assert tot == rtot and np.array_equal(per, rper), kinds
    if (kinds == 6).any():
        # task ranges (the shared-queue and host-driven --border paths) keep
        # the lane form
        per_h, tot_h, _ = dg.count_hybrid(t, len(t) // 2, nthreads=2)
        assert tot_h == rtot and np.array_equal(per_h, rper)
    pl.destroy()
    # a subset of the tasks, in another order: the others' triangles are not counted
    sub = t[::-2].copy()
    rsub, rstot = R.count(np.array(sub.tolist(), dtype=np.uint32))
    per, tot = dg.count(sub, "rows")
    assert tot == rstot and np.array_equal(per, rsub)
    dg.release()


def test_rows_form_many_bands(dev):
    # 160 uni bands (695,520 tasks, the band starts past the rule's
    # statistics limit of 128 cuts): every band lookup and task slot of the
    # row form against the oracle
    g = B.Graph.gnm(20000, 150000, 21)
    cuts = g.partition("uni", 160)
    L = B.Layout(g, cuts)
    t = L.tasks("lex")
    R = O.RefLayout(g.offsets(), g.cols(), cuts)
    rper, rtot = R.count(np.array(t.tolist(), dtype=np.uint32))
    dg = dev.upload(L)
    pl = dg.plan(t, "rows")
    assert (pl.kinds() == 6).any()
    pl.launch()
    per, tot = pl.results()
    pl.destroy()
    dg.release()
    assert tot == rtot and np.array_equal(per, rper)


def test_rows_form_declined(dev):
    # rows of more than 64 entries (G(n, m) of average degree 160): no row
    # table, the plan keeps the lane or other forms; a task listed twice: the
    # row form is declined for that plan
    g = B.Graph.gnm(4000, 320000, 5)
    cuts = g.partition("uni", 8)
    L = B.Layout(g, cuts)
    t = L.tasks("desc")
    R = O.RefLayout(g.offsets(), g.cols(), cuts)
    rper, rtot = R.count(np.array(t.tolist(), dtype=np.uint32))
    dg = dev.upload(L)
    pl = dg.plan(t, "rows")
    assert not (pl.kinds() == 6).any()
    pl.launch()
    per, tot = pl.results()
    pl.destroy()
    assert tot == rtot and np.array_equal(per, rper)
    dg.release()
    g = B.Graph.gnm(30000, 60000, 6)
    cuts = g.partition("pal", 4)
    L = B.Layout(g, cuts)
    t = L.tasks("desc")
    t2 = np.concatenate([t, t[:3]])
    R = O.RefLayout(g.offsets(), g.cols(), cuts)
    rper, rtot = R.count(np.array(t2.tolist(), dtype=np.uint32))
    dg = dev.upload(L)
    pl = dg.plan(t2, "rows")
    assert not (pl.kinds() == 6).any() and (pl.kinds() == 5).any()
    pl.launch()
    per, tot = pl.results()
    pl.destroy()
    assert tot == rtot and np.array_equal(per, rper)
    dg.release()


def test_complete_graph_rows_wider_than_lds(dev):
    # K_2100: out-degrees up to 2099 exceed the 1792-word LDS slice, so the
    # sorted-list forms search those rows in HBM
    n = 2100
    iu = np.triu_indices(n, 1)
    g = B.Graph.from_edges(n, iu[0].astype(np.uint32), iu[1].astype(np.uint32))
    expect = n * (n - 1) * (n - 2) // 6
    for p in (1, 2):
        L = B.Layout(g, g.partition("uni", p))
        t = L.tasks("desc")
        for form, (per, tot) in all_forms(dev, L, t).items():
            assert tot == expect, (form, p)


@pytest.mark.parametrize("sep_frac", [0.0, 0.5, 1.0, 2.0])
def test_hybrid_border(dev, sep_frac):
    g = B.Graph.rmat(16, 16, seed=12)
    L = B.Layout(g, g.partition("pal", 16))
    t = L.tasks("desc")
    cper, ctot = L.count_cpu(t)
    dg = dev.upload(L)
    sep = int(sep_frac * len(t))
    for nthreads, nstreams in [(1, 1), (8, 2), (16, 4)]:
        per, tot, gt = dg.count_hybrid(t, sep, nthreads=nthreads, nstreams=nstreams)
        assert tot == ctot and np.array_equal(per, cper)
        assert gt >= min(sep, len(t))
        if nthreads == 1:
            assert gt == len(t)  # no host workers: the device counts everything
    dg.release()


def test_plan_relaunch_and_timing(dev):
    g = B.Graph.rmat(15, 16, seed=4)
    L = B.Layout(g, g.partition("pal", 8))
    t = L.tasks("desc")
    dg = dev.upload(L)
    pl = dg.plan(t, "auto")
    pl.launch()
    a = pl.results()[0]
    ms = pl.time(3)
    b = pl.results()[0]
    assert np.array_equal(a, b) and len(ms) == 3 and (ms > 0).all()
    tt = pl.time_tasks()
    assert len(tt) == len(t) and np.array_equal(pl.results()[0], a)
    cyc = pl.profile()
    assert np.array_equal(pl.results()[0], a) and cyc.sum() > 0
    pl.destroy()
    dg.release()


def test_subset_tasks_and_empty(dev):
    g = B.Graph.rmat(12, 8, seed=2)
    L = B.Layout(g, g.partition("uni", 4))
    t = L.tasks("lex")
    dg = dev.upload(L)
    cper, _ = L.count_cpu(t)
    sub = t[::3]
    per, tot = dg.count(sub, "auto")
    assert np.array_equal(per, cper[::3])
    per, tot = dg.count(np.zeros((0, 3), np.uint32), "auto")
    assert tot == 0
    with pytest.raises(B.BbtcError):
        dg.count(np.array([[2, 1, 3]], np.uint32))
    dg.release()


def test_budget_is_enforced():
    g = B.Graph.rmat(12, 8, seed=2)
    L = B.Layout(g, g.partition("uni", 4))
    d = B.Device(0, hbm_budget_bytes=1024)
    with pytest.raises(B.BbtcError) as e:
        d.upload(L)
    assert e.value.status == 7  # BBTC_ERR_BUDGET
    assert d.bytes_in_use == 0
    d.close()


def test_read_bandwidth(dev):
    before = dev.bytes_in_use
    gbs = dev.read_bandwidth(1 << 30, 3)
    log("streaming read GB/s", gbs.tolist())
    # a streaming read of 1 GiB runs well above PCIe rates and below the 8 TB/s spec peak
    assert gbs.shape == (3,) and np.all(gbs > 500) and np.all(gbs < 8000)
    assert dev.bytes_in_use == before
    with pytest.raises(B.BbtcError) as e:
        dev.read_bandwidth(1024, 1)
    assert e.value.status == 1  # BBTC_ERR_INVALID


def test_cli_gpu(tmp_path):
    d = load_golden("c1_rmat_s16_ef8")[0]
    g = graph_from_spec(d["graph"])
    path = str(tmp_path / "c1.bbtc")
    g.save(path)
    for run in d["runs"]:
        for extra in ([], ["--border", "-1"], ["--border", "7", "--nstream", "2"], ["--copy-first"],
                      ["--kernel", "tuned", "--border", "-1"], ["--kernel", "rdense"], ["--kernel", "dense"],
                      ["--kernel", "lane"], ["--device-layout"], ["--device-layout", "--border", "-1"],
                      ["--device-layout", "--kernel", "tuned", "--border", "5"]):
            out = tmp_path / "pt.txt"
            cmd = [B.CLI_PATH, "--graph", path, "--partitioner", run["partitioner"], "--ncut", str(run["ncut"]),
                   "--repeat", "2", "--ngpu", "1", "--per-task-out", str(out)] + extra
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
            assert r.returncode == 0, r.stderr
            assert f"Number of Triangles: {run['total']}" in r.stdout
            assert "Number of GPUs: 1" in r.stdout
            assert ("# Layout built on the device" in r.stderr) == ("--device-layout" in extra)
            assert "Partitioning and Layout construction (s)" in r.stdout
            got = {tuple(map(int, ln.split()[:3])): int(ln.split()[3]) for ln in out.read_text().splitlines()}
            assert got == {tuple(x[:3]): x[3] for x in run["tasks"]}


def test_cli_rows_form(tmp_path):
    # a G(n, m) graph whose tasks all admit the lane form: --kernel rows
    # counts it with the row form (and with --border, the contested task
    # ranges in the lane form); per-task output equals the golden vectors
    d = load_golden("small_gnm_2k")[0]
    g = graph_from_spec(d["graph"])
    path = str(tmp_path / "g.bbtc")
    g.save(path)
    for run in d["runs"]:
        for extra in (["--kernel", "rows"], ["--kernel", "rows", "--border", "3"],
                      ["--kernel", "rows", "--device-layout", "--border", "-1"]):
            out = tmp_path / "pt.txt"
            cmd = [B.CLI_PATH, "--graph", path, "--partitioner", run["partitioner"], "--ncut", str(run["ncut"]),
                   "--repeat", "2", "--ngpu", "1", "--per-task-out", str(out)] + extra
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
            assert r.returncode == 0, r.stderr
            assert f"Number of Triangles: {run['total']}" in r.stdout
            got = {tuple(map(int, ln.split()[:3])): int(ln.split()[3]) for ln in out.read_text().splitlines()}
            assert got == {tuple(x[:3]): x[3] for x in run["tasks"]}


def test_cli_forms_out_and_in(tmp_path):
    # the forms measured by --kernel tuned, written once and reused: the same
    # per-task counts without autotuning; a file for other cuts is refused
    d = load_golden("c1_rmat_s16_ef8")[0]
    path = str(tmp_path / "c1.bbtc")
    graph_from_spec(d["graph"]).save(path)
    run = next(x for x in d["runs"] if x["partitioner"] == "pal")
    forms = tmp_path / "c1.forms"
    base = [B.CLI_PATH, "--graph", path, "--partitioner", "pal", "--ncut", str(run["ncut"]), "--repeat", "1",
            "--ngpu", "1", "--border", "-1"]
    per = {}
    for name, extra in (("out", ["--kernel", "tuned", "--forms-out", str(forms)]), ("in", ["--forms-in", str(forms)])):
        out = tmp_path / f"pt_{name}.txt"
        r = subprocess.run(base + extra + ["--per-task-out", str(out)], capture_output=True, text=True, timeout=120)
        log("cli forms", name, r.returncode)
        assert r.returncode == 0, r.stderr
        assert f"Number of Triangles: {run['total']}" in r.stdout
        assert ("Autotune kernel forms" in r.stdout) == (name == "out")
        per[name] = {tuple(map(int, ln.split()[:3])): int(ln.split()[3]) for ln in out.read_text().splitlines()}
    assert per["out"] == per["in"] == {tuple(x[:3]): x[3] for x in run["tasks"]}
    words = forms.read_text().split()
    g = graph_from_spec(d["graph"])
    assert words[:2] == ["bbtc-forms", "2"] and int(words[3]) == len(run["tasks"])
    assert (int(words[4]), int(words[5])) == (g.n, g.m)
    # the same cuts and task count but another graph's fingerprint: refused
    other = tmp_path / "other.forms"
    other.write_text(" ".join(words[:6] + [str(int(words[6]) ^ 1)] + words[7:]) + "\n")
    r = subprocess.run(base + ["--forms-in", str(other)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "written for another graph" in r.stderr
    r = subprocess.run([B.CLI_PATH, "--graph", path, "--partitioner", "uni", "--ncut", str(run["ncut"]),
                        "--repeat", "1", "--ngpu", "1", "--forms-in", str(forms)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 1 and "written for other cuts" in r.stderr


def test_cli_batch_and_serve(tmp_path):
    # golden graphs (totals written by the reference's code) as list files
    small = [d for d in GOLDEN if d["fingerprint"]["m"] < 3_000_000][:5]
    paths, totals = [], []
    for x, d in enumerate(small):
        p = str(tmp_path / f"g{x}.bbtc")
        graph_from_spec(d["graph"]).save(p)
        paths.append(p)
        totals.append(d["runs"][0]["total"])
    batch = tmp_path / "batch.txt"
    batch.write_text("".join(f"{p} {1 + x % 4}\n" if x % 2 else f"{p}\n" for x, p in enumerate(paths)))
    for extra in ([], ["--kernel", "tuned"], ["--ncut", "3", "--partitioner", "uni"]):
        r = subprocess.run([B.CLI_PATH, "--batch", str(batch), "--repeat", "2"] + extra, capture_output=True,
                           text=True, timeout=300)
        log("cli batch", extra, r.returncode)
        assert r.returncode == 0, r.stderr
        assert f"Number of Graphs: {len(paths)}" in r.stdout
        assert f"Number of Triangles: {sum(totals)}" in r.stdout
        for x, t in enumerate(totals):
            assert f"triangles {t}," in [ln for ln in r.stdout.splitlines() if ln.startswith(f"Graph {x} ")][0]
        assert r.stdout.count("Batch Counting Execution Time (s)") == 2
    serve = tmp_path / "serve.txt"
    serve.write_text("".join(f"{p} {0.01 * (len(paths) - x)}\n" for x, p in enumerate(paths)))  # out of order
    r = subprocess.run([B.CLI_PATH, "--serve", str(serve)], capture_output=True, text=True, timeout=300)
    log("cli serve", r.returncode)
    assert r.returncode == 0, r.stderr
    assert f"Number of Triangles: {sum(totals)}" in r.stdout
    for x, t in enumerate(totals):
        assert f"triangles {t}," in [ln for ln in r.stdout.splitlines() if ln.startswith(f"Request {x} ")][0]
    assert "Latency p99 (ms)" in r.stdout


@pytest.mark.parametrize("spec", [("rmat", 10, 8, 1), ("rmat", 18, 16, 7), ("rmat", 20, 32, 5),
                                  ("gnm", 5000, 40000, 3), ("gnm", 1 << 18, 1 << 21, 9)])
def test_device_generator_matches_host(dev, spec):
    from tests.golden.fingerprint import fingerprint
    if spec[0] == "rmat":
        h = B.Graph.rmat(spec[1], spec[2], seed=spec[3])
        d = B.Graph.rmat_device(dev, spec[1], spec[2], seed=spec[3])
    else:
        h = B.Graph.gnm(spec[1], spec[2], spec[3])
        d = B.Graph.gnm_device(dev, spec[1], spec[2], spec[3])
    assert np.array_equal(h.offsets(), d.offsets()) and np.array_equal(h.cols(), d.cols())
    assert dev.bytes_in_use == 0


def test_device_generator_golden_graphs(dev):
    from tests.golden.fingerprint import fingerprint
    for d in GOLDEN:
        s = d["graph"]
        if s["kind"] == "rmat":
            g = B.Graph.rmat_device(dev, s["scale"], s["ef"], s["a"], s["b"], s["c"], s["seed"])
        else:
            g = B.Graph.gnm_device(dev, s["n"], s["m"], s["seed"])
        assert fingerprint(g.offsets(), g.cols()) == d["fingerprint"], d["name"]


# ---- full size: RMAT scale 26 ------------------------------------------------
@pytest.fixture(scope="module")
def s26():
    if os.environ.get("BBTC_SKIP_S26"):
        pytest.skip("BBTC_SKIP_S26 set")
    d = B.Device(0)
    g = B.Graph.rmat_device(d, 26, 16, 0.57, 0.19, 0.19, 3)
    d.close()
    log("s26 generated", g.info)
    return g


def test_s26_properties(dev, s26):
    from tests.golden.fingerprint import fingerprint
    g = s26
    # the reference's own cuts, task order and per-task counts of this graph
    # (tests/golden/c3_rmat_s26_ef16.json, written by its CPU code)
    (fx,) = load_golden("c3_rmat_s26_ef16")
    (run,) = fx["runs"]
    assert fingerprint(g.offsets(), g.cols()) == fx["fingerprint"], "generator drifted"
    cuts = g.partition("pal", 32)
    assert cuts.tolist() == run["cuts"]
    L = B.Layout(g, cuts)
    t = L.tasks("desc")
    asc = L.tasks("asc")
    assert [tuple(x) for x in asc.tolist()] == [tuple(x) for x in run["order_asc"]]
    dg = dev.upload(L)
    per_a, tot_a = dg.count(t, "auto")
    per_r, tot_r = dg.count(t, "rdense")
    per_l, tot_l = dg.count(t, "lane")
    assert int(per_a.sum()) == tot_a == run["total"]
    assert tot_a == tot_r and np.array_equal(per_a, per_r)
    assert tot_a == tot_l and np.array_equal(per_a, per_l)
    # every task against the reference's count
    want = {tuple(x[:3]): x[3] for x in run["tasks"]}
    got = {tuple(x): int(c) for x, c in zip(t.tolist(), per_a)}
    assert got == want
    # and every 97th task (lexicographic) against the oracle, bit-exact
    lex = np.lexsort((t["k"], t["j"], t["i"]))[::97]
    R = O.RefLayout(g.offsets(), g.cols(), cuts)
    tt = np.stack([t["i"][lex], t["j"][lex], t["k"][lex]], 1).astype(np.uint32)
    rper, _ = R.count(tt)
    del R
    assert np.array_equal(rper, per_a[lex])
    dg.release()
    log("s26 pal total", tot_a)
    # the total does not depend on the cuts
    L2 = B.Layout(g, g.partition("uni", 16))
    t2 = L2.tasks("desc")
    dg2 = dev.upload(L2)
    per2, tot2 = dg2.count(t2, "auto")
    dg2.release()
    assert tot2 == tot_a


# ---- batches -----------------------------------------------------------------
def oracle_lex(g, L):
    """The oracle's per-task counts of L's tasks in lexicographic order, and
    their total."""
    R = O.RefLayout(g.offsets(), g.cols(), L.cuts)
    return R.count(np.array(L.tasks("lex").tolist(), dtype=np.uint32))


@pytest.mark.parametrize("order", ["heavy_first", "small_first"])
def test_batch_matches_per_graph(dev, order):
    from bbtc_amd import workloads as W
    specs = W.batched_specs(48, 5000)
    graphs = [W.make_graph(dev, s) for s in specs]
    ncut = [1, 2, 4, 8] * 12
    Ls = B.layouts_build(graphs, ncut, "pal")
    # per task: the oracle's count of each graph's tasks (lexicographic)
    ref = [oracle_lex(g, L) for g, L in zip(graphs, Ls)]
    b = B.Batch(dev, Ls, "auto", order=order)
    gidx, bt = b.tasks()
    assert len(gidx) == sum(len(r[0]) for r in ref)
    # the lane form only where no row of the task's three tiles holds more
    # than 32 entries (the batch's tile table carries each tile's longest row)
    kinds = b.kinds()
    for g, (i, j, k), kd in zip(gidx.tolist(), bt.tolist(), kinds.tolist()):
        longest = max(Ls[g].tile(i, j).max_row, Ls[g].tile(j, k).max_row, Ls[g].tile(i, k).max_row)
        assert kd != 5 or longest <= 32, (g, i, j, k)
    assert (kinds != 5).any()
    # per task of the batch: the oracle's count of that graph's task
    lex_index = [{tuple(x): q for q, x in enumerate(L.tasks("lex").tolist())} for L in Ls]
    want = np.array([ref[g][0][lex_index[g][tuple(x)]] for g, x in zip(gidx, bt.tolist())], dtype=np.uint64)
    for mode in ["auto", "tuned", "captured"]:
        if mode == "tuned":
            b.autotune()
        if mode == "captured":
            b.capture()
        for rep in range(3):  # replays, with host work between them
            b.launch()
            pg, pt, lat = b.results()
            assert np.array_equal(pg, np.array([r[1] for r in ref], dtype=np.uint64)), (mode, rep)
            assert np.array_equal(pt, want), (mode, rep)
            assert (lat > 0).all() and np.isfinite(lat).all()
    b.destroy()


@pytest.mark.timeout(900)
def test_batch_4096_graphs(dev):
    # the full batched workload: 4,096 uneven graphs (70% RMAT s10-18 ef8-32,
    # 30% G(n,m)), cut counts from the cut-count autotuner, forms autotuned
    # over the batch, one captured launch; every graph's per-task counts equal
    # the plain per-graph launch, and every 64th graph's equal the oracle
    from bbtc_amd import workloads as W
    specs = W.batched_specs(4096)
    graphs = [W.make_graph(dev, s) for s in specs]
    ms = np.array([g.m for g in graphs])
    log("c4 graphs", len(graphs), "edges", int(ms.sum()))
    by_size = np.argsort(ms)
    rule, _ = B.tune_ncut(dev, [graphs[i] for i in by_size[::64]], (1, 2, 4, 8))
    ncut = [B.rule_ncut(rule, int(m)) for m in ms]
    log("c4 ncut", dict(zip(*np.unique(ncut, return_counts=True))))
    Ls = B.layouts_build(graphs, ncut, "pal")
    b = B.Batch(dev, Ls, "auto")
    b.autotune()
    b.capture()
    b.launch()
    pg, pt, lat = b.results()
    gidx, bt = b.tasks()
    b.destroy()
    log("c4 batch counted", int(pg.sum()))
    per_graph = [dict() for _ in Ls]
    for x, (gi, tk) in enumerate(zip(gidx.tolist(), bt.tolist())):
        per_graph[gi][tuple(tk)] = int(pt[x])
    for x, L in enumerate(Ls):
        t = L.tasks("desc")
        dg = dev.upload(L)
        per, tot = dg.count(t, "auto")
        dg.release()
        assert tot == pg[x], x
        assert per_graph[x] == {tuple(k): int(c) for k, c in zip(t.tolist(), per)}, x
    log("c4 plain launches agree")
    for x in range(0, len(Ls), 64):
        rper, rtot = oracle_lex(graphs[x], Ls[x])
        lex = [tuple(k) for k in Ls[x].tasks("lex").tolist()]
        assert rtot == pg[x], x
        assert {k: int(c) for k, c in zip(lex, rper)} == per_graph[x], x
    assert (lat > 0).all()


# ---- serving -----------------------------------------------------------------
def _serving_requests():
    """Small requests of uneven size and cut count; per-task oracle counts."""
    specs = [("rmat", 12, 8, 1, "pal", 4), ("gnm", 3000, 20000, 2, "uni", 2), ("rmat", 14, 16, 3, "pal", 8),
             ("rmat", 10, 4, 4, "n1d", 1), ("gnm", 20000, 160000, 5, "pal", 3), ("rmat", 13, 24, 6, "uni", 5),
             ("gnm", 64, 0, 7, "uni", 2), ("rmat", 16, 18, 1492, "uni", 1)]
    reqs = []
    for kind, a, b, seed, part, p in specs:
        if kind == "rmat":
            g = B.Graph.rmat(a, b, seed=seed)
        elif b:
            g = B.Graph.gnm(a, b, seed)
        else:
            g = B.Graph.from_csr(np.zeros(a + 1, np.uint64), np.zeros(0, np.uint32))
        cuts = g.partition(part, p)
        L = B.Layout(g, cuts)
        t = L.tasks("desc")
        rper, _ = O.RefLayout(g.offsets(), g.cols(), cuts).count(np.array(t.tolist(), dtype=np.uint32))
        reqs.append((L, t, rper))
    return reqs


@pytest.mark.parametrize("ncut", [8, 160])
def test_tasks_shard(dev, ncut):
    # shares for 1-8 GPUs: every task in exactly one share; each share's
    # load within the greedy (LPT) bound total/G + the largest task; each
    # share counted alone equals
    # the oracle on its tasks; 160 cuts: no statistics, the reference's
    # task weight as the estimate
    g = B.Graph.rmat(14, 16, seed=11)
    cuts = g.partition("uni" if ncut > 128 else "pal", ncut)
    L = B.Layout(g, cuts)
    t = L.tasks("desc")
    if ncut > 128:
        t = t[:4000]
    rper, _ = O.RefLayout(g.offsets(), g.cols(), cuts).count(np.array(t.tolist(), dtype=np.uint32))
    dg = dev.upload(L)
    for G in (1, 3, 8):
        sh, cost = dg.shard(t, G)
        assert sh.shape == (len(t),) and (sh < G).all() and (cost >= 0).all()
        if G == 1:
            assert (sh == 0).all()
        load = np.bincount(sh, weights=cost, minlength=G)
        assert load.max() <= cost.sum() / G + cost.max() * (1 + 1e-9)
        got = np.zeros(len(t), dtype=np.uint64)
        for s in range(G):
            sel = sh == s
            if sel.any():
                got[sel], _ = dg.count(t[sel], "auto")
        assert np.array_equal(got, rper), G
    with pytest.raises(B.BbtcError):
        dg.shard(t, 0)
    dg.release()


def test_pinned_layouts(dev):
    # layouts moved to page-locked memory (bbtc_layout_pin): the upload, a
    # batch and the server copy straight from their arrays, small ones and
    # ones past the 8 MiB staging threshold (RMAT s18 columns: 15 MB); every
    # count against the oracle per task
    reqs = _serving_requests()[:4]
    g = B.Graph.rmat(18, 16, seed=5)
    cuts = g.partition("pal", 4)
    L = B.Layout(g, cuts)
    t = L.tasks("desc")
    rper, _ = O.RefLayout(g.offsets(), g.cols(), cuts).count(np.array(t.tolist(), dtype=np.uint32))
    reqs.append((L, t, rper))
    for L, _, _ in reqs:
        assert not L.pinned
        assert dev.pin(L) is L and L.pinned
        dev.pin(L)  # a second pin leaves it as it is
        assert L.pinned
    for L, t, cpu in reqs:
        dg = dev.upload(L)
        per, tot = dg.count(t, "auto")
        dg.release()
        assert np.array_equal(per, cpu) and tot == int(cpu.sum())
    srv = B.Server(dev, heap_bytes=1 << 30)
    ids = [srv.submit(L, t, 0.0) for L, t, _ in reqs]
    got = {r.id: r for r in srv.drain()}
    srv.destroy()
    for i, (_, _, cpu) in zip(ids, reqs):
        assert got[i].status == 0 and np.array_equal(got[i].per_task, cpu)
    Ls = [L for L, _, _ in reqs]
    b = B.Batch(dev, Ls, "auto")
    gidx, bt = b.tasks()
    b.launch()
    pg, pt, _ = b.results()
    b.destroy()
    assert np.array_equal(pg, np.array([int(c.sum()) for _, _, c in reqs], dtype=np.uint64))
    index = [{tuple(x): q for q, x in enumerate(t.tolist())} for _, t, _ in reqs]
    want = np.array([reqs[gi][2][index[gi][tuple(x)]] for gi, x in zip(gidx.tolist(), bt.tolist())],
                    dtype=np.uint64)
    assert np.array_equal(pt, want)


def test_server_rows_form(dev):
    # requests the plain path's AUTO counts in the row form (sparse G(n, m)
    # graphs) mixed with ones it does not (RMAT): the server counts the
    # first in launches of their own, as many as the plain plans take the
    # row form, taking turns with shared launches of the others; every
    # count per task against the oracle, with the lane kernel asked for
    # instead (no row form) as well
    specs = [("gnm", 30000, 150000, 31, "pal", 4), ("rmat", 13, 16, 32, "pal", 4), ("gnm", 5000, 30000, 33, "uni", 2),
             ("gnm", 40000, 200000, 34, "uni", 8), ("rmat", 12, 8, 35, "uni", 1), ("gnm", 8000, 60000, 36, "pal", 1)]
    reqs, want_rows = [], 0
    for kind, a, b, seed, part, p in specs:
        g = B.Graph.rmat(a, b, seed=seed) if kind == "rmat" else B.Graph.gnm(a, b, seed)
        cuts = g.partition(part, p)
        L = B.Layout(g, cuts)
        t = L.tasks("desc")
        rper, _ = O.RefLayout(g.offsets(), g.cols(), cuts).count(np.array(t.tolist(), dtype=np.uint32))
        dg = dev.upload(L)
        pl = dg.plan(t, "auto")
        want_rows += int((pl.kinds() == 6).any())
        pl.destroy()
        dg.release()
        reqs.append((L, t, rper))
    assert want_rows >= 3
    for kernel, rows in (("auto", want_rows), ("lane", 0)):
        srv = B.Server(dev, heap_bytes=1 << 30, launch_us=300)
        ids = [srv.submit(L, t, 0.0, kernel=kernel) for L, t, _ in reqs]
        got = {r.id: r for r in srv.drain()}
        st = srv.stats()
        srv.destroy()
        assert st["row_form_requests"] == rows, (kernel, st)
        for i, (_, _, cpu) in zip(ids, reqs):
            assert got[i].status == 0 and np.array_equal(got[i].per_task, cpu), (kernel, i)


def test_server_alone_vs_mix(dev):
    reqs = _serving_requests()
    # alone: one request in the server at a time
    srv = B.Server(dev, heap_bytes=1 << 30)
    alone = []
    for L, t, cpu in reqs:
        srv.submit(L, t)
        (r,) = srv.drain()
        assert r.status == 0 and np.array_equal(r.per_task, cpu)
        alone.append(r.per_task.copy())
    srv.destroy()
    # in the mix: everything at once, 64 items per launch, so launches mix
    # slices of several requests and large ones span many launches
    srv = B.Server(dev, heap_bytes=1 << 30, quantum_items=64)
    ids = [srv.submit(L, t, 0.0) for L, t, _ in reqs]
    got = {r.id: r for r in srv.drain()}
    st = srv.stats()
    srv.destroy()
    log("mix stats", st)
    assert st["peak_segments"] > 1 and st["launches"] > len(reqs)
    for x, i in enumerate(ids):
        assert got[i].status == 0
        assert np.array_equal(got[i].per_task, alone[x]), x
        assert got[i].total == int(alone[x].sum())
    assert max(r.launches for r in got.values()) > 1
    # launches composed by estimated work with a small target time, with and
    # without lane-form requests in launches of their own
    for lane_form in (False, True):
        srv = B.Server(dev, heap_bytes=1 << 30, launch_us=200, lane_form=lane_form, oldest_pct=60)
        ids = [srv.submit(L, t, 0.0) for L, t, _ in reqs]
        got = {r.id: r for r in srv.drain()}
        st = srv.stats()
        srv.destroy()
        log("work-budget mix stats", lane_form, st)
        for x, i in enumerate(ids):
            assert got[i].status == 0 and np.array_equal(got[i].per_task, alone[x]), (lane_form, x)
    # plain mode (one request at a time, whole plan per launch) agrees too
    srv = B.Server(dev, heap_bytes=1 << 30, sequential=True)
    ids = [srv.submit(L, t, 0.0) for L, t, _ in reqs]
    got = {r.id: r for r in srv.drain()}
    assert srv.stats()["peak_active"] == 1
    srv.destroy()
    for x, i in enumerate(ids):
        assert np.array_equal(got[i].per_task, alone[x])


def test_server_heap_admission(dev):
    reqs = _serving_requests()
    g = B.Graph.rmat(12, 8, seed=11)
    L = B.Layout(g, g.partition("pal", 4))
    t = L.tasks("desc")
    cpu, _ = O.RefLayout(g.offsets(), g.cols(), L.cuts).count(np.array(t.tolist(), dtype=np.uint32))
    srv = B.Server(dev, heap_bytes=4 << 20)
    # larger than the whole heap: refused at submission
    with pytest.raises(B.BbtcError) as e:
        srv.submit(reqs[-1][0], reqs[-1][1])
    assert e.value.status == 7  # BBTC_ERR_BUDGET
    # twelve copies that fit only a few at a time: admission waits for memory
    ids = [srv.submit(L, t, 0.0) for _ in range(12)]
    got = srv.drain()
    st = srv.stats()
    srv.destroy()
    assert len(got) == 12 and all(r.status == 0 and np.array_equal(r.per_task, cpu) for r in got)
    assert st["peak_resident_bytes"] <= st["heap_bytes"] and st["peak_active"] < 12
    assert sorted(r.id for r in got) == sorted(ids)


def test_server_admits_simultaneous_arrivals_heaviest_first(dev):
    # three requests arriving together (after submission): admitted by
    # estimated work, largest first; a later arrival waits for its time
    specs = [B.Graph.gnm(2000, 8000, 3), B.Graph.rmat(16, 16, seed=4), B.Graph.rmat(13, 8, seed=5)]
    reqs = []
    for g in specs:
        L = B.Layout(g, g.partition("pal", 4))
        t = L.tasks("desc")
        cpu, _ = O.RefLayout(g.offsets(), g.cols(), L.cuts).count(np.array(t.tolist(), dtype=np.uint32))
        reqs.append((L, t, cpu))
    srv = B.Server(dev, heap_bytes=1 << 30)
    srv.reset_clock()
    ids = [srv.submit(L, t, 0.3) for L, t, _ in reqs] + [srv.submit(reqs[0][0], reqs[0][1], 0.35)]
    got = {r.id: r for r in srv.drain()}
    srv.destroy()
    admit = [got[i].admit_s for i in ids]
    assert admit[1] < admit[2] < admit[0] <= admit[3], admit
    assert all(a >= 0.3 for a in admit) and admit[3] >= 0.35
    for i, (_, _, cpu) in zip(ids, reqs + [reqs[0]]):
        assert got[i].status == 0 and np.array_equal(got[i].per_task, cpu)


def test_server_poisson_arrivals(dev):
    from bbtc_amd.workloads import poisson_arrivals
    reqs = _serving_requests() * 3
    arr = poisson_arrivals(len(reqs), 200.0, seed=7)
    srv = B.Server(dev, heap_bytes=1 << 30, quantum_items=256)
    srv.reset_clock()
    ids = [srv.submit(L, t, a) for (L, t, _), a in zip(reqs, arr)]
    got = {r.id: r for r in srv.drain()}
    srv.destroy()
    for (L, t, cpu), i, a in zip(reqs, ids, arr):
        r = got[i]
        assert r.status == 0 and np.array_equal(r.per_task, cpu)
        assert abs(r.arrival_s - a) < 1e-9 and r.admit_s >= a and r.done_s >= r.start_s >= r.ready_s >= r.admit_s
