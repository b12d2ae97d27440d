"""The bbtc command line on the CPU path (--ngpu 0): flags, validation and
report lines of the reference driver (src/bbtc.cu:15-90, utils.hpp:21-154),
counts against the golden vectors."""
import subprocess

import numpy as np
import pytest

import bbtc_amd as B
from tests.conftest import graph_from_spec, load_golden


def run(args, timeout=120):
    return subprocess.run([B.CLI_PATH] + args, capture_output=True, text=True, timeout=timeout)


@pytest.fixture(scope="module")
def c1(tmp_path_factory):
    d = load_golden("c1_rmat_s16_ef8")[0]
    path = str(tmp_path_factory.mktemp("cli") / "c1.bbtc")
    graph_from_spec(d["graph"]).save(path)
    return d, path


def test_cpu_path_counts_and_report(c1, tmp_path):
    d, path = c1
    for r in d["runs"]:
        out = tmp_path / f"pt_{r['partitioner']}{r['ncut']}.txt"
        p = run(["--graph", path, "--partitioner", r["partitioner"], "--ncut", str(r["ncut"]), "--ngpu", "0",
                 "--repeat", "2", "--per-task-out", str(out)])
        assert p.returncode == 0, p.stderr
        lines = p.stdout.splitlines()
        assert f"Graph file: {path}" in lines
        assert f"Number of Cuts: {r['ncut']}" in lines
        assert f"Number of Tasks: {len(r['tasks'])}" in lines
        assert f"Number of Triangles: {r['total']}" in lines
        assert sum(ln.startswith("Triangle Counting Execution Time (s)") for ln in lines) == 2
        assert any(ln.startswith("Partitioning and Layout construction (s)") for ln in lines)
        got = {tuple(map(int, ln.split()[:3])): int(ln.split()[3]) for ln in out.read_text().splitlines()}
        assert got == {tuple(x[:3]): x[3] for x in r["tasks"]}


def test_rac_partitioner(c1, tmp_path):
    d, path = c1
    r = next(x for x in d["runs"] if x["partitioner"] == "rac")
    out = tmp_path / "pt.txt"
    p = run(["--graph", path, "--partitioner", "rac", "--ncut", str(r["ncut"]), "--ngpu", "0", "--repeat", "1",
             "--per-task-out", str(out)])
    assert p.returncode == 0, p.stderr
    assert f"Number of Triangles: {r['total']}" in p.stdout
    got = {tuple(map(int, ln.split()[:3])): int(ln.split()[3]) for ln in out.read_text().splitlines()}
    assert got == {tuple(x[:3]): x[3] for x in r["tasks"]}


def test_cuts_file(c1, tmp_path):
    d, path = c1
    r = d["runs"][0]
    cf = tmp_path / "cuts.txt"
    cf.write_text(" ".join(map(str, r["cuts"])))
    p = run(["--graph", path, "--ncut", str(r["ncut"]), "--cuts-file", str(cf), "--ngpu", "0", "--repeat", "1"])
    assert p.returncode == 0, p.stderr
    assert f"Number of Triangles: {r['total']}" in p.stdout


def test_flags_are_case_insensitive(c1):
    d, path = c1
    p = run(["--GRAPH", path, "--Partitioner", "uni", "--NCUT", "4", "--NGPU", "0", "--Repeat", "1"])
    assert p.returncode == 0, p.stderr
    assert f"Number of Triangles: {d['runs'][0]['total']}" in p.stdout


@pytest.mark.parametrize("args,msg", [
    (["--ncut", "1"], "--ncut has to be set in between"),
    (["--nstream", "0"], "--nstream has to be set in between"),
    (["--border", "-2"], "--border has to be set in between"),
    (["--repeat", "0"], "--repeat has to be set in between"),
    (["--partitioner", "xyz"], "Wrong partitioning algorithm."),
    (["--kernel", "xyz"], "Wrong kernel."),
    (["--bogus", "1"], "Wrong argument '--bogus' given."),
    (["--device-layout"], "--device-layout builds one graph's layout on a GPU"),
])
def test_validation(c1, args, msg):
    _, path = c1
    p = run(["--graph", path, "--ngpu", "0"] + args)
    assert p.returncode == 2
    assert msg in p.stdout and "--graph" in p.stdout  # message, then usage


def test_missing_graph_and_usage(tmp_path):
    p = run(["--graph", str(tmp_path / "none.bbtc"), "--ngpu", "0"])
    assert p.returncode == 2 and "Graph doesn't exist" in p.stdout
    p = run(["--help"])
    assert p.returncode == 0 and "--partitioner" in p.stdout


def test_batch_and_serve_modes_validate(c1, tmp_path):
    # the list modes count on the device; without one they refuse, as they do
    # for a missing list or both modes at once (exit 2, usage printed)
    d, path = c1
    lst = tmp_path / "list.txt"
    lst.write_text(f"{path}\n")
    for args, msg in [(["--batch", str(lst), "--ngpu", "0"], "need a GPU"),
                      (["--serve", str(lst), "--ngpu", "0"], "need a GPU"),
                      (["--batch", str(tmp_path / "none.txt"), "--ngpu", "0"], "List file doesn't exist"),
                      (["--batch", str(lst), "--serve", str(lst)], "exclusive")]:
        p = run(args)
        assert p.returncode == 2 and msg in p.stdout, (args, p.stdout)
        assert "--batch <list>" in p.stdout


def test_cpu_path_task_order(c1, tmp_path):
    # without a GPU the task list is in the CPU build's ascending weight order
    # (tasks.hpp:29-32), so the workers, who claim from the back, start with
    # the heaviest tasks
    d, path = c1
    r = next(x for x in d["runs"] if x["partitioner"] == "pal")
    out = tmp_path / "pt.txt"
    p = run(["--graph", path, "--partitioner", "pal", "--ncut", str(r["ncut"]), "--ngpu", "0", "--repeat", "1",
             "--per-task-out", str(out)])
    assert p.returncode == 0, p.stderr
    got = np.array([list(map(int, ln.split()[:3])) for ln in out.read_text().splitlines()], dtype=np.uint32)
    L = B.Layout(graph_from_spec(d["graph"]), np.array(r["cuts"], dtype=np.uint32))
    assert np.array_equal(B.tasks_array(got), L.tasks("asc"))


def test_random_arguments_never_crash(c1, tmp_path):
    # seeded random flag/value mixes (the GPU left out): every run ends with a
    # result (0) or the reference's message and usage (2), never a signal
    import random
    _, path = c1
    flags = ["--graph", "--partitioner", "--ncut", "--nstream", "--border", "--copy-first", "--repeat", "--help",
             "--kernel", "--threads", "--cuts-file", "--per-task-out", "--batch", "--serve", "--save-bcsr",
             "--keep-duplicates", "--bogus", "-x", ""]
    vals = [path, str(tmp_path / "nope.bbtc"), "pal", "uni", "n1d", "rac", "xyz", "0", "1", "-1", "-2", "2", "4",
            "99999999999", "abc", "", "1e3", "0x10", "-0", "4294967296", "lane", "tuned", str(tmp_path / "o.txt")]
    rng = random.Random(5)
    for _ in range(120):
        args = []
        for _ in range(rng.randint(0, 8)):
            args.append(rng.choice(flags))
            if rng.random() < 0.8:
                args.append(rng.choice(vals))
        p = run(args + ["--ngpu", "0"])
        assert p.returncode in (0, 2), (args, p.returncode, p.stderr[-300:])


def test_forms_flags_validation(c1, tmp_path):
    _, path = c1
    f = tmp_path / "x.forms"
    f.write_text("bbtc-forms 1\n")
    for extra, msg in ((["--forms-out", str(f)], "need a GPU"), (["--forms-in", str(f)], "need a GPU"),
                       (["--forms-in", str(f), "--batch", path], "count one graph")):
        p = run(["--graph", path, "--ngpu", "0"] + extra)
        assert p.returncode == 2 and msg in p.stdout, (extra, p.stdout[:200])
