"""Text ingest: edge lists and Matrix Market files converted to .bbtc.

Pinned by golden SHA-256 hashes of the .bbtc files that the reference's own
driver (utils::ReadGraph, built from its sources) wrote for the same text
files (tests/golden/ingest.json, from tests/golden/make_golden_ingest.py),
and by brute-force triangle counts of the converted graphs."""
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

import bbtc_amd as B
import oracle as O
from tests.golden import textgen

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "ingest.json")))["cases"]


def sha(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def write_case(tmp_path, case):
    p = str(tmp_path / textgen.filename(case))
    with open(p, "wb") as f:
        f.write(textgen.make(case))
    return p


@pytest.mark.parametrize("rec", GOLD, ids=[r["case"] for r in GOLD])
def test_bytes_equal_reference(tmp_path, rec):
    src = write_case(tmp_path, rec["case"])
    out = str(tmp_path / "ours.bbtc")
    B.convert_text(src, out, dedup=False)
    assert os.path.getsize(out) == rec["bytes"]
    assert sha(out) == rec["sha256"]


@pytest.mark.parametrize("rec", GOLD, ids=[r["case"] for r in GOLD])
def test_dedup_and_counts(tmp_path, rec):
    src = write_case(tmp_path, rec["case"])
    out = str(tmp_path / "dedup.bbtc")
    B.convert_text(src, out, dedup=True)
    if rec["simple"]:  # nothing to drop: the same file
        assert sha(out) == rec["sha256"]
    g = B.Graph.load_text(src)
    n, _, m = g.info
    assert n == rec["n"] and m <= rec["m"]
    off, col = g.offsets(), g.cols()
    assert all(np.all(np.diff(col[off[r]:off[r + 1]].astype(np.int64)) > 0) for r in range(n))
    L = B.Layout(g, g.partition("uni", 3))
    _, tot = L.count_cpu(L.tasks("lex"))
    if n <= 2048:  # the brute-force counter's adjacency matrix limit
        assert tot == O.brute_force(off, col)
    if rec["simple"]:  # the reference's own count of the same file
        assert tot == rec["triangles_reference"]


def test_source_ids_beyond_destinations(tmp_path):
    # every edge written as (larger id, smaller id): the reference reads past
    # its transpose here (it crashes); the conversion treats the matrix as
    # square and gives the graph of the same edges written the other way
    rng = np.random.default_rng(3)
    a = rng.integers(0, 400, 3000)
    b = rng.integers(0, 400, 3000)
    hi, lo = np.maximum(a, b), np.minimum(a, b)
    keep = hi != lo
    p1, p2 = tmp_path / "hi_lo.txt", tmp_path / "lo_hi.txt"
    p1.write_text("".join(f"{x} {y}\n" for x, y in zip(hi[keep], lo[keep])))
    p2.write_text("".join(f"{y} {x}\n" for x, y in zip(hi[keep], lo[keep])))
    o1, o2 = str(tmp_path / "1.bbtc"), str(tmp_path / "2.bbtc")
    B.convert_text(str(p1), o1)
    B.convert_text(str(p2), o2)
    assert sha(o1) == sha(o2)


def test_grammar_edge_cases(tmp_path):
    def conv(text, name="t.txt"):
        p = tmp_path / name
        p.write_bytes(text)
        g = B.Graph.load_text(str(p))
        off, col = g.offsets(), g.cols()
        return g.info[0], sorted(tuple(sorted((r, int(c)))) for r in range(g.info[0])
                                 for c in col[off[r]:off[r + 1]])

    # last entry dropped when it ends exactly at end of file; kept after a space
    n1, e1 = conv(b"0 1\n1 2\n2 0")
    n2, e2 = conv(b"0 1\n1 2\n2 0 ")
    assert len(e1) == 2 and len(e2) == 3
    # empty file: one isolated vertex, no edges
    assert conv(b"") == (1, [])
    # comments and a Matrix Market header; ids stay 1-based (vertex 0 isolated)
    n3, e3 = conv(b"%%MatrixMarket matrix coordinate pattern general\n% c\n3 3 3\n1 2\n2 3\n3 1\n", "t.mtx")
    assert n3 == 4 and len(e3) == 3
    with pytest.raises(B.BbtcError):
        conv(b"%%MatrixMarket matrix coordinate pattern general\n3 3 5\n1 2\n2 3\n", "bad.mtx")
    with pytest.raises(B.BbtcError):
        conv(b"%%MatrixMarket matrix array real general\n3 3\n1\n2\n3\n", "arr.mtx")
    with pytest.raises(B.BbtcError):
        conv(b"3 2\n0 1 2\n", "m.graph")


def test_cli_text_input(tmp_path):
    rec = next(r for r in GOLD if r["case"] == "simple_hubs")
    src = write_case(tmp_path, "simple_hubs")
    r = subprocess.run([B.CLI_PATH, "--graph", src, "--ngpu", "0", "--ncut", "4", "--repeat", "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert f"Number of Triangles: {rec['triangles_reference']}" in r.stdout
    assert sha(src + ".bbtc") == rec["sha256"]
    # the saved sibling is used on the next run
    r2 = subprocess.run([B.CLI_PATH, "--graph", src, "--ngpu", "0", "--ncut", "4", "--repeat", "1"],
                        capture_output=True, text=True, timeout=120)
    assert "Saving binary CSR graph" not in r2.stderr and r2.returncode == 0


def test_sparse_id_space(tmp_path):
    # a few thousand edges among ids up to 5M: with several threads, threads x
    # rows outgrows the edges, so the rows are built with atomic slot counters
    # rather than per-thread histograms; the graph must be the same up to the
    # relabeling
    rng = np.random.default_rng(11)
    ids = np.sort(rng.choice(5_000_000, 600, replace=False))
    a, b = ids[rng.integers(0, 600, 4000)], ids[rng.integers(0, 600, 4000)]
    p = tmp_path / "sparse.txt"
    p.write_text("".join(f"{x} {y}\n" for x, y in zip(a, b)))
    g = B.Graph.load_text(str(p))
    n = g.info[0]
    assert n == int(max(a.max(), b.max())) + 1
    off, col = g.offsets().astype(np.int64), g.cols().astype(np.int64)
    # expected: the simple graph of the input, by closed form trace(A^3)/6
    keep = a != b
    used = np.unique(np.concatenate([a[keep], b[keep]]))
    pos = {int(v): i for i, v in enumerate(used)}
    A = np.zeros((len(used), len(used)), dtype=np.int64)
    for x, y in zip(a[keep], b[keep]):
        A[pos[int(x)], pos[int(y)]] = A[pos[int(y)], pos[int(x)]] = 1
    # converted graph: keep the rows that carry edges, in order
    src = np.repeat(np.arange(n), np.diff(off))
    touched = np.unique(np.concatenate([src, col]))
    assert len(touched) == len(used)
    at = {int(v): i for i, v in enumerate(touched)}
    C = np.zeros_like(A)
    for x, y in zip(src, col):
        assert x < y
        C[at[int(x)], at[int(y)]] = C[at[int(y)], at[int(x)]] = 1
    assert sorted(C.sum(0)) == sorted(A.sum(0))
    assert np.trace(C @ C @ C) == np.trace(A @ A @ A)
    L = B.Layout(g, g.partition("uni", 3))
    _, tot = L.count_cpu(L.tasks("lex"))
    assert tot == np.trace(A @ A @ A) // 6
