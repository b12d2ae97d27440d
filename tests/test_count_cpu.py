"""Triangle counts on the CPU path, and the oracle itself.

The oracle restates the reference; it is pinned here by
  - golden per-task counts written by the reference's own BBTc (tests/golden),
  - brute force over the adjacency matrix,
  - the per-task definition counter (band of u, v, w of each triangle),
  - closed forms: K_n -> C(n,3), K_{a,b}, trees and C_n (n > 3) -> 0,
    wheel W_n -> n, friendship graph F_k -> k.
The product's CPU path (the --ngpu 0 drop-in and host tail) must equal all
of them.
"""
import itertools
from math import comb

import numpy as np
import pytest

import bbtc_amd as B
import oracle as O
from tests.conftest import graph_from_spec, load_golden

GOLDEN = load_golden()
COUNTED = [(d, r) for d in GOLDEN for r in d["runs"] if "tasks" in r]


def edges_graph(n, edges):
    e = np.array(edges, dtype=np.uint32).reshape(-1, 2)
    return B.Graph.from_edges(n, e[:, 0], e[:, 1])


def complete(n):
    return edges_graph(n, list(itertools.combinations(range(n), 2)))


def bipartite(a, b):
    return edges_graph(a + b, [(i, a + j) for i in range(a) for j in range(b)])


def cycle(n):
    return edges_graph(n, [(i, (i + 1) % n) for i in range(n)])


def wheel(n):  # hub n, rim 0..n-1
    return edges_graph(n + 1, [(i, (i + 1) % n) for i in range(n)] + [(i, n) for i in range(n)])


def friendship(k):  # k triangles sharing vertex 0
    e = []
    for t in range(k):
        a, b = 1 + 2 * t, 2 + 2 * t
        e += [(0, a), (0, b), (a, b)]
    return edges_graph(2 * k + 1, e)


def star(n):
    return edges_graph(n, [(0, i) for i in range(1, n)])


CLOSED = [
    ("K3", complete(3), 1), ("K7", complete(7), 35), ("K40", complete(40), comb(40, 3)),
    ("K5,7", bipartite(5, 7), 0), ("C4", cycle(4), 0), ("C9", cycle(9), 0), ("C3", cycle(3), 1),
    ("W4", wheel(4), 4), ("W11", wheel(11), 11), ("F1", friendship(1), 1), ("F9", friendship(9), 9),
    ("star", star(30), 0), ("empty", B.Graph.from_csr(np.zeros(6, np.uint64), np.zeros(0, np.uint32)), 0),
]


def count_all(g, part, p, nthreads=0):
    cuts = g.partition(part, p)
    L = B.Layout(g, cuts)
    t = L.tasks("desc")
    per, tot = L.count_cpu(t, nthreads)
    return cuts, L, t, per, tot


@pytest.mark.parametrize("name,g,expect", CLOSED, ids=[c[0] for c in CLOSED])
@pytest.mark.parametrize("part,p", [("uni", 1), ("uni", 3), ("pal", 2), ("n1d", 4)])
def test_closed_forms(name, g, expect, part, p):
    _, _, _, per, tot = count_all(g, part, p)
    assert tot == expect == int(per.sum())
    assert O.brute_force(g.offsets(), g.cols()) == expect


@pytest.mark.parametrize("seed", range(8))
def test_random_small_vs_brute_force_and_definition(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(5, 200))
    m = int(rng.integers(0, min(n * (n - 1) // 2 // 2, 3000)) )
    g = B.Graph.gnm(n, m, seed) if m else B.Graph.from_csr(np.zeros(n + 1, np.uint64), np.zeros(0, np.uint32))
    bf = O.brute_force(g.offsets(), g.cols())
    for part in ("uni", "n1d", "pal"):
        for p in (1, 2, 5):
            cuts, L, t, per, tot = count_all(g, part, p)
            assert tot == bf
            cube, dtot = O.count_definition(g.offsets(), g.cols(), cuts)
            assert dtot == bf
            for q, x in enumerate(t.tolist()):
                assert per[q] == cube[x[0], x[1], x[2]]
            R = O.RefLayout(g.offsets(), g.cols(), cuts)
            tt = np.array(t.tolist(), dtype=np.uint32)
            rper, rtot = R.count(tt)
            assert rtot == bf and np.array_equal(rper, per)


@pytest.mark.parametrize("d,r", COUNTED, ids=[f"{d['name']}-{r['partitioner']}{r['ncut']}" for d, r in COUNTED])
def test_per_task_counts_match_reference(d, r):
    g = graph_from_spec(d["graph"])
    cuts = np.array(r["cuts"], dtype=np.uint32)
    L = B.Layout(g, cuts)
    ref = np.array(r["tasks"], dtype=np.uint64)
    tasks = ref[:, :3].astype(np.uint32)
    per, tot = L.count_cpu(tasks)
    assert tot == r["total"]
    assert np.array_equal(per, ref[:, 3])


@pytest.mark.parametrize("d,r", [x for x in COUNTED if x[0]["graph"].get("n", 0) < (1 << 22)],
                         ids=lambda x: x["name"] if isinstance(x, dict) and "name" in x else "")
def test_oracle_matches_reference(d, r):
    """Pins the oracle restatement on the reference's own numbers."""
    g = graph_from_spec(d["graph"])
    cuts = np.array(r["cuts"], dtype=np.uint32)
    R = O.RefLayout(g.offsets(), g.cols(), cuts)
    ref = np.array(r["tasks"], dtype=np.uint64)
    per, tot = R.count(ref[:, :3].astype(np.uint32))
    assert tot == r["total"] and np.array_equal(per, ref[:, 3])
    cube, dtot = O.count_definition(g.offsets(), g.cols(), cuts)
    assert dtot == r["total"]
    assert all(cube[i, j, k] == c for i, j, k, c in r["tasks"])


def test_total_invariant_across_cuts():
    g = B.Graph.rmat(13, 16, seed=9)
    totals = {count_all(g, part, p)[4] for part in ("uni", "n1d", "pal") for p in (1, 2, 7, 16)}
    assert len(totals) == 1


def test_thread_counts_agree():
    g = B.Graph.rmat(12, 16, seed=6)
    a = count_all(g, "pal", 8, 1)[3]
    b = count_all(g, "pal", 8, 8)[3]
    assert np.array_equal(a, b)


def test_subset_and_bad_tasks():
    g = B.Graph.rmat(10, 8, seed=3)
    L = B.Layout(g, g.partition("uni", 3))
    per, tot = L.count_cpu(np.zeros((0, 3), np.uint32))
    assert tot == 0 and len(per) == 0
    with pytest.raises(B.BbtcError):
        L.count_cpu(np.array([[1, 0, 2]], np.uint32))  # i > j
    with pytest.raises(B.BbtcError):
        L.count_cpu(np.array([[0, 1, 3]], np.uint32))  # k out of range


def test_dense_task_split_over_workers():
    # 1,025 apex rows each joined to all 10,001 vertices of a path: the apex
    # band's tile to the path has average degree 10,001 > 10,000, so its task
    # takes the stamp-table form, and its 1,025 non-empty rows are more than
    # one claim unit; every apex closes one triangle per path edge
    apex, base = 1025, 10001
    n = apex + base
    row_len = np.concatenate([np.full(apex, base), np.ones(base - 1), [0]]).astype(np.uint64)
    off = np.concatenate([[0], np.cumsum(row_len)]).astype(np.uint64)
    cols = np.concatenate([np.tile(np.arange(apex, n, dtype=np.uint32), apex),
                           np.arange(apex + 1, n, dtype=np.uint32)])
    g = B.Graph.from_csr(off, cols)
    L = B.Layout(g, np.array([0, apex, n + 1], np.uint32))
    assert L.tile(0, 1).avg_degree > 10000 and L.tile(0, 1).nz_rows == apex
    t = L.tasks("desc")
    expect = apex * (base - 1)
    per1, tot1 = L.count_cpu(t, 1)
    per8, tot8 = L.count_cpu(t, 8)
    assert tot1 == tot8 == expect and np.array_equal(per1, per8)
    R = O.RefLayout(g.offsets(), g.cols(), np.array([0, apex, n + 1], np.uint32))
    rper, rtot = R.count(np.array(t.tolist(), dtype=np.uint32))
    assert rtot == expect and np.array_equal(rper, per1)


def test_long_bc_row_merges():
    # vertex 0 joins 1 and 2; vertex 1 joins 2..301. With one band, the ac row
    # of 0 holds 2 entries and the bc row of 1 holds 300, more than 128 times
    # as many, so that wedge is merged instead of probed against the marks;
    # the one triangle is (0, 1, 2)
    rows = [[1, 2], list(range(2, 302))] + [[] for _ in range(300)]
    off = np.cumsum([0] + [len(r) for r in rows]).astype(np.uint64)
    cols = np.array([c for r in rows for c in r], dtype=np.uint32)
    g = B.Graph.from_csr(off, cols)
    assert O.brute_force(g.offsets(), g.cols()) == 1
    for part, p in (("uni", 1), ("uni", 3)):
        _, _, _, per, tot = count_all(g, part, p)
        assert tot == 1 == int(per.sum())


@pytest.mark.parametrize("stride", [2, 5, 64])
def test_oracle_row_samples_add_up(stride):
    # the oracle's row samples (rows f % stride == phase of every task, the
    # CPU baseline's timing sample) add up to the whole tasks' counts, also
    # for a task of average degree above 10,000 (the stamp-table form): 1,025
    # apex rows joined to all 10,001 vertices of a path
    apex, base = 1025, 10001
    n = apex + base
    row_len = np.concatenate([np.full(apex, base), np.ones(base - 1), [0]]).astype(np.uint64)
    off = np.concatenate([[0], np.cumsum(row_len)]).astype(np.uint64)
    cols = np.concatenate([np.tile(np.arange(apex, n, dtype=np.uint32), apex),
                           np.arange(apex + 1, n, dtype=np.uint32)])
    dense = (B.Graph.from_csr(off, cols), None, np.array([0, apex, n + 1], np.uint32))
    for g, part, p in ((B.Graph.rmat(14, 16, seed=3), "pal", 6), (B.Graph.gnm(600, 60000, 15), "uni", 3), dense):
        cuts = g.partition(part, p) if part else p
        R = O.RefLayout(g.offsets(), g.cols(), cuts)
        t = R.tasks("lex")
        full, tot = R.count(t)
        parts = [R.count_rows(t, stride, ph) for ph in range(stride)]
        assert sum(x[1] for x in parts) == tot
        assert np.array_equal(sum(x[0] for x in parts), full)
