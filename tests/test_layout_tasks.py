"""Block layout, task list and task order.

Layout: reconstructing the input CSR from the tiles, tile sizes and the
reference's float average degree (block_csr.cpp:272-278, restated by the
oracle's BlockCSR). Tasks: C(p+2,3) lexicographic triples; the ascending
order must equal the reference's std::sort order written into the golden
vectors by tasks::SortTasks itself.
"""
import numpy as np
import pytest

import bbtc_amd as B
import oracle as O
from tests.conftest import graph_from_spec, load_golden

GOLDEN = load_golden()
COUNTED = [(d, r) for d in GOLDEN for r in d["runs"] if "tasks" in r]


def reconstruct(L, g):
    """Rebuild each row from the tiles through the count-free accessors of
    the oracle restatement is not used here: tile sizes come from the
    library, and the rows from the per-tile nnz must sum to the CSR."""
    p = L.p
    total = 0
    for i in range(p):
        for k in range(p):
            t = L.tile(i, k)
            if k < i:
                assert t.nnz == 0
            total += t.nnz
    assert total == g.m


@pytest.mark.parametrize("spec,p,kind", [
    (dict(kind="rmat", scale=12, ef=16, seed=3, a=.57, b=.19, c=.19), 8, "pal"),
    (dict(kind="rmat", scale=10, ef=8, seed=1, a=.57, b=.19, c=.19), 1, "uni"),
    (dict(kind="gnm", n=3000, m=30000, seed=5), 5, "n1d"),
])
def test_tiles_match_reference_blockcsr(spec, p, kind):
    g = graph_from_spec(spec)
    cuts = g.partition(kind, p)
    L = B.Layout(g, cuts)
    R = O.RefLayout(g.offsets(), g.cols(), cuts)
    reconstruct(L, g)
    for i in range(p):
        for k in range(i, p):
            t = L.tile(i, k)
            tid = i * p + k
            assert t.nnz == R.tile_nnz(tid)
            assert t.height == R.tile_height(tid) == cuts[i + 1] - cuts[i]
            assert t.width == cuts[k + 1] - cuts[k]
            a, b = np.float32(t.avg_degree), np.float32(R.tile_deg(tid))
            assert (np.isnan(a) and np.isnan(b)) or a == b


def test_rows_reconstruct_input():
    g = B.Graph.rmat(11, 12, seed=8)
    p = 6
    cuts = g.partition("pal", p)
    R = O.RefLayout(g.offsets(), g.cols(), cuts)
    off, col = g.offsets(), g.cols()
    c = cuts.astype(np.int64)
    for u in range(0, g.n, 7):
        i = int(np.searchsorted(c, u, side="right") - 1)
        parts = [R.row(i * p + k, u - c[i]) for k in range(p)]
        assert np.array_equal(np.concatenate(parts), col[off[u]:off[u + 1]])


def test_nonempty_row_counts():
    g = B.Graph.rmat(10, 16, seed=2)
    cuts = g.partition("uni", 4)
    L = B.Layout(g, cuts)
    off, col = g.offsets().astype(np.int64), g.cols().astype(np.int64)
    c = cuts.astype(np.int64)
    rows = np.repeat(np.arange(g.n), np.diff(off))
    bi = np.searchsorted(c, rows, side="right") - 1
    bk = np.searchsorted(c, col, side="right") - 1
    for i in range(4):
        for k in range(i, 4):
            sel = (bi == i) & (bk == k)
            assert L.tile(i, k).nz_rows == len(np.unique(rows[sel]))
            # entries of the longest row, which decides where the lane form applies
            longest = int(np.bincount(rows[sel]).max()) if sel.any() else 0
            assert L.tile(i, k).max_row == longest


@pytest.mark.parametrize("p", [1, 2, 3, 4, 8, 16, 32])
def test_task_list(p):
    g = B.Graph.rmat(8, 8, seed=1)
    L = B.Layout(g, g.partition("uni", p))
    t = L.tasks("lex")
    assert len(t) == p * (p + 1) * (p + 2) // 6 == B.lib().bbtc_task_count(p)
    expect = [(i, j, k) for i in range(p) for j in range(i, p) for k in range(j, p)]
    assert [tuple(x) for x in t.tolist()] == expect


@pytest.mark.parametrize("d,r", COUNTED, ids=[f"{d['name']}-{r['partitioner']}{r['ncut']}" for d, r in COUNTED])
def test_sort_order_matches_reference(d, r):
    if d["graph"].get("n", 0) >= (1 << 22):
        pytest.skip("large graph covered by the count tests")
    g = graph_from_spec(d["graph"])
    L = B.Layout(g, np.array(r["cuts"], dtype=np.uint32))
    asc = L.tasks("asc")
    assert [list(x) for x in asc.tolist()] == r["order_asc"]
    desc = L.tasks("desc")
    assert sorted(map(tuple, desc.tolist())) == sorted(map(tuple, asc.tolist()))


def test_empty_band_layout():
    # cuts with an empty band: every band-of lookup must skip it
    g = B.Graph.rmat(9, 8, seed=4)
    cuts = np.array([0, 100, 100, 300, g.n + 1], dtype=np.uint32)
    L = B.Layout(g, cuts)
    assert L.tile(1, 1).nnz == 0 and L.tile(1, 1).height == 0
    assert np.isnan(L.tile(1, 1).avg_degree)
    reconstruct(L, g)


def test_bad_cuts_rejected():
    g = B.Graph.rmat(8, 8, seed=4)
    with pytest.raises(B.BbtcError):
        B.Layout(g, np.array([0, 50, 20, g.n + 1], dtype=np.uint32))
    with pytest.raises(B.BbtcError):
        B.Layout(g, np.array([0, 50, 100], dtype=np.uint32))  # does not cover all rows


def _stats_numpy(g, cuts, tasks):
    """Per-task rows, wedges, sum d_ac, bc probes, ac probes and the bitmap
    lines those meet, straight from the CSR and the cuts (a task's rows: u in
    band i with neighbours in band j and in band k)."""
    off, col = g.offsets().astype(np.int64), g.cols().astype(np.int64)
    band = np.searchsorted(cuts, np.arange(g.n), side="right") - 1
    rows_of = [col[off[u]:off[u + 1]] for u in range(g.n)]
    out = []
    for i, j, k in tasks.tolist():
        R = W = A = Pbc = Pac = Lac = 0
        width = min(cuts[k + 1], g.n + 1) - cuts[k]
        lines = ((((width + 31) // 32) + 3) // 4 * 4 + 31) // 32  # 16-byte bitmap rows, 128-byte lines
        for u in range(cuts[i], min(cuts[i + 1], g.n)):
            nb = rows_of[u]
            ab = nb[band[nb] == j]
            ac = nb[band[nb] == k]
            if len(ab) == 0 or len(ac) == 0:
                continue
            R += 1
            W += len(ab)
            A += len(ac)
            Pac += len(ab) * len(ac)
            Lac += len(ab) * min(len(ac), lines)
            for v in ab:
                nv = rows_of[v]
                Pbc += int(np.sum(band[nv] == k))
        out.append((R, W, A, Pbc, Pac, Lac))
    return np.array(out, dtype=np.uint64).reshape(-1, 6)


@pytest.mark.parametrize("p,kind", [(1, "uni"), (4, "pal"), (6, "n1d")])
def test_task_stats(p, kind):
    g = B.Graph.rmat(9, 12, seed=31)
    cuts = g.partition(kind, p)
    L = B.Layout(g, cuts)
    t = L.tasks("lex")
    st = L.task_stats(t)
    want = _stats_numpy(g, cuts, t)
    # tasks the reference skips (a tile of average degree 0) carry no work
    for q, x in enumerate(t.tolist()):
        tiles = [L.tile(x[0], x[1]), L.tile(x[1], x[2]), L.tile(x[0], x[2])]
        if any(ti.nnz == 0 for ti in tiles):
            want[q] = 0
    assert np.array_equal(st, want)
    # the algorithmic bytes are these terms: 4o R + 4 W + 4 A + 2o W + 4 Pbc
    for o in (4, 8):
        per, _, wedges = L.task_bytes(t, o)
        R, W, A, Pbc = (st[:, c].astype(np.int64) for c in range(4))
        assert np.array_equal(per.astype(np.int64), 4 * o * R + 4 * W + 4 * A + 2 * o * W + 4 * Pbc)
        assert wedges == int(W.sum())


def test_column_beyond_last_cut_is_refused():
    """A graph whose column count exceeds its rows may hold a column at or
    beyond cuts[P]; the layout refuses it instead of placing it in a tile that
    does not exist."""
    off = np.array([0, 2, 3, 3, 3], dtype=np.uint64)
    cols = np.array([1, 8, 2], dtype=np.uint32)
    g = B.Graph.from_csr(off, cols, n=4, ncols=10)
    with pytest.raises(B.BbtcError, match="beyond"):
        B.Layout(g, [0, 2, 5])
    L = B.Layout(g, [0, 2, 9])
    assert L.tile(0, 0).nnz == 1 and L.tile(0, 1).nnz == 2 and L.tile(1, 1).nnz == 0
    rp, cl, nz = L.arrays()
    assert list(nz) == [0, 0, 1]  # tile (0,0): row 0; tile (0,1): rows 0 and 1
