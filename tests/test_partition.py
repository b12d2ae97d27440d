"""uni / n1d / pal cut vectors.

Pinned by golden vectors that the reference's own partitioner code wrote
(tests/golden, via oracle/_ref/ref_harness), by the closed form of uni, and
by agreement with the oracle's independent restatement on random graphs.
"""
import numpy as np
import pytest

import bbtc_amd as B
import oracle as O
from tests.conftest import graph_from_spec, load_golden
from tests.golden.fingerprint import fingerprint

GOLDEN = load_golden()
RUNS = [(d, r) for d in GOLDEN for r in d["runs"]]


@pytest.fixture(scope="module")
def graphs():
    cache = {}

    def get(d):
        if d["name"] not in cache:
            g = graph_from_spec(d["graph"])
            assert fingerprint(g.offsets(), g.cols()) == d["fingerprint"], "generator drifted"
            cache[d["name"]] = g
        return cache[d["name"]]
    return get


@pytest.mark.parametrize("d,r", RUNS, ids=[f"{d['name']}-{r['partitioner']}{r['ncut']}" for d, r in RUNS])
def test_cuts_match_reference(graphs, d, r):
    g = graphs(d)
    cuts = g.partition(r["partitioner"], r["ncut"])
    assert cuts.tolist() == r["cuts"]


@pytest.mark.parametrize("d,r", [x for x in RUNS if x[0]["graph"].get("n", 1 << 30) < 100000
                                 or x[0]["graph"].get("scale", 99) <= 16],
                         ids=lambda x: str(x)[:20] if not isinstance(x, dict) else x.get("name", ""))
def test_oracle_cuts_match_reference(graphs, d, r):
    g = graphs(d)
    c = O.cuts(r["partitioner"], g.offsets(), g.cols(), r["ncut"])
    assert c.tolist() == r["cuts"]


@pytest.mark.parametrize("n,p", [(10, 3), (65536, 4), (1000003, 7), (5, 8)])
def test_uni_closed_form(n, p):
    g = B.Graph.from_csr(np.zeros(n + 1, dtype=np.uint64), np.zeros(0, dtype=np.uint32))
    cuts = g.partition("uni", p)
    assert cuts.tolist() == [i * n // p for i in range(p)] + [n + 1]


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("kind", ["n1d", "pal"])
def test_product_agrees_with_oracle_random(seed, kind):
    rng = np.random.default_rng(seed)
    if seed % 2:
        g = B.Graph.rmat(int(rng.integers(6, 12)), int(rng.integers(4, 24)), seed=seed + 100)
    else:
        n = int(rng.integers(20, 3000))
        g = B.Graph.gnm(n, int(rng.integers(1, min(8 * n, n * (n - 1) // 4))), seed + 200)
    for p in (1, 2, 3, 5, 9):
        a = g.partition(kind, p)
        b = O.cuts(kind, g.offsets(), g.cols(), p)
        assert a.tolist() == b.tolist(), (kind, p)
        assert a[0] == 0 and a[-1] == g.n + 1 and np.all(np.diff(a.astype(np.int64)) >= 0)


def test_pal_balances_load():
    g = B.Graph.rmat(14, 16, seed=3)
    p = 8
    cuts = g.partition("pal", p).astype(np.int64)
    cuts[-1] = g.n
    off, col = g.offsets().astype(np.int64), g.cols().astype(np.int64)
    rows = np.repeat(np.arange(g.n), np.diff(off))
    bi = np.searchsorted(cuts, rows, side="right") - 1
    bk = np.searchsorted(cuts, col, side="right") - 1
    loads = np.zeros((p, p), dtype=np.int64)
    np.add.at(loads, (bi, bk), 1)
    ucut = g.partition("uni", p).astype(np.int64)
    ucut[-1] = g.n
    ub = np.searchsorted(ucut, rows, side="right") - 1
    uk = np.searchsorted(ucut, col, side="right") - 1
    uloads = np.zeros((p, p), dtype=np.int64)
    np.add.at(uloads, (ub, uk), 1)
    assert loads.max() < uloads.max()


@pytest.mark.parametrize("n", [3000, 3_000_000])
def test_rac_parallel_and_serial_transpose(n):
    # rac's transpose uses per-thread histograms unless threads x rows outgrow
    # the entries; a few thousand edges over 3M ids take the one-thread path.
    # Both give the cuts of the oracle's restatement.
    rng = np.random.default_rng(n)
    ids = np.sort(rng.choice(n, 400, replace=False)).astype(np.uint32)
    a, b = ids[rng.integers(0, 400, 3000)], ids[rng.integers(0, 400, 3000)]
    keep = a != b
    g = B.Graph.from_edges(n, a[keep], b[keep])
    for p in (2, 5, 16):
        got = g.partition("rac", p)
        assert got.tolist() == O.cuts("rac", g.offsets(), g.cols(), p).tolist(), p


@pytest.mark.slow
def test_s26_pal32_cuts_match_reference():
    # the headline graph (RMAT scale 26, edge factor 16, seed 3; about 2.5
    # minutes and 25 GB on 8 cores): the reference's own pal cuts with 32 cuts
    # and the ascending task order its tasks::SortTasks wrote
    (fx,) = load_golden("c3_rmat_s26_ef16")
    (run,) = fx["runs"]
    g = graph_from_spec(fx["graph"])
    assert fingerprint(g.offsets(), g.cols()) == fx["fingerprint"], "generator drifted"
    cuts = g.partition("pal", 32)
    assert cuts.tolist() == run["cuts"]
    L = B.Layout(g, cuts)
    assert [tuple(x) for x in L.tasks("asc").tolist()] == [tuple(x) for x in run["order_asc"]]
    assert len(run["tasks"]) == 5984 and sum(x[3] for x in run["tasks"]) == run["total"]
