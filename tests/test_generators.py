"""Seeded generators and the canonical-input pipeline.

Properties the canonical input must have (upper triangle, strictly
increasing rows, ranks ordered by (degree, id)), checked against an
independent numpy construction from the same edge list.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import bbtc_amd as B
from tests.golden.fingerprint import fingerprint


def canonical_numpy(n, src, dst):
    """Independent restatement: symmetrise, drop loops and duplicates,
    rank by (degree, id), orient to the lower rank, sort."""
    s = np.asarray(src, dtype=np.int64)
    d = np.asarray(dst, dtype=np.int64)
    keep = s != d
    lo = np.minimum(s, d)[keep]
    hi = np.maximum(s, d)[keep]
    key = np.unique(lo * n + hi)
    lo, hi = key // n, key % n
    deg = np.bincount(lo, minlength=n) + np.bincount(hi, minlength=n)
    order = np.lexsort((np.arange(n), deg))  # by degree, then id
    rank = np.empty(n, dtype=np.int64)
    rank[order] = np.arange(n)
    a, b = rank[lo], rank[hi]
    r, c = np.minimum(a, b), np.maximum(a, b)
    idx = np.lexsort((c, r))
    r, c = r[idx], c[idx]
    off = np.zeros(n + 1, dtype=np.uint64)
    np.add.at(off, r + 1, 1)
    return np.cumsum(off).astype(np.uint64), c.astype(np.uint32)


def check_canonical(g):
    off, col = g.offsets().astype(np.int64), g.cols().astype(np.int64)
    n = g.n
    rows = np.repeat(np.arange(n), np.diff(off))
    assert np.all(col > rows), "upper triangle"
    same_row = rows[1:] == rows[:-1]
    assert np.all(col[1:][same_row] > col[:-1][same_row]), "strictly increasing rows"
    deg = np.bincount(rows, minlength=n) + np.bincount(col, minlength=n)
    assert np.all(np.diff(deg) >= 0), "ranks ordered by degree"


def test_from_edges_matches_numpy():
    rng = np.random.default_rng(5)
    n = 300
    src = rng.integers(0, n, 5000)
    dst = rng.integers(0, n, 5000)
    g = B.Graph.from_edges(n, src, dst)
    off, col = canonical_numpy(n, src, dst)
    assert np.array_equal(g.offsets(), off)
    assert np.array_equal(g.cols(), col)
    check_canonical(g)


@pytest.mark.parametrize("scale,ef,seed", [(8, 8, 1), (12, 16, 3), (14, 4, 9)])
def test_rmat_is_canonical_and_deterministic(scale, ef, seed):
    g1 = B.Graph.rmat(scale, ef, seed=seed)
    g2 = B.Graph.rmat(scale, ef, seed=seed)
    assert g1.n == 1 << scale
    check_canonical(g1)
    assert fingerprint(g1.offsets(), g1.cols()) == fingerprint(g2.offsets(), g2.cols())
    assert g1.m < ef << scale


def test_rmat_seed_changes_graph():
    a, b = B.Graph.rmat(10, 8, seed=1), B.Graph.rmat(10, 8, seed=2)
    assert fingerprint(a.offsets(), a.cols()) != fingerprint(b.offsets(), b.cols())


def test_rmat_skew_follows_quadrants():
    # with a = 1 every sample is (0, 0): a self loop, so no edges survive
    assert B.Graph.rmat(8, 8, 1.0, 0.0, 0.0, 1).m == 0
    # uniform quadrants give an Erdos-Renyi-like graph with few duplicates
    g = B.Graph.rmat(12, 4, 0.25, 0.25, 0.25, 1)
    assert g.m > 0.95 * (4 << 12)


@pytest.mark.parametrize("n,m,seed", [(1000, 5000, 1), (50, 600, 2), (4096, 81920, 1007)])
def test_gnm_exact_edge_count(n, m, seed):
    g = B.Graph.gnm(n, m, seed)
    assert g.m == m and g.n == n
    check_canonical(g)


def test_gnm_rejects_impossible():
    with pytest.raises(B.BbtcError):
        B.Graph.gnm(10, 46, 1)


def test_thread_count_independence():
    code = ("import sys; sys.path.insert(0, %r); import bbtc_amd as B; "
            "from tests.golden.fingerprint import fingerprint; "
            "g = B.Graph.rmat(13, 8, seed=21); h = B.Graph.gnm(3000, 20000, 22); "
            "print(fingerprint(g.offsets(), g.cols()), fingerprint(h.offsets(), h.cols()))"
            % os.path.abspath(os.path.join(os.path.dirname(__file__), "..")))
    outs = []
    for t in ("1", "3", "8"):
        env = dict(os.environ, OMP_NUM_THREADS=t)
        outs.append(subprocess.run([sys.executable, "-c", code], env=env, check=True,
                                   capture_output=True, text=True).stdout)
    assert outs[0] == outs[1] == outs[2]
