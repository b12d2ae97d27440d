"""The block layout built on the device (bbtc_graph_build) against the host
build (bbtc_layout_build + bbtc_graph_upload): the same arrays bit for bit,
the same per-tile fields, the same work statistics (computed on the device),
the same plans and the same per-task counts, which also equal the oracle and
the reference-written golden vectors."""
import numpy as np
import pytest

import bbtc_amd as B
import oracle as O
from tests.conftest import graph_from_spec, load_golden
from tests.test_count_cpu import CLOSED

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def dev():
    assert B.device_count() >= 1, "no HIP device"
    d = B.Device(0)
    yield d
    d.close()


def empty_band_cuts(n):
    # bands 1 and 3 are empty, the last band ends at n + 1 (the driver's convention)
    a, b = n // 3, 2 * n // 3
    return np.array([0, a, a, b, b, n + 1], dtype=np.uint32)


def graphs():
    out = [
        ("rmat14", B.Graph.rmat(14, 8, seed=1)),
        ("rmat16", B.Graph.rmat(16, 16, seed=2)),
        ("gnm", B.Graph.gnm(20000, 200000, 3)),
        ("empty", B.Graph.from_edges(100, np.zeros(0, np.uint32), np.zeros(0, np.uint32))),
    ]
    out += [(name, g) for name, g, _ in CLOSED[:3]]
    return out


CASES = []
for name, g in graphs():
    n = g.n
    for kind, p in (("uni", 1), ("uni", 3), ("pal", 8), ("uni", 16), ("empty-bands", 5)):
        CASES.append((name, g, kind, p))


def cuts_of(g, kind, p):
    if kind == "empty-bands":
        return empty_band_cuts(g.n)
    return g.partition(kind, p)


def same_layout(Lh, Ld, p):
    ah, ad = Lh.arrays(), Ld.arrays()
    for x, y, what in zip(ah, ad, ("rowptr", "cols", "nzrows")):
        assert x.shape == y.shape and np.array_equal(x, y), what
    for i in range(p):
        for k in range(p):
            th, td = Lh.tile(i, k), Ld.tile(i, k)
            assert th.nnz == td.nnz and th.nz_rows == td.nz_rows and th.max_row == td.max_row, (i, k)
            assert th.height == td.height and th.width == td.width
            assert (np.isnan(th.avg_degree) and np.isnan(td.avg_degree)) or th.avg_degree == td.avg_degree
    assert Lh.device_bytes == Ld.device_bytes
    for order in ("lex", "desc", "asc"):
        assert np.array_equal(Lh.tasks(order), Ld.tasks(order)), order


@pytest.mark.parametrize("name,g,kind,p", CASES, ids=[f"{c[0]}-{c[2]}{c[3]}" for c in CASES])
def test_device_build_equals_host_build(dev, name, g, kind, p):
    log("case", name, kind, p)
    cuts = cuts_of(g, kind, p)
    p = len(cuts) - 1
    Lh = B.Layout(g, cuts)
    before = dev.bytes_in_use
    dg = dev.build(g, cuts)
    same_layout(Lh, dg.layout, p)
    tasks = Lh.tasks("desc")
    # work statistics on the device equal the host's, exact and sampled
    for s in (0, 1, 16):
        assert np.array_equal(dg.task_stats(tasks, s), Lh.task_stats(tasks, s)), s
    # the same plans (AUTO forms from the same statistics) and counts
    up = dev.upload(Lh)
    assert up.bytes == dg.bytes
    for form in ("auto", "merge", "lane"):
        per_u, tot_u = up.count(tasks, form)
        per_d, tot_d = dg.count(tasks, form)
        assert tot_u == tot_d and np.array_equal(per_u, per_d), form
    pu, pd = up.plan(tasks), dg.plan(tasks)
    assert np.array_equal(pu.kinds(), pd.kinds())
    pu.destroy()
    pd.destroy()
    ref = O.RefLayout(g.offsets(), g.cols(), cuts)
    want, _ = ref.count(np.array(tasks.tolist(), dtype=np.uint32))
    assert np.array_equal(dg.count(tasks)[0], want)
    up.release()
    dg.release()
    assert dev.bytes_in_use == before


GOLDEN = [(d, r) for d in load_golden() for r in d["runs"] if "tasks" in r]


@pytest.mark.parametrize("d,r", GOLDEN, ids=[f"{d['name']}-{r['partitioner']}{r['ncut']}" for d, r in GOLDEN])
def test_device_build_golden(dev, d, r):
    """Per-task counts on layouts built on the device, against the vectors
    the reference wrote, without host arrays."""
    log("case", d["name"], r["partitioner"], r["ncut"])
    g = graph_from_spec(d["graph"])
    dg = dev.build(g, np.array(r["cuts"], dtype=np.uint32), host_arrays=False)
    ref = np.array(r["tasks"], dtype=np.uint64)
    per, tot = dg.count(ref[:, :3].astype(np.uint32))
    assert tot == r["total"]
    assert np.array_equal(per, ref[:, 3])
    dg.release()


def test_without_host_arrays(dev):
    g = B.Graph.rmat(13, 8, seed=5)
    cuts = g.partition("pal", 4)
    dg = dev.build(g, cuts, host_arrays=False)
    L = dg.layout
    assert L.arrays() is None
    tasks = L.tasks("desc")
    Lh = B.Layout(g, cuts)
    assert np.array_equal(tasks, Lh.tasks("desc"))
    per, tot = dg.count(tasks)
    assert np.array_equal(per, Lh.count_cpu(tasks)[0])
    # device-only --border works; whatever reads the host arrays is refused
    per_h, tot_h, gt = dg.count_hybrid(tasks, len(tasks))
    assert tot_h == tot and gt == len(tasks)
    for call in (lambda: L.task_bytes(tasks, 4), lambda: L.task_stats(tasks), lambda: L.count_cpu(tasks),
                 lambda: dev.upload(L), lambda: dg.count_hybrid(tasks, 0)):
        with pytest.raises(B.BbtcError, match="host"):
            call()
    dg.release()


def test_column_beyond_cuts_is_refused(dev):
    off = np.array([0, 2, 3, 3, 3], dtype=np.uint64)
    cols = np.array([1, 8, 2], dtype=np.uint32)
    g = B.Graph.from_csr(off, cols, n=4, ncols=10)
    before = dev.bytes_in_use
    with pytest.raises(B.BbtcError, match="beyond"):
        dev.build(g, [0, 2, 5])
    assert dev.bytes_in_use == before
    dg = dev.build(g, [0, 2, 9])
    assert dg.count(dg.layout.tasks())[1] == 0
    dg.release()


def test_device_build_large(dev):
    """RMAT scale 21 with 32 pal cuts (tiles of more than one scan chunk)."""
    g = B.Graph.rmat_device(dev, 21, 16, seed=7)
    cuts = g.partition("pal", 32)
    Lh = B.Layout(g, cuts)
    dg = dev.build(g, cuts)
    same_layout(Lh, dg.layout, 32)
    tasks = Lh.tasks("desc")
    assert np.array_equal(dg.task_stats(tasks, 16), Lh.task_stats(tasks, 16))
    per, tot = dg.count(tasks)
    up = dev.upload(Lh)
    assert np.array_equal(per, up.count(tasks)[0])
    up.release()
    dg.release()
