"""bbtc_count_noalloc: the count without device allocation (a fixed
per-device scratch array), checked per task against bbtc_count on the same
inputs, bit-exact, on two calls in a row (the second with a smaller input,
so stale counters of the first would show), and for its limits."""
import numpy as np
import pytest

import bbtc_amd as B

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def dev():
    assert B.device_count() >= 1, "no HIP device"
    d = B.Device(0)
    yield d
    d.close()


def _case(g, part, p):
    L = B.Layout(g, g.partition(part, p))
    return L, L.tasks("desc")


def test_noalloc_matches_count_two_calls(dev):
    # first call: hub-heavy RMAT s16 with 2 bands (items split into spilled
    # chunks); second: a smaller graph with fewer tasks
    big = _case(B.Graph.rmat(16, 18, seed=1492), "pal", 2)
    small = _case(B.Graph.rmat(11, 8, seed=5), "uni", 3)
    for kernel in ("auto", "merge", "bitmap", "lane"):
        for name, (L, t) in (("big", big), ("small", small)):
            log("case noalloc", name, kernel, len(t))
            dg = dev.upload(L)
            want, wtot = dg.count(t, kernel)
            before = dev.bytes_in_use
            got, gtot = dg.count_noalloc(t, kernel)
            _, only_total = dg.count_noalloc(t, kernel, per_task=False)
            assert dev.bytes_in_use == before
            dg.release()
            assert gtot == wtot == only_total, (name, kernel)
            assert np.array_equal(got, want), (name, kernel)


def test_noalloc_many_tasks_and_empty(dev):
    g = B.Graph.gnm(20000, 200000, 3)
    L, t = _case(g, "pal", 40)  # 11,480 tasks
    dg = dev.upload(L)
    want, wtot = dg.count(t, "auto")
    got, gtot = dg.count_noalloc(t, "auto")
    assert gtot == wtot and np.array_equal(got, want)
    per, tot = dg.count_noalloc(np.zeros((0, 3), np.uint32), "auto")
    assert tot == 0 and len(per) == 0
    dg.release()


def test_noalloc_refuses_more_tasks_than_scratch(dev):
    g = B.Graph.rmat(10, 8, seed=2)
    L, t = _case(g, "uni", 73)  # C(75, 3) = 67,525 tasks > 65,536
    assert len(t) == 67525
    dg = dev.upload(L)
    with pytest.raises(B.BbtcError) as e:
        dg.count_noalloc(t, "auto")
    assert e.value.status == 1 and "65536" in str(e.value)
    # the first 65,536 tasks fit
    want, _ = dg.count(t[:65536], "auto")
    got, _ = dg.count_noalloc(t[:65536], "auto")
    assert np.array_equal(got, want)
    dg.release()
