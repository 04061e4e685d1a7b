"""Admission by HBM footprint in the serving loop (pure host logic, no GPU): requests are taken in
arrival order while their footprint fits the budget and the batch limits; a request that does not
fit waits until memory frees up; a batch holds one size class."""
import threading
from collections import deque

import numpy as np


def _server(budget, max_docs=100, max_bytes=1 << 30, large=1 << 40):
    from alpacc_amd import serving
    s = serving.Server.__new__(serving.Server)
    s.budget = budget
    s.max_docs = max_docs
    s.max_bytes = max_bytes
    s.large = large
    s._pending = deque()
    s._inflight = 0
    s._cv = threading.Condition()
    return s


def _req(i, n):
    from alpacc_amd.serving import Request
    return Request(i, np.zeros(n, dtype=np.uint8))


def test_admit_in_order_within_budget():
    from alpacc_amd.serving import footprint
    s = _server(budget=footprint(1000) + footprint(2000) + 10)
    for i, n in enumerate((1000, 2000, 500, 10)):
        s._pending.append(_req(i, n))
    b = s._admit()
    assert [r.id for r in b] == [0, 1]          # 2 would exceed the budget
    assert s._inflight == footprint(1000) + footprint(2000)
    assert s._admit() == []                     # nothing fits until memory is released
    s._inflight = 0
    assert [r.id for r in s._admit()] == [2, 3]


def test_batch_limits():
    s = _server(budget=1 << 40, max_docs=3, max_bytes=2500)
    for i, n in enumerate((1000, 1000, 1000, 1000)):
        s._pending.append(_req(i, n))
    assert [r.id for r in s._admit()] == [0, 1]  # a third would pass max_bytes
    s2 = _server(budget=1 << 40, max_docs=3)
    for i in range(5):
        s2._pending.append(_req(i, 10))
    assert [r.id for r in s2._admit()] == [0, 1, 2]


def test_size_classes():
    """Small requests batch without the large ones between them, and the large ones keep their
    order for the next batch."""
    s = _server(budget=1 << 40, large=1000)
    for i, n in enumerate((10, 5000, 20, 3000, 30)):
        s._pending.append(_req(i, n))
    assert [r.id for r in s._admit()] == [0, 2, 4]
    assert [r.id for r in s._pending] == [1, 3]
    assert [r.id for r in s._admit()] == [1, 3]
    s2 = _server(budget=1 << 40, large=1000, max_bytes=6000)
    for i, n in enumerate((5000, 10, 3000, 20)):
        s2._pending.append(_req(i, n))
    assert [r.id for r in s2._admit()] == [0]   # 3000 more would pass max_bytes
    assert [r.id for r in s2._pending] == [1, 2, 3]


def test_workload_shapes():
    from alpacc_amd import serving
    sz = serving.request_sizes(10_000, seed=1)
    assert sz.min() >= 256 and sz.max() <= 64 << 20
    # log-uniform: the median sits near the geometric mean of the bounds
    assert abs(np.log(np.median(sz)) - 0.5 * (np.log(256) + np.log(64 << 20))) < 0.2
    arr = serving.poisson_arrivals(10_000, rate=100.0, seed=1)
    assert np.all(np.diff(arr) > 0) and abs(arr[-1] / 10_000 - 0.01) < 0.001


def test_parallel_copy():
    """The host copies of the serving loop and of Context's large transfers: small ones directly,
    large ones in parallel slices; every byte lands, at any offset and length."""
    from alpacc_amd import runtime
    rng = np.random.Generator(np.random.PCG64(5))
    src = rng.integers(0, 256, size=(runtime.COPY_SPLIT * 3) + 12345, dtype=np.uint8)
    for n in (0, 1, runtime.COPY_SPLIT, runtime.COPY_SPLIT + 1, src.shape[0]):
        dst = np.zeros(src.shape[0] + 7, dtype=np.uint8)
        runtime.copy_into(dst[3:3 + n], src[:n])
        assert np.array_equal(dst[3:3 + n], src[:n])
        assert not dst[:3].any() and not dst[3 + n:].any()
