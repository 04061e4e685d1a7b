"""Admission by HBM footprint in the serving loop (pure host logic, no GPU): requests are taken in
arrival order while the batch limits hold and the device memory fits the budget (every worker's
measured holdings, the growth reserved by running batches, the growth this batch needs); a request
that does not fit waits until memory frees up; a batch holds one size class."""
import threading
from collections import deque

import numpy as np


def _server(budget, max_docs=100, max_bytes=1 << 30, large=1 << 40, workers=1):
    from alpacc_amd import serving
    s = serving.Server.__new__(serving.Server)
    s.budget = budget
    s.max_docs = max_docs
    s.max_bytes = max_bytes
    s.large = large
    s._pending = deque()
    s._resident = [0] * workers
    s._sized_for = [0] * workers
    s._per_byte = float(serving.FOOTPRINT_PER_BYTE)
    s._inflight = 0
    s._cv = threading.Condition()
    return s


def _req(i, n):
    from alpacc_amd.serving import Request
    return Request(i, np.zeros(n, dtype=np.uint8))


def test_admit_in_order_within_budget():
    from alpacc_amd.serving import FOOTPRINT_PER_BYTE as F
    s = _server(budget=F * 3000 + 10)
    for i, n in enumerate((1000, 2000, 500, 10)):
        s._pending.append(_req(i, n))
    b, grow = s._admit()
    assert [r.id for r in b] == [0, 1]          # 2 would exceed the budget
    assert grow == F * 3000 and s._inflight == grow
    assert s._admit() == ([], 0)                # nothing fits until memory is released
    s._inflight = 0
    b, grow = s._admit()
    assert [r.id for r in b] == [2, 3] and grow == F * 510


def test_admission_by_measured_holdings():
    """A worker already sized for a batch needs no growth: its batch fits while the workers'
    measured holdings do; the per-byte rate becomes what the sized workers hold."""
    s = _server(budget=1000 << 20, workers=2)
    s._measured(0, 600 << 20, 10 << 20)         # worker 0 holds 600 MiB, sized for 10 MiB
    assert s._per_byte == 60.0
    for i, n in enumerate((4 << 20, 4 << 20, 4 << 20)):
        s._pending.append(_req(i, n))
    b, grow = s._admit(0)                       # 12 MiB: 2 MiB growth at 60 B/byte = 120 MiB
    assert [r.id for r in b] == [0, 1, 2] and grow == 120 << 20
    s._inflight = 0
    s._measured(1, 500 << 20, 0)                # worker 1 holds 500 MiB, sized for nothing yet
    for i, n in enumerate((8 << 20, 1 << 20)):
        s._pending.append(_req(10 + i, n))
    assert s._admit(1) == ([], 0)               # 1100 MiB held + 480 MiB growth > budget
    b, grow = s._admit(0)                       # worker 0: 9 MiB fit its size, no growth
    assert [r.id for r in b] == [10, 11] and grow == 0


def test_batch_limits():
    s = _server(budget=1 << 40, max_docs=3, max_bytes=2500)
    for i, n in enumerate((1000, 1000, 1000, 1000)):
        s._pending.append(_req(i, n))
    assert [r.id for r in s._admit()[0]] == [0, 1]  # a third would pass max_bytes
    s2 = _server(budget=1 << 40, max_docs=3)
    for i in range(5):
        s2._pending.append(_req(i, 10))
    assert [r.id for r in s2._admit()[0]] == [0, 1, 2]


def test_size_classes():
    """Small requests batch without the large ones between them, and the large ones keep their
    order for the next batch."""
    s = _server(budget=1 << 40, large=1000)
    for i, n in enumerate((10, 5000, 20, 3000, 30)):
        s._pending.append(_req(i, n))
    assert [r.id for r in s._admit()[0]] == [0, 2, 4]
    assert [r.id for r in s._pending] == [1, 3]
    assert [r.id for r in s._admit()[0]] == [1, 3]
    s2 = _server(budget=1 << 40, large=1000, max_bytes=6000)
    for i, n in enumerate((5000, 10, 3000, 20)):
        s2._pending.append(_req(i, n))
    assert [r.id for r in s2._admit()[0]] == [0]   # 3000 more would pass max_bytes
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


def test_transfer_slices_cover_every_part():
    """The slices of the pipelined transfers: in order, at most the requested count (when the
    parts allow), every slice but a part's last at least COPY_SPLIT bytes, and together exactly
    the parts' bytes at their offsets."""
    from alpacc_amd import runtime
    S = runtime.COPY_SPLIT
    cases = [[(1000, 0, 0)], [(1000, 5, 0)], [(1 << 40, 3 * S + 7, 0)],
             [(1 << 40, 4 * 10, 0), (1 << 41, 4 * 10, 40), (1 << 42, 8 * 10, 80)],
             [(1 << 40, 100 * S + 1, 0), (1 << 41, 50 * S, 100 * S + 16)]]
    for parts in cases:
        for nmax in (1, 4, 16, 64):
            sl = runtime._slices([(d, nb, off) for d, nb, off in parts], nmax)
            for d, nb, off in parts:
                mine = [x for x in sl if off <= x[2] < off + nb]
                assert sum(k for _, k, _ in mine) == nb
                pos = off
                for dev, k, o in mine:
                    assert o == pos and dev == d + (o - off) and k > 0
                    pos += k
                for _, k, _ in mine[:-1]:
                    assert k >= S
            assert len(sl) <= nmax + len(parts)
