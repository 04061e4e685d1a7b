"""Serving loop for lex+parse requests: continuous batching over device-resident tables.

The reference has no serving layer (its runtime is one entry call per document,
src/Alpacc/Generator/Futhark/Generator.hs:35-44). This loop keeps one grammar's tables on the GPU
and serves a stream of documents:

* every worker owns a context (tables in HBM, grow-only scratch arenas), a non-blocking stream and
  pinned staging buffers for the documents going in and the trees coming back (a device-to-host
  copy into pinned memory runs at DMA speed; the fresh per-request arrays are then filled by a
  pool of host threads);
* a worker that becomes free takes the requests waiting at that moment, in arrival order, as long
  as the batch limits hold and the device memory fits the shared budget (admission by footprint):
  every worker's measured holdings (context tables and arenas, its buffers), the growth reserved by
  running batches larger than their worker is sized for, and this batch's growth at the measured
  device bytes per input byte; it runs them as one ``alpacc_parse_batch`` call. A batch holds
  requests of one size class (below ``large_bytes``, or from it up): small requests are not held
  up by a large one's copies;
* with ``warmup`` (a document as large as the largest batch), every worker runs it once before
  serving, so that its device arenas, its pinned staging and its output buffers already have their
  final sizes: growing them while serving frees device memory, and a free waits for every stream
  of the device;
* requests that arrive while batches run wait for the next free worker, so several batches overlap
  on the device (one per worker stream) and arrivals are coalesced into the next launch;
* each request gets exactly what ``parse_int`` gives it alone: the batched entry processes every
  document on its own (tests/test_gpu_serving.py checks alone vs in the mix).

``serve_plain`` is the comparison: one ``alpacc_parse`` call per request in arrival order.
"""
from __future__ import annotations

import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from .runtime import (ACCEPT, Context, DeviceBuffer, HostBuffer, SliceEvents, load_library,
                      pipelined_download, pipelined_upload)

# Device bytes per input byte of a worker sized for a batch, before any batch has measured it:
# the input, the token slots (3 x u32 per byte), key positions and pair entries, brackets /
# productions with their scan values and minima, and the tree (u32 parent, node, span x2) at up to
# one node per byte. Workers replace it with what their allocations hold (Server._per_byte).
FOOTPRINT_PER_BYTE = 48


_CREATE_LOCK = threading.Lock()  # alpacc_create reports errors through one global string


@dataclass
class Request:
    id: int
    data: np.ndarray
    arrival: float = 0.0
    status: int = -99
    tree: tuple | None = None       # (parent, node, span[N, 2]) when accepted
    t_done: float = 0.0
    batch: int = -1                 # which launch served it
    # when its batch was taken, staged in pinned memory, parsed on the device (copy in included)
    # and copied back (perf_counter seconds)
    t_taken: float = 0.0
    t_staged: float = 0.0
    t_parsed: float = 0.0
    done: threading.Event = field(default_factory=threading.Event)

    @property
    def latency(self) -> float:
        return self.t_done - self.arrival


class Server:
    """Continuous batching of parse requests on one GPU."""

    def __init__(self, bundle, workers: int = 4, hbm_budget: float = 200e9,
                 max_batch_docs: int = 4096, max_batch_bytes: int = 256 << 20, device: int = 0,
                 large_bytes: int = 4 << 20, warmup=None):
        self.bundle = bundle
        self.budget = int(hbm_budget)
        self.max_docs = max_batch_docs
        self.max_bytes = max_batch_bytes
        self.large = large_bytes
        self._warm = None if warmup is None else np.ascontiguousarray(
            np.frombuffer(bytes(warmup), dtype=np.uint8) if isinstance(warmup, (bytes, bytearray))
            else warmup, dtype=np.uint8)
        self.device = device
        self._cv = threading.Condition()
        self._pending: deque = deque()
        # admission by measured device memory: what every worker holds (context tables and arenas,
        # its input and output buffers), the largest batch each is sized for, the device bytes
        # per input byte of a worker so sized (measured), and the growth reserved by running
        # batches that exceed their worker's size
        self._resident = [0] * workers
        self._sized_for = [0] * workers
        self._per_byte = float(FOOTPRINT_PER_BYTE)
        self._inflight = 0
        self._closing = False
        self._next_id = 0
        self._launches = 0
        self.batch_sizes: list = []
        self._threads = [threading.Thread(target=self._worker, args=(w,), daemon=True)
                         for w in range(workers)]
        self._ready = threading.Barrier(workers + 1)
        self._errors: list = []
        for t in self._threads:
            t.start()
        self._ready.wait()
        if self._errors:
            # the workers that did start wait for work: stop them, so that each frees its
            # context, stream and pinned buffers, before the error goes to the caller
            self.close()
            raise self._errors[0]

    # -- client side ----------------------------------------------------------------------------
    def submit(self, data, arrival: float | None = None) -> Request:
        arr = np.frombuffer(bytes(data), dtype=np.uint8) if isinstance(data, (bytes, bytearray)) \
            else np.ascontiguousarray(data, dtype=np.uint8)
        if self._per_byte * arr.shape[0] > self.budget:
            raise ValueError("request larger than the HBM budget")
        with self._cv:
            r = Request(self._next_id, arr, time.perf_counter() if arrival is None else arrival)
            self._next_id += 1
            self._pending.append(r)
            self._cv.notify()
        return r

    def close(self):
        with self._cv:
            self._closing = True
            self._cv.notify_all()
        for t in self._threads:
            t.join()

    # -- worker side ----------------------------------------------------------------------------
    def _growth(self, w: int, nbytes: int) -> int:
        """Device bytes worker w must add to run a batch of nbytes input bytes."""
        return int(max(0, nbytes - self._sized_for[w]) * self._per_byte)

    def _admit(self, w: int = 0) -> tuple[list, int]:
        """Requests waiting now, in arrival order, of the first one's size class, while the batch
        limits hold and the device memory fits the budget: every worker's measured holdings, the
        growth reserved by running batches, and the growth this batch needs on worker w. Returns
        the batch and the growth it reserves."""
        batch, nbytes, skipped = [], 0, []
        large = None
        held = sum(self._resident) + self._inflight
        while self._pending and len(batch) < self.max_docs:
            r = self._pending[0]
            n = r.data.shape[0]
            if large is not None and (n >= self.large) != large:
                skipped.append(self._pending.popleft())  # the other class: left for later
                continue
            # a batch within its worker's size allocates nothing and always fits
            g = self._growth(w, nbytes + n)
            over = g > 0 and held + g > self.budget
            if batch and (nbytes + n > self.max_bytes or over):
                break
            if not batch and over:
                break  # wait for memory to free up
            self._pending.popleft()
            batch.append(r)
            large = n >= self.large
            nbytes += n
        self._pending.extendleft(reversed(skipped))  # back in arrival order
        grow = self._growth(w, nbytes) if batch else 0
        self._inflight += grow
        return batch, grow

    def _measured(self, w: int, held: int, nbytes: int) -> None:
        """Worker w holds `held` device bytes after a batch of nbytes input bytes (lock held)."""
        self._resident[w] = held
        self._sized_for[w] = max(self._sized_for[w], nbytes)
        per = [self._resident[i] / self._sized_for[i] for i in range(len(self._resident))
               if self._sized_for[i] >= (1 << 20)]
        if per:
            self._per_byte = max(per)

    def _worker(self, w: int):
        ctx = stream = None
        state = {"in": None, "pin_in": None, "out": None, "pin_out": None, "ratio": 1.0,
                 "events": None}
        try:
            lib = load_library()
            with _CREATE_LOCK:
                ctx = Context(self.bundle, self.device)
            stream = lib.alpacc_stream_create()
            ctx.stream = stream
            state["events"] = SliceEvents(lib)
            if self._warm is not None:
                self._run(ctx, lib, stream, state, [Request(-1, self._warm)], -1)
            held = self._held(ctx, state)
            with self._cv:
                self._measured(w, held, 0 if self._warm is None else self._warm.shape[0])
        except Exception as exc:  # reported to the constructor
            self._errors.append(exc)
            self._release(ctx, stream, state)
            self._ready.wait()
            return
        self._ready.wait()
        while True:
            with self._cv:
                while not self._pending and not self._closing:
                    self._cv.wait()
                if not self._pending and self._closing:
                    break
                batch, grow = self._admit(w)
                if not batch:
                    self._cv.wait(0.001)
                    continue
                launch = self._launches
                self._launches += 1
                self.batch_sizes.append(len(batch))
            try:
                self._run(ctx, lib, stream, state, batch, launch)
            except Exception as exc:  # fail the batch's requests, keep serving
                self._errors.append(exc)
                now = time.perf_counter()
                for r in batch:
                    if not r.done.is_set():
                        r.status = -2
                        r.t_done = now
                        r.done.set()
            finally:
                held = self._held(ctx, state)
                with self._cv:
                    self._inflight -= grow
                    self._measured(w, held, sum(r.data.shape[0] for r in batch))
                    self._cv.notify_all()
        self._release(ctx, stream, state)

    @staticmethod
    def _held(ctx, state) -> int:
        """Device bytes a worker holds: its context's tables and arenas, its input and output
        buffers."""
        return ctx.device_bytes() + sum(b.nbytes for k in ("in", "out") for b in (state[k] or []))

    @staticmethod
    def _release(ctx, stream, state):
        """Free a worker's buffers, context and stream (whatever of them exists)."""
        for k in ("in", "out"):
            for b in (state[k] or []):
                b.free()
            state[k] = None
        for k in ("pin_in", "pin_out", "events"):
            if state[k] is not None:
                state[k].free()
            state[k] = None
        if ctx is not None:
            ctx.close()
        if stream is not None:
            load_library().alpacc_stream_destroy(stream)

    @staticmethod
    def _grow_dev(state, key, sizes, lib):
        bufs = state[key]
        if bufs is None or any(b.nbytes < s for b, s in zip(bufs, sizes)):
            for b in bufs or []:
                b.free()
            state[key] = [DeviceBuffer(max(s, 256), lib) for s in sizes]
        return state[key]

    @staticmethod
    def _grow_pin(state, key, size, lib):
        b = state[key]
        if b is None or b.nbytes < size:
            if b is not None:
                b.free()
            state[key] = HostBuffer(max(size, 256), lib)
        return state[key]

    def _run(self, ctx, lib, stream, state, batch, launch):
        t_taken = time.perf_counter()
        lengths = np.array([r.data.shape[0] for r in batch], dtype=np.uint64)
        offsets = np.zeros(len(batch), dtype=np.uint64)
        if len(batch) > 1:
            offsets[1:] = np.cumsum(lengths[:-1])
        total = int(lengths.sum())
        pin = self._grow_pin(state, "pin_in", total, lib)
        (d_in,) = self._grow_dev(state, "in", [total], lib)
        # the documents go in through pinned memory, each slice's DMA queued as soon as the host
        # has copied it there
        pipelined_upload(lib, stream, [(r.data, o) for r, o in zip(batch, offsets.tolist())],
                         pin, d_in.ptr)
        t_staged = time.perf_counter()
        cap = max(16, int(total * state["ratio"] * 1.25) + 16)
        while True:
            outs = self._grow_dev(state, "out", [4 * cap, 4 * cap, 8 * cap], lib)
            rc, nodes, res = ctx.parse_batch_device(d_in.ptr, offsets, lengths, outs[0].ptr,
                                                    outs[1].ptr, outs[2].ptr, cap, stream)
            if rc == -3:
                cap = int(nodes * 1.25) + 16
                continue
            break
        t_parsed = time.perf_counter()
        if total:
            state["ratio"] = max(state["ratio"] * 0.9, nodes / total)
        # the batch's trees come back through pinned memory into fresh arrays, in slices whose
        # DMA overlaps the host copies of the slices before; every request gets views of its range
        pin = self._grow_pin(state, "pin_out", 16 * nodes, lib)
        mem = np.empty(16 * nodes, dtype=np.uint8)
        if nodes:
            pipelined_download(lib, stream, [(outs[0].ptr, 4 * nodes, 0),
                                             (outs[1].ptr, 4 * nodes, 4 * nodes),
                                             (outs[2].ptr, 8 * nodes, 8 * nodes)],
                               pin, mem, state["events"])
        par = mem[:4 * nodes].view(np.uint32)
        nd = mem[4 * nodes:8 * nodes].view(np.uint32)
        sp = mem[8 * nodes:].view(np.uint32).reshape(-1, 2)
        now = time.perf_counter()
        for r, x in zip(batch, res):
            r.status = int(x["status"])
            if r.status == ACCEPT:
                o, c = int(x["offset"]), int(x["count"])
                r.tree = (par[o:o + c], nd[o:o + c], sp[o:o + c])
            r.batch = launch
            r.t_taken, r.t_staged, r.t_parsed = t_taken, t_staged, t_parsed
            r.t_done = now
            r.done.set()


def serve_plain(ctx: Context, requests, arrivals) -> None:
    """One alpacc_parse call per request, in arrival order (the comparison for the server)."""
    lib = ctx.lib
    t0 = time.perf_counter()
    for r, a in zip(requests, arrivals):
        target = t0 + a
        now = time.perf_counter()
        if target > now:
            time.sleep(target - now)
        r.arrival = target
        res = ctx.parse_int(r.data)
        r.status = ACCEPT if res is not None else 0
        r.tree = res
        r.t_done = time.perf_counter()
        r.done.set()
    lib.alpacc_stream_sync(None)


def serve_poisson(server: Server, datas, arrivals) -> list:
    """Submit every document at its arrival offset (seconds from now); wait for all of them."""
    reqs = []
    t0 = time.perf_counter()
    for d, a in zip(datas, arrivals):
        target = t0 + a
        now = time.perf_counter()
        if target > now:
            time.sleep(target - now)
        reqs.append(server.submit(d, arrival=target))
    for r in reqs:
        r.done.wait()
    return reqs


def poisson_arrivals(count: int, rate: float, seed: int = 42) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.cumsum(rng.exponential(1.0 / rate, size=count))


def request_sizes(count: int, lo: int = 256, hi: int = 64 << 20, seed: int = 42) -> np.ndarray:
    """Log-uniform sizes in [lo, hi]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size=count)).astype(np.int64)


def latency_stats(reqs) -> dict:
    lat = np.array([r.latency for r in reqs]) * 1e3
    span = max(r.t_done for r in reqs) - min(r.arrival for r in reqs)
    nbytes = sum(int(r.data.shape[0]) for r in reqs)
    out = {"p50_ms": round(float(np.percentile(lat, 50)), 3),
           "p99_ms": round(float(np.percentile(lat, 99)), 3),
           "mean_ms": round(float(lat.mean()), 3),
           "GBps": round(nbytes / span / 1e9, 3), "requests": len(reqs), "bytes": nbytes,
           "seconds": round(span, 3)}
    # where the slowest 1 % of requests spent their time (continuous batching only)
    tail = [r for r, x in zip(reqs, lat) if x >= np.percentile(lat, 99) and r.t_taken > 0]
    if tail:
        ph = np.array([[r.t_taken - r.arrival, r.t_staged - r.t_taken, r.t_parsed - r.t_staged,
                        r.t_done - r.t_parsed, r.data.shape[0] / 2**20] for r in tail])
        out["p99_tail"] = dict(zip(("wait_ms", "stage_ms", "parse_ms", "copy_back_ms", "MiB"),
                                   (round(float(x), 2) for x in ph.mean(axis=0) *
                                    np.array([1e3, 1e3, 1e3, 1e3, 1.0]))))
    return out
