"""Host-side mirror of the reference's generated entries, backed by the HIP C ABI.

The reference's Futhark artifact exposes (src/Alpacc/Generator/Futhark/Generator.hs:19-92):
  * ``lex_int``   : bytes -> (terminal ids, starts, ends), empty arrays when rejected;
  * ``lex``       : the same with terminal names (lexer.fut:180-191);
  * ``parse_int`` : bytes -> opt [(parent, #terminal t (i, j) | #production p)];
  * ``parse``     : the same with names (parser.fut:338-350);
  * ``format_lex`` / ``format_parse`` print those results in the README's form (README.md:84-106).
Here ``lex_int``/``parse_int`` return ``None`` for the reference's ``#none`` (rejection) instead of
empty arrays, because "accepted with zero tokens" and "rejected" must stay distinguishable.

The library is ``alpacc_amd/libalpacc_hip.so`` (built by ``alpacc_amd.build``); every call runs the
HIP kernels. There is no CPU fallback: a missing library or GPU raises.
"""
from __future__ import annotations

import ctypes
import os
import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from .gen.tables import Bundle

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libalpacc_hip.so")

ACCEPT, REJECT = 1, 0
ERRORS = {-1: "bad argument", -2: "HIP error", -3: "capacity", -4: "look-back timeout",
          -5: "bundle has no parser"}
NODE_TERMINAL = 0x80000000

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u16p = ctypes.POINTER(ctypes.c_uint16)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)


class AlpaccTables(ctypes.Structure):
    _fields_ = [
        ("num_endos", ctypes.c_uint32), ("num_classes", ctypes.c_uint32),
        ("num_states", ctypes.c_uint32), ("init_state", ctypes.c_uint32),
        ("ignore_terminal", ctypes.c_int32), ("dead_terminal", ctypes.c_uint32),
        ("byte_class", _u8p), ("endo_step", _u16p), ("endo_comp", _u16p),
        ("endo_apply", _u8p), ("dfa_step", _u16p), ("state_accept", _u8p),
        ("q", ctypes.c_uint32), ("k", ctypes.c_uint32), ("num_terminals", ctypes.c_uint32),
        ("start_terminal", ctypes.c_uint32), ("end_terminal", ctypes.c_uint32),
        ("empty_terminal", ctypes.c_uint32), ("num_productions", ctypes.c_uint32),
        ("arities", _i32p), ("prod_to_terminal", _i32p),
        ("pair_index_len", ctypes.c_uint64), ("pair_index", _u32p),
        ("num_pair_entries", ctypes.c_uint32), ("pair_entry", _u32p),
        ("stacks_len", ctypes.c_uint32), ("stacks", _u16p),
        ("productions_len", ctypes.c_uint32), ("productions", _u16p),
        ("pair_hash_len", ctypes.c_uint64), ("pair_hash_keys", ctypes.POINTER(ctypes.c_uint64)),
        ("pair_hash_vals", _u32p),
    ]


EXPORTS = ("alpacc_create", "alpacc_create_error", "alpacc_destroy", "alpacc_last_error",
           "alpacc_lex", "alpacc_parse", "alpacc_parse_tokens", "alpacc_last_launch_count",
           "alpacc_device_count", "alpacc_set_device", "alpacc_device_alloc", "alpacc_device_free",
           "alpacc_copy_to_device", "alpacc_copy_to_host", "alpacc_device_memset",
           "alpacc_stream_create", "alpacc_stream_destroy", "alpacc_stream_sync",
           "alpacc_event_create", "alpacc_event_destroy", "alpacc_event_record",
           "alpacc_event_elapsed_ms", "alpacc_profile_enable", "alpacc_profile_read",
           "alpacc_lex_chunked", "alpacc_parse_chunked", "alpacc_lex_batch", "alpacc_parse_batch",
           "alpacc_last_counts", "alpacc_host_alloc", "alpacc_host_free", "alpacc_copy_device",
           "alpacc_test", "alpacc_free", "alpacc_test_force_timeout", "alpacc_lex_noalloc",
           "alpacc_set_tuning", "alpacc_get_tuning", "alpacc_event_sync", "alpacc_device_bytes")

# the chunk size of the reference's generated entries (Generator.hs:24-76: lexer.lex_int 16777216)
REFERENCE_CHUNK = 16777216

_lib = None


def load_library():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"alpacc HIP library missing at {LIB_PATH}; build it with "
                           "`python __graft_entry__.py build`")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    u64 = ctypes.c_uint64
    L.alpacc_create.argtypes = [ctypes.POINTER(AlpaccTables), ctypes.c_int]
    L.alpacc_create.restype = vp
    L.alpacc_create_error.restype = ctypes.c_char_p
    L.alpacc_destroy.argtypes = [vp]
    L.alpacc_last_error.argtypes = [vp]
    L.alpacc_last_error.restype = ctypes.c_char_p
    L.alpacc_last_launch_count.argtypes = [vp]
    L.alpacc_test_force_timeout.argtypes = [vp, ctypes.c_int]
    L.alpacc_lex.argtypes = [vp, vp, u64, vp, vp, vp, u64, ctypes.POINTER(u64), vp]
    L.alpacc_lex_noalloc.argtypes = [vp, vp, u64, vp, vp, vp, u64, ctypes.POINTER(u64), vp]
    L.alpacc_parse.argtypes = [vp, vp, u64, vp, vp, vp, u64, ctypes.POINTER(u64), vp]
    L.alpacc_parse_tokens.argtypes = [vp, vp, vp, vp, u64, vp, vp, vp, u64, ctypes.POINTER(u64), vp]
    L.alpacc_lex_chunked.argtypes = [vp, vp, u64, u64, vp, vp, vp, u64, ctypes.POINTER(u64), vp]
    L.alpacc_parse_chunked.argtypes = [vp, vp, u64, u64, vp, vp, vp, u64, ctypes.POINTER(u64), vp]
    L.alpacc_last_counts.argtypes = [vp, vp]
    L.alpacc_device_bytes.argtypes = [vp]
    L.alpacc_device_bytes.restype = u64
    L.alpacc_set_tuning.argtypes = [vp, ctypes.POINTER(AlpaccTuning)]
    L.alpacc_get_tuning.argtypes = [vp, ctypes.POINTER(AlpaccTuning)]
    L.alpacc_copy_device.argtypes = [vp, vp, u64, vp]
    L.alpacc_test.argtypes = [vp, vp, u64, ctypes.POINTER(vp), ctypes.POINTER(u64), vp]
    L.alpacc_free.argtypes = [vp]
    L.alpacc_host_alloc.argtypes = [u64]
    L.alpacc_host_alloc.restype = vp
    L.alpacc_host_free.argtypes = [vp]
    L.alpacc_lex_batch.argtypes = [vp, vp, vp, vp, ctypes.c_uint32, vp, vp, vp, u64, vp, vp]
    L.alpacc_parse_batch.argtypes = [vp, vp, vp, vp, ctypes.c_uint32, vp, vp, vp, u64,
                                     ctypes.POINTER(u64), vp, vp]
    L.alpacc_device_count.restype = ctypes.c_int
    L.alpacc_set_device.argtypes = [ctypes.c_int]
    L.alpacc_device_alloc.argtypes = [u64]
    L.alpacc_device_alloc.restype = vp
    L.alpacc_device_free.argtypes = [vp]
    L.alpacc_copy_to_device.argtypes = [vp, vp, u64, vp]
    L.alpacc_copy_to_host.argtypes = [vp, vp, u64, vp]
    L.alpacc_device_memset.argtypes = [vp, ctypes.c_int, u64, vp]
    L.alpacc_stream_create.restype = vp
    L.alpacc_stream_destroy.argtypes = [vp]
    L.alpacc_stream_sync.argtypes = [vp]
    L.alpacc_event_create.restype = vp
    L.alpacc_event_destroy.argtypes = [vp]
    L.alpacc_event_record.argtypes = [vp, vp]
    L.alpacc_event_sync.argtypes = [vp]
    L.alpacc_event_elapsed_ms.argtypes = [vp, vp]
    L.alpacc_event_elapsed_ms.restype = ctypes.c_float
    L.alpacc_profile_enable.argtypes = [vp, ctypes.c_int]
    L.alpacc_profile_read.argtypes = [vp, ctypes.POINTER(ctypes.c_char_p),
                                      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(u64), ctypes.c_int]
    _lib = L
    return L


class AlpaccTuning(ctypes.Structure):
    """alpacc_tuning (include/alpacc_hip.h): a context's launch choices."""
    _fields_ = [("graph_mode", ctypes.c_int32), ("lex_blocks_per_cu", ctypes.c_uint32),
                ("keys_blocks_per_cu", ctypes.c_uint32), ("search_max_blocks", ctypes.c_uint32),
                ("brackets_per_kib", ctypes.c_uint32), ("search_block_threads", ctypes.c_uint32),
                ("lexer_tables_in_lds", ctypes.c_int32), ("reserved", ctypes.c_uint32),
                ("graph_launches", ctypes.c_uint64), ("graph_captures", ctypes.c_uint64)]


TUNING_KEYS = ("graph_mode", "lex_blocks_per_cu", "keys_blocks_per_cu", "search_max_blocks",
               "brackets_per_kib", "search_block_threads", "lexer_tables_in_lds")


# alpacc_doc_result: status, reserved, offset, count
DOC_RESULT = np.dtype([("status", "<i4"), ("reserved", "<u4"), ("offset", "<u8"), ("count", "<u8")])


class DeviceBuffer:
    """A device allocation owned by this library's HIP runtime."""

    def __init__(self, nbytes: int, lib=None):
        self.lib = lib or load_library()
        self.nbytes = int(nbytes)
        p = self.lib.alpacc_device_alloc(max(self.nbytes, 1))
        if not p:
            raise MemoryError(f"device allocation of {nbytes} bytes failed")
        self.ptr = p

    def upload(self, arr: np.ndarray, stream=None, offset: int = 0):
        arr = np.ascontiguousarray(arr)
        if arr.nbytes + offset > self.nbytes:
            raise ValueError("upload larger than the buffer")
        if self.lib.alpacc_copy_to_device(ctypes.c_void_p(self.ptr + offset), arr.ctypes.data,
                                          arr.nbytes, stream) != 0:
            raise RuntimeError("host to device copy failed")

    def download(self, dtype, count: int, stream=None, offset: int = 0) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if out.nbytes + offset > self.nbytes:
            raise ValueError("download larger than the buffer")
        if self.lib.alpacc_copy_to_host(out.ctypes.data, ctypes.c_void_p(self.ptr + offset),
                                        out.nbytes, stream) != 0:
            raise RuntimeError("device to host copy failed")
        return out

    def free(self):
        if self.ptr:
            self.lib.alpacc_device_free(ctypes.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


COPY_SPLIT = 16 << 20     # bytes per host copy task
PINNED_MIN = 1 << 20      # transfers from this size go through pinned staging
_COPY_POOL = None
_COPY_POOL_LOCK = threading.Lock()


def copy_into(dst: np.ndarray, src: np.ndarray) -> None:
    """dst[:] = src for uint8 arrays; large ones in parallel slices on a pool of 8 threads (numpy
    drops the GIL): one thread writing fresh pages moves about 9 GB/s, and the tree of a 64 MiB
    JSON document is about 250 MB."""
    global _COPY_POOL
    n = src.shape[0]
    if n <= COPY_SPLIT:
        np.copyto(dst, src)
        return
    with _COPY_POOL_LOCK:
        if _COPY_POOL is None:
            _COPY_POOL = ThreadPoolExecutor(max_workers=8)
    cuts = list(range(0, n, COPY_SPLIT)) + [n]
    list(_COPY_POOL.map(lambda ab: np.copyto(dst[ab[0]:ab[1]], src[ab[0]:ab[1]]),
                        zip(cuts[:-1], cuts[1:])))


def _copy_pool() -> ThreadPoolExecutor:
    global _COPY_POOL
    with _COPY_POOL_LOCK:
        if _COPY_POOL is None:
            _COPY_POOL = ThreadPoolExecutor(max_workers=8)
    return _COPY_POOL


class SliceEvents:
    """HIP events for the slices of one pipelined transfer (created once, reused)."""

    def __init__(self, lib, count: int = 64):
        self.lib = lib
        self.events = [lib.alpacc_event_create() for _ in range(count)]
        if not all(self.events):
            self.free()
            raise RuntimeError("event creation failed")

    def free(self):
        for e in self.events or []:
            if e:
                self.lib.alpacc_event_destroy(ctypes.c_void_p(e))
        self.events = []


def _slices(parts, nslices_max: int):
    """(device address, bytes, host offset) parts cut into at most nslices_max slices of at least
    COPY_SPLIT bytes, in order."""
    total = sum(nb for _, nb, _ in parts)
    size = max(COPY_SPLIT, -(-total // max(1, nslices_max)))
    size = (size + 4095) & ~4095
    out = []
    for dev, nb, hoff in parts:
        o = 0
        while o < nb:
            k = min(size, nb - o)
            out.append((dev + o, k, hoff + o))
            o += k
    return out


def pipelined_download(lib, stream, parts, pin, dst: np.ndarray, events: SliceEvents) -> None:
    """Device ranges into the uint8 host array dst through the pinned buffer pin (same offsets):
    every slice is DMA'd into pin and its event recorded at once; the host copies each slice out
    of pin as soon as its event completes, so the DMA of later slices overlaps the host copies of
    earlier ones. parts: (device address, bytes, offset in dst)."""
    # at most (slice count asked for) + (number of parts) slices: one event each
    sl = _slices(parts, max(1, len(events.events) - len(parts)))
    if len(sl) > len(events.events):
        raise ValueError("too many parts for the transfer's events")
    for i, (dev, nb, off) in enumerate(sl):
        if lib.alpacc_copy_to_host(ctypes.c_void_p(pin.ptr + off), ctypes.c_void_p(dev), nb,
                                   stream) != 0 or \
                lib.alpacc_event_record(ctypes.c_void_p(events.events[i]), stream) != 0:
            raise RuntimeError("device to host copy failed")
    pool = _copy_pool()
    futs = []
    for i, (_, nb, off) in enumerate(sl):
        if lib.alpacc_event_sync(ctypes.c_void_p(events.events[i])) != 0:
            raise RuntimeError("device to host copy failed")
        futs.append(pool.submit(np.copyto, dst[off:off + nb], pin.array[off:off + nb]))
    for f in futs:
        f.result()


def pipelined_upload(lib, stream, srcs, pin, dev: int) -> None:
    """uint8 host arrays to the device buffer at address dev through the pinned buffer pin (same
    offsets): a pool of threads copies slices into pin, and each slice's DMA is queued as soon as
    it is there. srcs: an array (at offset 0) or a list of (array, offset)."""
    if isinstance(srcs, np.ndarray):
        srcs = [(srcs, 0)]
    sl = []
    for src, base in srcs:
        for _, nb, off in _slices([(0, int(src.shape[0]), 0)], 64):
            sl.append((src[off:off + nb], base + off, nb))
    pool = _copy_pool()
    futs = [pool.submit(np.copyto, pin.array[o:o + nb], part) for part, o, nb in sl]
    for (_, o, nb), f in zip(sl, futs):
        f.result()
        if nb and lib.alpacc_copy_to_device(ctypes.c_void_p(dev + o), ctypes.c_void_p(pin.ptr + o),
                                            nb, stream) != 0:
            raise RuntimeError("host to device copy failed")


class HostBuffer:
    """Pinned host memory (for asynchronous copies), viewed as a numpy uint8 array."""

    def __init__(self, nbytes: int, lib=None):
        self.lib = lib or load_library()
        self.nbytes = int(nbytes)
        p = self.lib.alpacc_host_alloc(max(self.nbytes, 1))
        if not p:
            raise MemoryError(f"pinned host allocation of {nbytes} bytes failed")
        self.ptr = p
        self.array = np.ctypeslib.as_array((ctypes.c_uint8 * max(self.nbytes, 1)).from_address(p))

    def free(self):
        if self.ptr:
            self.array = None
            self.lib.alpacc_host_free(ctypes.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_count() -> int:
    return load_library().alpacc_device_count()


def _ptr(arr: np.ndarray, typ):
    return arr.ctypes.data_as(typ)


def fill_tables(bundle: Bundle) -> tuple[AlpaccTables, list]:
    """The C struct of a bundle plus the numpy arrays it points into (keep them alive)."""
    t = AlpaccTables()
    keep = []

    def hold(a, dtype):
        a = np.ascontiguousarray(a, dtype=dtype)
        keep.append(a)
        return a

    lx = bundle.lexer
    if lx is not None:
        t.num_endos = lx.endo.num_endos
        t.num_classes = lx.num_classes
        t.num_states = lx.num_states
        t.init_state = lx.init_state
        t.ignore_terminal = lx.ignore
        t.dead_terminal = lx.dead_terminal
        t.byte_class = _ptr(hold(lx.byte_class, np.uint8), _u8p)
        t.endo_step = _ptr(hold(lx.endo_step, np.uint16), _u16p)
        t.endo_comp = _ptr(hold(lx.endo_comp, np.uint16), _u16p)
        t.endo_apply = _ptr(hold(lx.endo_apply, np.uint8), _u8p)
        t.dfa_step = _ptr(hold(lx.dfa_step, np.uint16), _u16p)
        t.state_accept = _ptr(hold(lx.state_info, np.uint8), _u8p)
    px = bundle.parser
    if px is not None:
        t.q, t.k = px.q, px.k
        t.num_terminals = px.num_terminals
        t.start_terminal = px.start_terminal
        t.end_terminal = px.end_terminal
        t.empty_terminal = px.empty_terminal
        t.num_productions = len(px.production_names)
        t.arities = _ptr(hold(px.arities, np.int32), _i32p)
        t.prod_to_terminal = _ptr(hold(px.production_to_terminal, np.int32), _i32p)
        t.pair_index_len = px.pair_index.shape[0]
        t.pair_index = _ptr(hold(px.pair_index, np.uint32), _u32p) if px.pair_index.shape[0] \
            else ctypes.cast(None, _u32p)
        t.pair_hash_len = px.pair_hash_keys.shape[0]
        if px.pair_hash_keys.shape[0]:
            t.pair_hash_keys = _ptr(hold(px.pair_hash_keys, np.uint64),
                                    ctypes.POINTER(ctypes.c_uint64))
            t.pair_hash_vals = _ptr(hold(px.pair_hash_vals, np.uint32), _u32p)
        t.num_pair_entries = px.pair_entry.shape[0]
        t.pair_entry = _ptr(hold(px.pair_entry.reshape(-1), np.uint32), _u32p)
        t.stacks_len = px.stacks.shape[0]
        t.stacks = _ptr(hold(px.stacks, np.uint16), _u16p)
        t.productions_len = px.productions.shape[0]
        t.productions = _ptr(hold(px.productions, np.uint16), _u16p)
    return t, keep


class Context:
    """A device-resident table bundle (``alpacc_create``) and the four reference entries."""

    def __init__(self, bundle: Bundle, device: int = 0):
        self.bundle = bundle
        self.lib = load_library()
        if self.lib.alpacc_device_count() <= device:
            raise RuntimeError("no ROCm device available for the alpacc HIP backend")
        self.device = device
        self._tables, self._keep = fill_tables(bundle)
        h = self.lib.alpacc_create(ctypes.byref(self._tables), device)
        if not h:
            raise RuntimeError("alpacc_create failed: " + self.lib.alpacc_create_error().decode())
        self.handle = h
        self.stream = None          # default stream
        self._bufs: dict = {}
        self._pins: dict = {}  # grow-only pinned staging for large transfers
        self._events = None    # events of pipelined downloads

    def close(self):
        for b in list(self._bufs.values()) + list(self._pins.values()):
            b.free()
        self._bufs.clear()
        self._pins.clear()
        if getattr(self, "_events", None) is not None:
            self._events.free()
            self._events = None
        if getattr(self, "handle", None):
            self.lib.alpacc_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- helpers ---------------------------------------------------------------------------
    def _err(self, rc):
        msg = self.lib.alpacc_last_error(self.handle).decode()
        return RuntimeError(f"alpacc error {rc} ({ERRORS.get(rc, '?')}): {msg}")

    def buffer(self, name: str, nbytes: int) -> DeviceBuffer:
        b = self._bufs.get(name)
        if b is None or b.nbytes < nbytes:
            if b is not None:
                b.free()
            b = DeviceBuffer(max(nbytes, 256), self.lib)
            self._bufs[name] = b
        return b

    def pinned(self, name: str, nbytes: int) -> HostBuffer:
        b = self._pins.get(name)
        if b is None or b.nbytes < nbytes:
            if b is not None:
                b.free()
            b = HostBuffer(max(nbytes, 256), self.lib)
            self._pins[name] = b
        return b

    def upload_doc(self, data) -> tuple[DeviceBuffer, int]:
        if isinstance(data, (bytes, bytearray, memoryview)):
            arr = np.frombuffer(bytes(data), dtype=np.uint8)
        else:
            arr = np.ascontiguousarray(data, dtype=np.uint8)
        n = int(arr.shape[0])
        buf = self.buffer("doc", n)
        if n >= PINNED_MIN:  # through pinned memory: the device copy runs at DMA speed
            pipelined_upload(self.lib, self.stream, arr, self.pinned("doc", n), buf.ptr)
        elif n:
            buf.upload(arr, self.stream)
        return buf, n

    def download_arrays(self, parts) -> list:
        """[(DeviceBuffer, dtype, count), ...] -> host arrays, one fresh allocation for all of them;
        from PINNED_MIN bytes up through pinned staging (DMA) and parallel host copies."""
        sizes = [np.dtype(dt).itemsize * int(c) for _, dt, c in parts]
        offs = [0]
        for size in sizes:
            offs.append(offs[-1] + ((size + 15) & ~15))  # every part 16-byte aligned
        total = offs[-1]
        if sum(sizes) < PINNED_MIN:
            return [b.download(dt, int(c), self.stream) for b, dt, c in parts]
        pin = self.pinned("out", total)
        if self._events is None:
            self._events = SliceEvents(self.lib)
        mem = np.empty(total, dtype=np.uint8)
        pipelined_download(self.lib, self.stream,
                           [(b.ptr, size, off) for (b, _, _), size, off in zip(parts, sizes, offs)
                            if size], pin, mem, self._events)
        return [mem[off:off + size].view(dt) for (_, dt, _), size, off in zip(parts, sizes, offs)]

    def download_tree(self, pb: DeviceBuffer, nb: DeviceBuffer, spb: DeviceBuffer, nn: int):
        """(parent, node, span[N, 2]) u32 arrays of N = nn nodes from the device buffers."""
        par, node, span = self.download_arrays(((pb, np.uint32, nn), (nb, np.uint32, nn),
                                                (spb, np.uint32, 2 * nn)))
        return par, node, span.reshape(-1, 2)

    # -- launch choices ----------------------------------------------------------------------
    def tuning(self) -> dict:
        """The context's launch choices and graph counters (alpacc_get_tuning)."""
        t = AlpaccTuning()
        if self.lib.alpacc_get_tuning(self.handle, ctypes.byref(t)) != 0:
            raise self._err(-1)
        return {name: int(getattr(t, name)) for name, _ in AlpaccTuning._fields_}

    def set_tuning(self, **kw) -> dict:
        """Change some launch choices (alpacc_set_tuning); returns the new ones."""
        cur = self.tuning()
        for k, v in kw.items():
            if k not in TUNING_KEYS:
                raise KeyError(k)
            cur[k] = int(v)
        t = AlpaccTuning(*(cur[name] for name, _ in AlpaccTuning._fields_))
        rc = self.lib.alpacc_set_tuning(self.handle, ctypes.byref(t))
        if rc != 0:
            raise self._err(rc)
        return self.tuning()

    # -- profiling ---------------------------------------------------------------------------
    def profile(self, on: bool = True):
        """Record HIP events around every stage of later calls (on their stream)."""
        if self.lib.alpacc_profile_enable(self.handle, 1 if on else 0) != 0:
            raise self._err(-2)

    def read_profile(self) -> dict:
        """{stage: (total ms, calls)} accumulated since the last read."""
        m = 16
        names = (ctypes.c_char_p * m)()
        ms = (ctypes.c_double * m)()
        calls = (ctypes.c_uint64 * m)()
        k = self.lib.alpacc_profile_read(self.handle, names, ms, calls, m)
        return {names[i].decode(): (ms[i], calls[i]) for i in range(k)}

    def device_bytes(self) -> int:
        """Device memory of the context (tables and scratch arenas) plus its Python-side buffers."""
        return int(self.lib.alpacc_device_bytes(self.handle)) + sum(
            b.nbytes for b in self._bufs.values())

    def last_counts(self) -> dict:
        """Element counts of the last parser call: tokens, positions, brackets, productions."""
        out = np.zeros(4, dtype=np.uint64)
        self.lib.alpacc_last_counts(self.handle, out.ctypes.data)
        return dict(zip(("tokens", "positions", "brackets", "productions"), (int(x) for x in out)))

    # -- device-level calls -----------------------------------------------------------------
    def lex_device(self, d_in: int, n: int, kinds: int, starts: int, ends: int, cap: int,
                   stream=None) -> tuple[int, int]:
        """``alpacc_lex`` on device pointers; returns (status, token count)."""
        nt = ctypes.c_uint64(0)
        rc = self.lib.alpacc_lex(self.handle, ctypes.c_void_p(d_in), n, ctypes.c_void_p(kinds),
                                 ctypes.c_void_p(starts), ctypes.c_void_p(ends), cap,
                                 ctypes.byref(nt), stream)
        if rc < 0:
            raise self._err(rc)
        return rc, nt.value

    def parse_device(self, d_in: int, n: int, parent: int, node: int, span: int, cap: int,
                     stream=None) -> tuple[int, int]:
        """``alpacc_parse`` on device pointers; returns (status, node count); status -3 means
        the capacity was too small and the count is the size needed."""
        nn = ctypes.c_uint64(0)
        rc = self.lib.alpacc_parse(self.handle, ctypes.c_void_p(d_in), n, ctypes.c_void_p(parent),
                                   ctypes.c_void_p(node), ctypes.c_void_p(span), cap,
                                   ctypes.byref(nn), stream)
        if rc == -3:
            return rc, nn.value
        if rc < 0:
            raise self._err(rc)
        return rc, nn.value

    def lex_chunked_device(self, d_in: int, n: int, chunk: int, kinds: int, starts: int, ends: int,
                           cap: int, stream=None) -> tuple[int, int]:
        """``alpacc_lex_chunked`` on device pointers; returns (status, token count)."""
        nt = ctypes.c_uint64(0)
        rc = self.lib.alpacc_lex_chunked(self.handle, ctypes.c_void_p(d_in), n, chunk,
                                         ctypes.c_void_p(kinds), ctypes.c_void_p(starts),
                                         ctypes.c_void_p(ends), cap, ctypes.byref(nt), stream)
        if rc < 0:
            raise self._err(rc)
        return rc, nt.value

    # -- reference entries ------------------------------------------------------------------
    def lex_int(self, data, chunk_size: int | None = None):
        """``lex_int``: None when rejected, else (kinds, starts, ends) numpy arrays.

        ``chunk_size=None``: unchunked semantics, uint32 spans. ``chunk_size=c``: the reference's
        chunk loop over chunks of c bytes (``REFERENCE_CHUNK`` is what its entries use), uint64
        spans.
        """
        if chunk_size is not None:
            return self._lex_int_chunked(data, int(chunk_size))
        d_in, n = self.upload_doc(data)
        cap = max(n, 1)
        kb = self.buffer("kinds", 4 * cap)
        sb = self.buffer("starts", 4 * cap)
        eb = self.buffer("ends", 4 * cap)
        rc, nt = self.lex_device(d_in.ptr, n, kb.ptr, sb.ptr, eb.ptr, cap, self.stream)
        if rc != ACCEPT:
            return None
        return tuple(self.download_arrays(((kb, np.uint32, nt), (sb, np.uint32, nt),
                                           (eb, np.uint32, nt))))

    def _lex_int_chunked(self, data, chunk: int):
        d_in, n = self.upload_doc(data)
        c = min(chunk, max(n, 1))
        cap = n + (n + c - 1) // c + 1
        kb = self.buffer("kinds", 4 * cap)
        sb = self.buffer("starts64", 8 * cap)
        eb = self.buffer("ends64", 8 * cap)
        rc, nt = self.lex_chunked_device(d_in.ptr, n, chunk, kb.ptr, sb.ptr, eb.ptr, cap, self.stream)
        if rc != ACCEPT:
            return None
        return tuple(self.download_arrays(((kb, np.uint32, nt), (sb, np.uint64, nt),
                                           (eb, np.uint64, nt))))

    def lex(self, data, chunk_size: int | None = None):
        r = self.lex_int(data, chunk_size)
        if r is None:
            return None
        names = self.bundle.terminal_names
        return [(names[k], (int(s), int(e))) for k, s, e in zip(*r)]

    def parse_int(self, data, cap: int | None = None, chunk_size: int | None = None):
        """``parse_int``: None when rejected, else (parent, node, span[N,2]) numpy arrays.

        ``node`` has bit 31 set for terminal nodes (low bits = terminal id), else a production id.
        ``chunk_size=None``: unchunked lexer, uint32 spans; ``chunk_size=c``: the reference's
        chunked lexer (``REFERENCE_CHUNK`` for its generated ``parse_int``), uint64 spans.
        """
        if chunk_size is not None:
            return self._parse_int_chunked(data, cap, int(chunk_size))
        d_in, n = self.upload_doc(data)
        if cap is None:
            cap = max(16, 4 * n + 16)
        while True:
            pb = self.buffer("parent", 4 * cap)
            nb = self.buffer("node", 4 * cap)
            spb = self.buffer("span", 8 * cap)
            rc, nn = self.parse_device(d_in.ptr, n, pb.ptr, nb.ptr, spb.ptr, cap, self.stream)
            if rc == -3:
                cap = nn
                continue
            break
        if rc != ACCEPT:
            return None
        return self.download_tree(pb, nb, spb, nn)

    def parse_tokens(self, kinds, starts, ends, cap: int | None = None):
        """``parser.parse_int`` (parser.fut:334-336) on an existing token stream: None when
        rejected, else (parent, node, span[N,2]) as ``parse_int`` returns them."""
        kinds = np.ascontiguousarray(kinds, dtype=np.uint32)
        starts = np.ascontiguousarray(starts, dtype=np.uint32)
        ends = np.ascontiguousarray(ends, dtype=np.uint32)
        t = int(kinds.shape[0])
        if starts.shape[0] != t or ends.shape[0] != t:
            raise ValueError("kinds, starts and ends differ in length")
        toks = self.buffer("tokens", 12 * t + 12)
        for i, arr in enumerate((kinds, starts, ends)):
            if t:
                toks.upload(arr, self.stream, offset=4 * t * i)
        if cap is None:
            cap = max(16, 8 * t + 16)
        while True:
            pb = self.buffer("parent", 4 * cap)
            nb = self.buffer("node", 4 * cap)
            spb = self.buffer("span", 8 * cap)
            nn = ctypes.c_uint64(0)
            rc = self.lib.alpacc_parse_tokens(self.handle, ctypes.c_void_p(toks.ptr),
                                              ctypes.c_void_p(toks.ptr + 4 * t),
                                              ctypes.c_void_p(toks.ptr + 8 * t), t,
                                              ctypes.c_void_p(pb.ptr), ctypes.c_void_p(nb.ptr),
                                              ctypes.c_void_p(spb.ptr), cap, ctypes.byref(nn),
                                              self.stream)
            if rc == -3:
                cap = nn.value
                continue
            if rc < 0:
                raise self._err(rc)
            break
        if rc != ACCEPT:
            return None
        nn = nn.value
        return (pb.download(np.uint32, nn, self.stream), nb.download(np.uint32, nn, self.stream),
                spb.download(np.uint32, 2 * nn, self.stream).reshape(-1, 2))

    def _parse_int_chunked(self, data, cap, chunk: int):
        d_in, n = self.upload_doc(data)
        if cap is None:
            cap = max(16, 4 * n + 16)
        while True:
            pb = self.buffer("parent", 4 * cap)
            nb = self.buffer("node", 4 * cap)
            spb = self.buffer("span64", 16 * cap)
            nn = ctypes.c_uint64(0)
            rc = self.lib.alpacc_parse_chunked(self.handle, ctypes.c_void_p(d_in.ptr), n, chunk,
                                               ctypes.c_void_p(pb.ptr), ctypes.c_void_p(nb.ptr),
                                               ctypes.c_void_p(spb.ptr), cap, ctypes.byref(nn),
                                               self.stream)
            if rc == -3:
                cap = nn.value
                continue
            if rc < 0:
                raise self._err(rc)
            break
        if rc != ACCEPT:
            return None
        nn = nn.value
        return (pb.download(np.uint32, nn, self.stream), nb.download(np.uint32, nn, self.stream),
                spb.download(np.uint64, 2 * nn, self.stream).reshape(-1, 2))

    def test(self, blob: bytes) -> bytes:
        """The artifact's ``test`` entry (futhark/test.fut): wire-format inputs to wire-format
        outputs (lexer, lexer+parser or parser-only form, from the bundle)."""
        blob = bytes(blob)
        buf = ctypes.create_string_buffer(blob, len(blob))
        out = ctypes.c_void_p()
        n = ctypes.c_uint64(0)
        rc = self.lib.alpacc_test(self.handle, buf, len(blob), ctypes.byref(out), ctypes.byref(n),
                                  self.stream)
        if rc < 0:
            raise self._err(rc)
        try:
            return ctypes.string_at(out, n.value)
        finally:
            self.lib.alpacc_free(out)

    # -- batched entries ---------------------------------------------------------------------
    def upload_batch(self, docs) -> tuple[DeviceBuffer, np.ndarray, np.ndarray]:
        """Concatenate documents into one device buffer; returns (buffer, offsets, lengths)."""
        lengths = np.array([len(d) for d in docs], dtype=np.uint64)
        offsets = np.zeros(len(docs), dtype=np.uint64)
        if len(docs) > 1:
            offsets[1:] = np.cumsum(lengths[:-1])
        blob = np.frombuffer(b"".join(bytes(d) for d in docs), dtype=np.uint8)
        buf = self.buffer("batch", max(int(blob.shape[0]), 1))
        if blob.shape[0]:
            buf.upload(blob, self.stream)
        return buf, offsets, lengths

    def lex_batch_device(self, d_in: int, offsets: np.ndarray, lengths: np.ndarray, kinds: int,
                         starts: int, ends: int, cap: int, stream=None) -> np.ndarray:
        """``alpacc_lex_batch`` on device pointers; returns the per-document results."""
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        lengths = np.ascontiguousarray(lengths, dtype=np.uint64)
        res = np.zeros(len(lengths), dtype=DOC_RESULT)
        rc = self.lib.alpacc_lex_batch(self.handle, ctypes.c_void_p(d_in), offsets.ctypes.data,
                                       lengths.ctypes.data, len(lengths), ctypes.c_void_p(kinds),
                                       ctypes.c_void_p(starts), ctypes.c_void_p(ends), cap,
                                       res.ctypes.data, stream)
        if rc < 0:
            raise self._err(rc)
        return res

    def parse_batch_device(self, d_in: int, offsets: np.ndarray, lengths: np.ndarray, parent: int,
                           node: int, span: int, cap: int, stream=None):
        """``alpacc_parse_batch`` on device pointers; returns (status, total nodes, results);
        status -3 means the capacity was too small and total nodes is the size needed."""
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        lengths = np.ascontiguousarray(lengths, dtype=np.uint64)
        res = np.zeros(len(lengths), dtype=DOC_RESULT)
        tn = ctypes.c_uint64(0)
        rc = self.lib.alpacc_parse_batch(self.handle, ctypes.c_void_p(d_in), offsets.ctypes.data,
                                         lengths.ctypes.data, len(lengths), ctypes.c_void_p(parent),
                                         ctypes.c_void_p(node), ctypes.c_void_p(span), cap,
                                         ctypes.byref(tn), res.ctypes.data, stream)
        if rc < 0 and rc != -3:
            raise self._err(rc)
        return rc, tn.value, res

    def lex_int_batch(self, docs) -> list:
        """``lex_int`` of every document in one batched pass: a list of None / (kinds, starts,
        ends)."""
        buf, offsets, lengths = self.upload_batch(docs)
        cap = int(lengths.sum()) + len(docs) + 1
        kb = self.buffer("bkinds", 4 * cap)
        sb = self.buffer("bstarts", 4 * cap)
        eb = self.buffer("bends", 4 * cap)
        res = self.lex_batch_device(buf.ptr, offsets, lengths, kb.ptr, sb.ptr, eb.ptr, cap,
                                    self.stream)
        k = kb.download(np.uint32, cap, self.stream)
        s_ = sb.download(np.uint32, cap, self.stream)
        e = eb.download(np.uint32, cap, self.stream)
        out = []
        for r in res:
            if r["status"] != ACCEPT:
                out.append(None)
                continue
            o, c = int(r["offset"]), int(r["count"])
            out.append((k[o:o + c], s_[o:o + c], e[o:o + c]))
        return out

    def parse_int_batch(self, docs, cap: int | None = None) -> list:
        """``parse_int`` of every document in one batched pass: a list of None /
        (parent, node, span[N,2])."""
        buf, offsets, lengths = self.upload_batch(docs)
        if cap is None:
            cap = max(16, 4 * int(lengths.sum()) + 16)
        while True:
            pb = self.buffer("bparent", 4 * cap)
            nb = self.buffer("bnode", 4 * cap)
            spb = self.buffer("bspan", 8 * cap)
            rc, total, res = self.parse_batch_device(buf.ptr, offsets, lengths, pb.ptr, nb.ptr,
                                                     spb.ptr, cap, self.stream)
            if rc == -3:
                cap = total
                continue
            break
        par = pb.download(np.uint32, total, self.stream)
        nd = nb.download(np.uint32, total, self.stream)
        sp = spb.download(np.uint32, 2 * total, self.stream).reshape(-1, 2)
        out = []
        for r in res:
            if r["status"] != ACCEPT:
                out.append(None)
                continue
            o, c = int(r["offset"]), int(r["count"])
            out.append((par[o:o + c], nd[o:o + c], sp[o:o + c]))
        return out

    def parse(self, data, chunk_size: int | None = None):
        r = self.parse_int(data, chunk_size=chunk_size)
        if r is None:
            return None
        tnames = self.bundle.terminal_names
        pnames = self.bundle.parser.production_names
        out = []
        for par, nd, sp in zip(*r):
            if nd & NODE_TERMINAL:
                out.append((int(par), ("terminal", tnames[nd & 0x7FFFFFFF], (int(sp[0]), int(sp[1])))))
            else:
                out.append((int(par), ("production", pnames[nd])))
        return out


def tree_tuples(r):
    """(parent, is_terminal, id, start, end) tuples of a parse_int result (the wire-format view)."""
    if r is None:
        return None
    par, node, span = r
    out = []
    for p, nd, sp in zip(par.tolist(), node.tolist(), span.tolist()):
        if nd & NODE_TERMINAL:
            out.append((p, True, nd & 0x7FFFFFFF, sp[0], sp[1]))
        else:
            out.append((p, False, nd, 0, 0))
    return out


def format_parse(tree) -> str:
    """The printed form of a ``parse`` result as README.md:84-106 shows it: ``#some [(parent,
    #production (#name)), (parent, #terminal (#name) (start, end)), ...]``, one node a line, or
    ``#none`` for a rejected input (``tree`` is None)."""
    if tree is None:
        return "#none"
    items = []
    for parent, node in tree:
        if node[0] == "terminal":
            _, name, (s, e) = node
            items.append(f"({parent}, #terminal (#{name}) ({s}, {e}))")
        else:
            items.append(f"({parent}, #production (#{node[1]}))")
    return "#some [" + ",\n       ".join(items) + "]"


def format_lex(tokens) -> str:
    """The same form for a ``lex`` result (lexer.fut:180-191: opt [(terminal, (start, end))]):
    ``#some [(#name, (start, end)), ...]`` or ``#none``."""
    if tokens is None:
        return "#none"
    return "#some [" + ",\n       ".join(f"(#{t}, ({s}, {e}))" for t, (s, e) in tokens) + "]"
