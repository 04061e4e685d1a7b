"""ctypes wrapper of the C/OpenMP oracle (oracle/c/oracle.c)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import build as _build

_lib = None
_i64p = ctypes.POINTER(ctypes.c_int64)
_u32p = ctypes.POINTER(ctypes.c_uint32)


class OTables(ctypes.Structure):
    _fields_ = [
        ("e", ctypes.c_int64), ("comp", ctypes.c_void_p), ("trans", ctypes.c_void_p),
        ("accept", ctypes.c_void_p), ("identity", ctypes.c_uint32), ("index_mask", ctypes.c_uint32),
        ("t_mask", ctypes.c_uint32), ("t_off", ctypes.c_uint32), ("p_off", ctypes.c_uint32),
        ("ignore", ctypes.c_int32),
        ("q", ctypes.c_int64), ("k", ctypes.c_int64), ("start_t", ctypes.c_int64),
        ("end_t", ctypes.c_int64), ("empty_t", ctypes.c_int64), ("hsize", ctypes.c_int64),
        ("max_iters", ctypes.c_int64), ("hkeys", ctypes.c_void_p), ("hused", ctypes.c_void_p),
        ("hvals", ctypes.c_void_p), ("stacks", ctypes.c_void_p), ("top_bit", ctypes.c_int64),
        ("prods", ctypes.c_void_p), ("arity", ctypes.c_void_p), ("p2t", ctypes.c_void_p),
    ]


def load():
    global _lib
    if _lib is None:
        so = _build.OUT
        if not os.path.exists(so) or os.path.getmtime(so) < os.path.getmtime(_build.SRC):
            _build.build()
        L = ctypes.CDLL(so)
        vp = ctypes.c_void_p
        L.oracle_lex.argtypes = [vp, vp, ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(vp),
                                 ctypes.POINTER(vp), ctypes.POINTER(vp), _i64p]
        L.oracle_parse.argtypes = [vp, vp, vp, vp, ctypes.c_int64, ctypes.POINTER(vp),
                                   ctypes.POINTER(vp), ctypes.POINTER(vp), _i64p]
        L.oracle_free.argtypes = [vp]
        L.oracle_threads.restype = ctypes.c_int
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_set_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _take(lib, ptr, dtype, count):
    if count == 0:
        lib.oracle_free(ptr)
        return np.zeros(0, dtype=dtype)
    buf = (ctypes.c_char * (count * np.dtype(dtype).itemsize)).from_address(ptr.value)
    out = np.frombuffer(buf, dtype=dtype).copy()
    lib.oracle_free(ptr)
    return out


class Oracle:
    def __init__(self, bundle):
        self.lib = load()
        self.threads = self.lib.oracle_threads()
        self.keep = []
        t = OTables()

        def hold(a, dtype):
            a = np.ascontiguousarray(a, dtype=dtype)
            self.keep.append(a)
            return a.ctypes.data

        lx = bundle.lexer
        if lx is not None:
            et = lx.endo
            t.e = et.num_endos
            t.comp = hold(et.compositions, np.uint32)
            t.trans = hold(et.transitions_to_states, np.uint32)
            t.accept = hold(et.accept, np.uint8)
            t.identity = et.identity
            t.index_mask = et.index_mask
            t.t_mask = et.terminal_mask
            t.t_off = et.terminal_offset
            t.p_off = et.produce_offset
            t.ignore = lx.ignore
        px = bundle.parser
        if px is not None:
            t.q, t.k = px.q, px.k
            t.start_t, t.end_t, t.empty_t = px.start_terminal, px.end_terminal, px.empty_terminal
            t.hsize = px.hash_used.shape[0]
            t.max_iters = px.max_iters
            t.hkeys = hold(px.hash_keys, np.int64)
            t.hused = hold(px.hash_used, np.uint8)
            t.hvals = hold(px.hash_vals, np.int64)
            t.stacks = hold(px.flat_stacks, np.int64)
            t.top_bit = 1 << (px.bracket_bits - 1)
            t.prods = hold(px.flat_productions, np.int64)
            t.arity = hold(px.arities, np.int32)
            t.p2t = hold(px.production_to_terminal, np.int32)
        self.t = t

    def lex(self, data, chunk: int = 0, wide: bool = False):
        """None when rejected, else (kinds u32, starts, ends) (u32, or int64 with wide=True)."""
        arr = np.frombuffer(bytes(data), dtype=np.uint8) if isinstance(data, (bytes, bytearray)) \
            else np.ascontiguousarray(data, dtype=np.uint8)
        vp = ctypes.c_void_p
        k, s, e = vp(), vp(), vp()
        nt = ctypes.c_int64()
        ok = self.lib.oracle_lex(ctypes.byref(self.t), arr.ctypes.data, arr.shape[0], chunk,
                                 ctypes.byref(k), ctypes.byref(s), ctypes.byref(e), ctypes.byref(nt))
        kinds = _take(self.lib, k, np.uint32, nt.value)
        starts = _take(self.lib, s, np.int64, nt.value)
        ends = _take(self.lib, e, np.int64, nt.value)
        if not ok:
            return None
        if wide:
            return kinds, starts, ends
        return kinds, starts.astype(np.uint32), ends.astype(np.uint32)

    def parse_tokens(self, kinds, starts, ends):
        kinds = np.ascontiguousarray(kinds, dtype=np.uint32)
        starts = np.ascontiguousarray(starts, dtype=np.int64)
        ends = np.ascontiguousarray(ends, dtype=np.int64)
        vp = ctypes.c_void_p
        p, nd, sp = vp(), vp(), vp()
        nn = ctypes.c_int64()
        ok = self.lib.oracle_parse(ctypes.byref(self.t), kinds.ctypes.data, starts.ctypes.data,
                                   ends.ctypes.data, kinds.shape[0], ctypes.byref(p), ctypes.byref(nd),
                                   ctypes.byref(sp), ctypes.byref(nn))
        if not ok:
            return None
        N = nn.value
        return (_take(self.lib, p, np.uint32, N), _take(self.lib, nd, np.uint32, N),
                _take(self.lib, sp, np.uint32, 2 * N))

    def set_threads(self, n: int) -> int:
        """OpenMP threads of the following calls (process-wide); returns the count now in use."""
        self.threads = self.lib.oracle_set_threads(int(n))
        return self.threads

    def parse(self, data, chunk: int = 0):
        """None when rejected, else (parent u32, node u32, span u32[2N]) like alpacc_parse.
        chunk > 0: the lexer's chunk loop (lexer.fut:144-172) over chunks of that many bytes, as
        the generated parse_int runs it with 16 MiB chunks."""
        r = self.lex(data, chunk=chunk, wide=True)
        if r is None:
            return None
        return self.parse_tokens(*r)
