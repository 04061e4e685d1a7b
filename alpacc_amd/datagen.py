"""Seeded synthetic workloads (C generators in csrc/datagen.c, wrapped with ctypes).

Workloads:
  * ``json_doc(n)``      - a JSON document of exactly n bytes (lexer / lex+parse workloads);
  * ``arith_doc(n)``     - an arithmetic expression chain of exactly n bytes;
  * ``sexp_doc(n)``      - one s-expression list of exactly n bytes;
  * ``flip_bytes(doc)``  - one uniformly chosen byte per MiB block XORed with 1 + rng(255)
                           (invalid-input variants; the verdict comes from the oracle, not assumed);
  * ``sexp_batch_sizes`` - 4096 log-normal document lengths (median 64 KiB, clipped to
                           [1 KiB, 16 MiB]) whose clipped mean is 512 KiB, summing to exactly 2^31.
The seed defaults to 42.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "datagen.c")
LIB = os.path.join(HERE, "libalpacc_datagen.so")
SEED = 42

_lib = None


def build() -> str:
    tmp = LIB + ".tmp"
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", SRC, "-o", tmp], check=True)
    os.replace(tmp, LIB)
    return LIB


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
            build()
        L = ctypes.CDLL(LIB)
        for f in ("gen_json", "gen_arith", "gen_sexp"):
            fn = getattr(L, f)
            fn.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64]
            fn.restype = ctypes.c_size_t
        _lib = L
    return _lib


def _gen(name: str, n: int, seed: int, out: np.ndarray | None = None) -> np.ndarray:
    L = _load()
    buf = np.empty(n, dtype=np.uint8) if out is None else out
    assert buf.shape[0] == n and buf.dtype == np.uint8 and buf.flags.c_contiguous
    got = getattr(L, name)(buf.ctypes.data, n, seed)
    if got != n:
        raise RuntimeError(f"{name} produced {got} of {n} bytes")
    return buf


def json_doc(n: int, seed: int = SEED, out=None) -> np.ndarray:
    return _gen("gen_json", n, seed, out)


def arith_doc(n: int, seed: int = SEED, out=None) -> np.ndarray:
    return _gen("gen_arith", n, seed, out)


def sexp_doc(n: int, seed: int = SEED, out=None) -> np.ndarray:
    return _gen("gen_sexp", n, seed, out)


def flip_bytes(doc: np.ndarray, seed: int = SEED, block: int = 1 << 20) -> np.ndarray:
    out = doc.copy()
    rng = np.random.Generator(np.random.PCG64(seed))
    nb = (out.shape[0] + block - 1) // block
    for b in range(nb):
        lo = b * block
        hi = min(lo + block, out.shape[0])
        i = lo + int(rng.integers(0, hi - lo))
        out[i] ^= np.uint8(1 + int(rng.integers(0, 255)))
    return out


def sexp_batch_sizes(count: int = 4096, total: int = 1 << 31, median: int = 64 << 10,
                     lo: int = 1 << 10, hi: int = 16 << 20, seed: int = SEED) -> np.ndarray:
    """Log-normal lengths clipped to [lo, hi] with sigma solved so the clipped mean is total/count.

    The draws use one fixed standard-normal sample; sigma is found by bisection, then the residual
    to the exact total is spread over the largest documents within the clip.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    z = rng.standard_normal(count)
    mu = np.log(median)
    target = total / count

    def clipped(sig):
        return np.clip(np.exp(mu + sig * z), lo, hi)

    a, b = 0.01, 6.0
    for _ in range(100):
        m = 0.5 * (a + b)
        if clipped(m).mean() < target:
            a = m
        else:
            b = m
    sizes = np.floor(clipped(0.5 * (a + b))).astype(np.int64)
    resid = total - int(sizes.sum())
    order = np.argsort(-sizes, kind="stable")
    i = 0
    while resid != 0:
        j = order[i % count]
        step = 1 if resid > 0 else -1
        room = (hi - sizes[j]) if step > 0 else (sizes[j] - lo)
        take = min(abs(resid), room)
        sizes[j] += step * take
        resid -= step * take
        i += 1
        if i > 4 * count and resid != 0:
            raise RuntimeError("cannot reach the exact total within the clip")
    return sizes
