"""alpacc_lex_noalloc (the allocation-free lexer entry, scratch in one static device array per GPU)
against alpacc_lex through the C ABI: identical status, token count and (kind, start, end) on two
calls in a row, the second with a smaller document (the shared scratch keeps the first call's
contents), then a rejected document, then a document above the limit, which is refused before any
GPU work."""
import ctypes
import os

import numpy as np
import pytest

from conftest import GRAMMARS

pytestmark = pytest.mark.gpu


def _lex(ctx, fn, d_in, n, outs):
    nt = ctypes.c_uint64(0)
    rc = fn(ctx.handle, ctypes.c_void_p(d_in), n, ctypes.c_void_p(outs[0].ptr),
            ctypes.c_void_p(outs[1].ptr), ctypes.c_void_p(outs[2].ptr), n, ctypes.byref(nt), None)
    got = tuple(o.download(np.uint32, nt.value) for o in outs) if rc == 1 else None
    return rc, nt.value, got


def test_noalloc_equals_lex():
    from alpacc_amd import datagen
    from alpacc_amd.gen.tables import generate
    from alpacc_amd.runtime import Context, DeviceBuffer
    b = generate(os.path.join(GRAMMARS, "json.alp"))
    ctx = Context(b)
    lib = ctx.lib
    big = datagen.json_doc(24 << 20, seed=71)
    small = datagen.json_doc(300_001, seed=72)
    bad = datagen.flip_bytes(datagen.json_doc(1 << 20, seed=73), seed=5, block=1 << 18)
    n_max = big.shape[0]
    d_in = DeviceBuffer(n_max, lib)
    ref = [DeviceBuffer(4 * n_max, lib) for _ in range(3)]
    out = [DeviceBuffer(4 * n_max, lib) for _ in range(3)]
    try:
        for name, doc in (("24 MiB", big), ("300 KB", small), ("flipped 1 MiB", bad)):
            print(f"alpacc_lex_noalloc vs alpacc_lex: {name}", flush=True)
            n = doc.shape[0]
            d_in.upload(doc)
            want = _lex(ctx, lib.alpacc_lex, d_in.ptr, n, ref)
            # poison the outputs so that a token the entry does not write cannot match by chance
            for o in out:
                lib.alpacc_device_memset(ctypes.c_void_p(o.ptr), 0xAB, 4 * n_max, None)
            got = _lex(ctx, lib.alpacc_lex_noalloc, d_in.ptr, n, out)
            assert got[0] == want[0] and got[1] == want[1], name
            if want[2] is not None:
                for g_, w_ in zip(got[2], want[2]):
                    assert np.array_equal(g_, w_), name
        # above the limit: refused before any GPU work (the input pointer is never read)
        nt = ctypes.c_uint64(0)
        rc = lib.alpacc_lex_noalloc(ctx.handle, ctypes.c_void_p(d_in.ptr), (64 << 20) + 1,
                                    ctypes.c_void_p(out[0].ptr), ctypes.c_void_p(out[1].ptr),
                                    ctypes.c_void_p(out[2].ptr), 1 << 40, ctypes.byref(nt), None)
        assert rc == -1
        assert b"64 MiB" in lib.alpacc_last_error(ctx.handle)
    finally:
        for o in ref + out:
            o.free()
        d_in.free()
        ctx.close()
