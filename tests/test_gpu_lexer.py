"""GPU lexer parity: alpacc_lex (HIP) against the sequential oracle (tokenize, DFA.hs:415-447) and
the literal restatement of futhark/lexer.fut, bit-exact, including rejections."""
import itertools
import os

import numpy as np
import pytest

from conftest import GRAMMARS

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctxs():
    from alpacc_amd.gen.tables import generate
    from alpacc_amd.runtime import Context
    out = {}
    for g in ("json", "arithmetic", "sexp"):
        b = generate(os.path.join(GRAMMARS, f"{g}.alp"))
        out[g] = (b, Context(b))
    lb = generate(os.path.join(GRAMMARS, "json.alp"), mode="lexer")
    out["json-lexer"] = (lb, Context(lb))
    yield out
    for _, c in out.values():
        c.close()


def _tokens(r):
    if r is None:
        return None
    return [tuple(int(v) for v in t) for t in zip(*r)]


def _oracle(bundle, data):
    from oracle import sequential
    r = sequential.lex(bundle, data)
    return None if r is None else [tuple(t) for t in r]


def _class_reps(bundle):
    bc = bundle.lexer.byte_class
    reps = []
    for c in range(bundle.lexer.num_classes):
        reps.append(int(np.nonzero(bc == c)[0][0]))
    return reps


@pytest.mark.parametrize("g,maxlen", [("sexp", 5), ("arithmetic", 3), ("json", 2)])
def test_exhaustive_short(ctxs, g, maxlen):
    """All strings up to maxlen over one byte per class (the reference's exhaustive mode,
    src/Alpacc/Test/Lexer.hs:154-179, on a class-reduced alphabet)."""
    bundle, ctx = ctxs[g]
    reps = _class_reps(bundle)
    n_ok = 0
    for L in range(0, maxlen + 1):
        for tup in itertools.product(reps, repeat=L):
            data = bytes(tup)
            got = _tokens(ctx.lex_int(data))
            want = _oracle(bundle, data)
            assert got == want, (g, data)
            n_ok += want is not None
    assert n_ok > 0


@pytest.mark.parametrize("g", ["json", "arithmetic", "sexp", "json-lexer"])
def test_documents_tile_boundaries(ctxs, g):
    from alpacc_amd import datagen
    bundle, ctx = ctxs[g]
    gen = {"json": datagen.json_doc, "json-lexer": datagen.json_doc,
           "arithmetic": datagen.arith_doc, "sexp": datagen.sexp_doc}[g]
    for n in (1, 2, 63, 64, 65, 2047, 2048, 2049, 4095, 4096, 4097, 2048 * 64 + 1, 300_001):
        try:
            doc = gen(n, seed=n).tobytes()
        except RuntimeError:
            continue
        got = _tokens(ctx.lex_int(doc))
        want = _oracle(bundle, doc)
        assert got == want, (g, n)


def test_literal_lexer_agrees(ctxs):
    from alpacc_amd import datagen
    from oracle.lexer_fut import LexerFut
    bundle, ctx = ctxs["json"]
    lf = LexerFut(bundle.lexer)
    for n in (5000, 70_000):
        doc = datagen.json_doc(n, seed=3).tobytes()
        got = _tokens(ctx.lex_int(doc))
        r = lf.lex_int(doc)
        want = None if r is None else [tuple(int(v) for v in t) for t in zip(*r)]
        assert got == want


def test_rejections(ctxs):
    from alpacc_amd import datagen
    bundle, ctx = ctxs["json"]
    doc = datagen.json_doc(200_000, seed=11)
    bad = datagen.flip_bytes(doc, seed=5, block=4096)
    n_rej = 0
    for k in range(0, bad.shape[0], 20_000):
        d = bad[: k + 20_000].tobytes()
        got = _tokens(ctx.lex_int(d))
        want = _oracle(bundle, d)
        assert got == want
        n_rej += want is None
    assert ctx.lex_int(b"") is None
    assert ctx.lex_int(b"\x00\x01") is None


def test_large_document_invariants(ctxs):
    """8 MiB: size-independent properties (spans increasing and disjoint, gaps are ignore bytes,
    kinds in range) plus the token count of the literal restatement."""
    from alpacc_amd import datagen
    bundle, ctx = ctxs["json"]
    n = 8 << 20
    doc = datagen.json_doc(n, seed=99)
    r = ctx.lex_int(doc)
    assert r is not None
    kinds, starts, ends = (x.astype(np.int64) for x in r)
    assert np.all(starts < ends)
    assert np.all(ends[:-1] <= starts[1:])
    gaps = np.ones(n, dtype=bool)
    # bytes covered by tokens
    cover = np.zeros(n + 1, dtype=np.int64)
    np.add.at(cover, starts, 1)
    np.add.at(cover, ends, -1)
    covered = np.cumsum(cover)[:n] > 0
    gaps &= ~covered
    assert set(np.unique(doc[gaps]).tolist()) <= {0x20, 0x0A, 0x09, 0x0D}
    assert kinds.max() < len(bundle.terminal_names)
    want = _oracle(bundle, doc[: 1 << 20].tobytes())
    got = _tokens(ctx.lex_int(doc[: 1 << 20]))
    assert got == want


@pytest.mark.parametrize("g", ["json", "sexp"])
def test_unaligned_input(ctxs, g):
    """Input pointers 1..15 bytes past a 16-byte boundary: every lane reads the aligned 16-byte
    words around its bytes (at the document's edges clamped to the words holding its first and
    last byte); tokens must equal the C restatement's."""
    from alpacc_amd import datagen
    from alpacc_amd.runtime import DeviceBuffer
    from oracle.c import coracle
    bundle, ctx = ctxs[g]
    orc = coracle.Oracle(bundle)
    gen = {"json": datagen.json_doc, "sexp": datagen.sexp_doc}[g]
    for n in (40, 2049, 70_001):
        doc = gen(n, seed=n + 7)
        n = len(doc)
        want = orc.lex(doc.tobytes())
        want = None if want is None else [tuple(int(v) for v in t) for t in zip(*want)]
        buf = DeviceBuffer(n + 64)
        cap = max(n, 1)
        outs = [DeviceBuffer(4 * cap) for _ in range(3)]
        try:
            for off in range(1, 16):
                buf.upload(doc, offset=off)
                rc, cnt = ctx.lex_device(buf.ptr + off, n, outs[0].ptr, outs[1].ptr, outs[2].ptr,
                                         cap)
                got = None
                if rc == 1:
                    k, s, e = (o.download(np.uint32, cnt) for o in outs)
                    got = [tuple(int(v) for v in t) for t in zip(k, s, e)]
                assert got == want, (g, n, off)
        finally:
            buf.free()
            for o in outs:
                o.free()


@pytest.mark.parametrize("g", ["json", "sexp"])
def test_document_end_at_every_lane_offset(ctxs, g):
    """The document's last byte at every offset of a lane's 32 bytes (lengths 1-40, around the
    first wave tile's end and near the second's), the input 0, 5 or 13 bytes past a 16-byte
    boundary. The last lane walks all its classes like the others, its marks past the end are
    cleared and the state after its last byte is walked again for the verdict; cuts inside a
    token change the verdict. Tokens and verdicts must equal the C restatement's."""
    from alpacc_amd import datagen
    from alpacc_amd.runtime import DeviceBuffer
    from oracle.c import coracle
    bundle, ctx = ctxs[g]
    orc = coracle.Oracle(bundle)
    gen = {"json": datagen.json_doc, "sexp": datagen.sexp_doc}[g]
    base = gen(6000, seed=31)
    lengths = list(range(1, 41)) + list(range(2048 - 33, 2048 + 34)) + [4064 + r for r in range(33)]
    assert len(base) >= max(lengths)
    buf = DeviceBuffer(len(base) + 64)
    cap = len(base) + 1
    outs = [DeviceBuffer(4 * cap) for _ in range(3)]
    verdicts = set()
    try:
        for n in lengths:
            doc = base[:n]
            want = orc.lex(doc.tobytes())
            want = None if want is None else [tuple(int(v) for v in t) for t in zip(*want)]
            verdicts.add(want is None)
            for off in (0, 5, 13):
                buf.upload(doc, offset=off)
                rc, cnt = ctx.lex_device(buf.ptr + off, n, outs[0].ptr, outs[1].ptr, outs[2].ptr,
                                         cap)
                got = None
                if rc == 1:
                    k, s, e = (o.download(np.uint32, cnt) for o in outs)
                    got = [tuple(int(v) for v in t) for t in zip(k, s, e)]
                assert got == want, (g, n, off)
    finally:
        buf.free()
        for o in outs:
            o.free()
    if g == "json":  # every prefix of an s-expression lexes
        assert verdicts == {True, False}, "the cuts should give both verdicts"
