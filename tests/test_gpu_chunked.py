"""GPU parity of the chunked mode: alpacc_lex_chunked / alpacc_parse_chunked against the literal
restatements of the reference's chunk loop (lex_int_flag, futhark/lexer.fut:144-172): the numpy one
(oracle/lexer_fut.py, then oracle/parser_fut.py) and the C one (oracle/c/oracle.c). Chunk sizes
from 1 byte up; the reference's own 16 MiB chunk on a 64 MiB document and its invalid variants."""
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
    yield out
    for _, c in out.values():
        c.close()


def _rows(r):
    return None if r is None else [tuple(int(v) for v in t) for t in zip(*r)]


@pytest.mark.parametrize("g,maxlen", [("sexp", 4), ("arithmetic", 4), ("json", 2)])
def test_exhaustive_short_strings(ctxs, g, maxlen):
    """Every string up to maxlen over one byte per class, chunk sizes 1, 2 and 3 (chunk loops of up
    to 4 iterations, overlap bytes at every position)."""
    from oracle.lexer_fut import LexerFut
    b, ctx = ctxs[g]
    lf = LexerFut(b.lexer)
    reps = [int(np.nonzero(b.lexer.byte_class == c)[0][0]) for c in range(b.lexer.num_classes)]
    print(f"exhaustive {g} maxlen={maxlen}", flush=True)
    for L in range(1, maxlen + 1):
        for tup in itertools.product(reps, repeat=L):
            d = bytes(tup)
            for C in (1, 2, 3):
                want = _rows(lf.lex_int(d, chunk_size=C))
                got = _rows(ctx.lex_int(d, chunk_size=C))
                assert got == want, (d, C)


@pytest.mark.parametrize("g", ["json", "arithmetic", "sexp"])
def test_documents_small_chunks(ctxs, g):
    from alpacc_amd import datagen
    from oracle.c import coracle
    b, ctx = ctxs[g]
    orc = coracle.Oracle(b)
    gen = {"json": datagen.json_doc, "arithmetic": datagen.arith_doc, "sexp": datagen.sexp_doc}[g]
    for n, seed in ((1000, 1), (5000, 2), (70_001, 3)):
        doc = gen(n, seed=seed)
        for d in (doc, datagen.flip_bytes(doc, seed=seed, block=max(64, n // 4))):
            d = d.tobytes()
            for C in (1, 7, 64, 333, 2047, 2048, 4096, n - 1, n, n + 1):
                print(f"doc {g} n={n} seed={seed} chunk={C}", flush=True)
                want = _rows(orc.lex(d, chunk=C, wide=True))
                got = _rows(ctx.lex_int(d, chunk_size=C))
                assert got == want, (g, n, C)


@pytest.mark.parametrize("g", ["json", "arithmetic", "sexp"])
def test_parse_chunked(ctxs, g):
    """parse_int of the lexer+parser artifact (lex_int chunked, then parser.parse_int)."""
    from alpacc_amd import datagen
    from oracle.c import coracle
    from oracle.parser_fut import ParserFut
    b, ctx = ctxs[g]
    orc = coracle.Oracle(b)
    pf = ParserFut(b.parser)
    gen = {"json": datagen.json_doc, "arithmetic": datagen.arith_doc, "sexp": datagen.sexp_doc}[g]
    for n, seed in ((300, 4), (3000, 5), (20_000, 6)):
        d = gen(n, seed=seed).tobytes()
        for C in (1, 5, 64, 1000, n, 1 << 24):
            print(f"parse {g} n={n} chunk={C}", flush=True)
            toks = orc.lex(d, chunk=C, wide=True)
            want = None if toks is None else pf.parse_int(*toks)
            got = ctx.parse_int(d, chunk_size=C)
            assert (got is None) == (want is None), (g, n, C)
            if got is None:
                continue
            par, node, span = got
            wp, wt, wid, ws, we = want
            assert np.array_equal(par, wp)
            assert np.array_equal((node & 0x80000000) != 0, wt)
            assert np.array_equal(node & 0x7FFFFFFF, wid)
            assert np.array_equal(span[:, 0], np.where(wt, ws, 0))
            assert np.array_equal(span[:, 1], np.where(wt, we, 0))


def test_unchunked_equivalence(ctxs):
    """A chunk covering the document gives the unchunked result, with 64-bit spans."""
    from alpacc_amd import datagen
    b, ctx = ctxs["json"]
    d = datagen.json_doc(100_000, seed=8).tobytes()
    u = ctx.lex_int(d)
    c = ctx.lex_int(d, chunk_size=len(d))
    assert _rows(u) == _rows(c)
    assert c[1].dtype == np.uint64


def test_reference_chunk_64mib(ctxs):
    """The reference's entries at 64 MiB: four 16 MiB chunks, against the C restatement of the
    chunk loop. Seed 42: the doubled overlap byte breaks a keyword, so lex_int rejects; its invalid
    variant (one byte flipped per MiB). Seed 40: lex_int accepts and every span of the first chunk
    carries the loop's re-shift of 0+16+32+48 MiB; parse_int rejects, because the element ending
    each chunk overwrites the last token of the chunk before (lexer.fut:108-111, 131-132)."""
    from alpacc_amd import datagen
    from alpacc_amd.runtime import REFERENCE_CHUNK
    from oracle.c import coracle
    b, ctx = ctxs["json"]
    orc = coracle.Oracle(b)
    n = 64 << 20
    doc42 = datagen.json_doc(n, seed=42)
    doc40 = datagen.json_doc(n, seed=40)
    cases = (("seed 42", doc42, False), ("seed 42 flipped",
             datagen.flip_bytes(doc42, seed=42, block=1 << 20), None), ("seed 40", doc40, True))
    for label, d, accepted in cases:
        print(f"64 MiB {label}", flush=True)
        want = orc.lex(d, chunk=REFERENCE_CHUNK, wide=True)
        if accepted is not None:
            assert (want is not None) == accepted, label
        got = ctx.lex_int(d, chunk_size=REFERENCE_CHUNK)
        assert (got is None) == (want is None), label
        if want is not None:
            for x, y in zip(got, want):
                assert np.array_equal(x.astype(np.int64), y.astype(np.int64)), label
            un = ctx.lex_int(d)
            assert int(got[1][0]) == int(un[1][0]) + 6 * REFERENCE_CHUNK
            assert orc.parse_tokens(*want) is None
        assert ctx.parse_int(d, chunk_size=REFERENCE_CHUNK) is None, label
