"""GPU lex+parse parity: alpacc_parse (HIP) against the sequential LLP oracle (llpParse +
flattenTree, LLP.hs:600-745) and the literal restatement of futhark/parser.fut, bit-exact."""
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


def _tree(ctx, data):
    from alpacc_amd.runtime import tree_tuples
    return tree_tuples(ctx.parse_int(data))


def test_readme_tree(ctxs):
    """README.md:81-106: 12+(3*4) parsed with grammars/arithmetic.alp."""
    bundle, ctx = ctxs["arithmetic"]
    got = _tree(ctx, b"12+(3*4)")
    parents = [0, 0, 1, 2, 1, 0, 5, 5, 7, 8, 8, 10, 11, 12, 11, 14, 14, 16, 14, 10, 8, 7, 5]
    assert [p for p, *_ in got] == parents
    names = ctx.parse(b"12+(3*4)")
    assert names[0] == (0, ("production", "E0_0"))
    assert names[3] == (2, ("terminal", "num", (0, 2)))
    assert names[6] == (5, ("terminal", "literal_5", (2, 3)))
    assert names[15] == (14, ("terminal", "literal_4", (5, 6)))
    assert names[-1] == (5, ("production", "Empty"))
    # printed as the README prints it (README.md:84-106, whitespace collapsed)
    import json
    from alpacc_amd.runtime import format_parse
    kat = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                      "haskell_vectors.json")))["readme_tree"]
    assert " ".join(format_parse(names).split()) == kat["text"]
    assert format_parse(ctx.parse(b"12+(3*")) == "#none"


@pytest.mark.parametrize("g", ["json", "arithmetic", "sexp"])
def test_documents(ctxs, g):
    from alpacc_amd import datagen
    from oracle import sequential
    bundle, ctx = ctxs[g]
    gen = {"json": datagen.json_doc, "arithmetic": datagen.arith_doc, "sexp": datagen.sexp_doc}[g]
    for n in (2, 10, 100, 2047, 2049, 5000, 65_537, 250_000):
        try:
            doc = gen(n, seed=1000 + n).tobytes()
        except RuntimeError:
            continue
        got = _tree(ctx, doc)
        want = sequential.parse(bundle, doc)
        assert got == want, (g, n)


def test_literal_parser_agrees(ctxs):
    from alpacc_amd import datagen
    from oracle.lexer_fut import LexerFut
    from oracle.parser_fut import ParserFut
    bundle, ctx = ctxs["json"]
    doc = datagen.json_doc(100_000, seed=5).tobytes()
    k, s, e = LexerFut(bundle.lexer).lex_int(doc)
    par, is_t, ident, st, en = ParserFut(bundle.parser).parse_int(k, s, e)
    want = [(int(a), bool(b), int(c), int(d), int(f)) for a, b, c, d, f in zip(par, is_t, ident, st, en)]
    assert _tree(ctx, doc) == want


def test_invalid_variants(ctxs):
    """Byte flips: the verdict must equal the oracle's (some flips stay valid)."""
    from alpacc_amd import datagen
    from oracle import sequential
    bundle, ctx = ctxs["json"]
    doc = datagen.json_doc(64_000, seed=21)
    for seed in range(12):
        bad = datagen.flip_bytes(doc, seed=seed, block=16_000).tobytes()
        assert _tree(ctx, bad) == sequential.parse(bundle, bad), seed
    # structurally wrong but lexically valid
    for d in (b"[1,2", b"[1,,2]", b"{\"a\" 1}", b"]", b"[]]", b"{}", b"[{}]", b"1 2", b"   "):
        assert _tree(ctx, d) == sequential.parse(bundle, d), d
