"""CPU tests of the oracles: the README known-answer tree, agreement of the three independent
restatements (sequential tokenize/llpParse, literal numpy lexer.fut/parser.fut, C/OpenMP), the
chunked lexer mode, wire formats and the data generators (no GPU)."""
import itertools
import os
import struct

import numpy as np
import pytest

from conftest import GRAMMARS, ROOT

GOLDEN_TREE = None


@pytest.fixture(scope="module")
def bundles():
    from alpacc_amd.gen.tables import generate
    return {g: generate(os.path.join(GRAMMARS, f"{g}.alp")) for g in ("json", "arithmetic", "sexp")}


def _tree_tuples(par, node, span):
    out = []
    for i, (p, nd) in enumerate(zip(par.tolist(), node.tolist())):
        if nd & 0x80000000:
            out.append((p, True, nd & 0x7FFFFFFF, int(span[2 * i]), int(span[2 * i + 1])))
        else:
            out.append((p, False, nd, 0, 0))
    return out


def test_readme_tree_all_oracles(bundles):
    """README.md:81-106 (nodes named as in the README; the input is 12+(3*4))."""
    import json
    from oracle import sequential
    from oracle.c import coracle
    from oracle.lexer_fut import LexerFut
    from oracle.parser_fut import ParserFut
    kat = json.load(open(os.path.join(ROOT, "tests", "golden", "haskell_vectors.json")))["readme_tree"]
    b = bundles["arithmetic"]
    data = kat["input"].encode()
    tnames = b.terminal_names
    pnames = b.parser.production_names
    want = []
    for nd in kat["nodes"]:
        if nd["kind"] == "terminal":
            want.append((nd["parent"], True, tnames.index(nd["name"]), nd["span"][0], nd["span"][1]))
        else:
            want.append((nd["parent"], False, nd["name"], 0, 0))
    seq = sequential.parse(b, data)
    assert [(p, t, i if t else pnames[i], s, e) for p, t, i, s, e in seq] == want
    k, s, e = LexerFut(b.lexer).lex_int(data)
    lit = ParserFut(b.parser).parse_int(k, s, e)
    lit = [(int(a), bool(c), int(d), int(f), int(g)) for a, c, d, f, g in zip(*lit)]
    assert lit == seq
    assert _tree_tuples(*coracle.Oracle(b).parse(data)) == seq


def test_readme_printed_form(bundles):
    """format_parse prints the README's block (README.md:84-106, whitespace collapsed) for the
    sequential parser's tree of 12+(3*4); format_lex prints the tokens the same way."""
    import json
    import re
    from alpacc_amd.runtime import format_lex, format_parse
    from oracle import sequential
    from oracle.c import coracle
    kat = json.load(open(os.path.join(ROOT, "tests", "golden", "haskell_vectors.json")))["readme_tree"]
    b = bundles["arithmetic"]
    data = kat["input"].encode()
    tnames, pnames = b.terminal_names, b.parser.production_names
    tree = [(p, ("terminal", tnames[i], (s, e)) if t else ("production", pnames[i]))
            for p, t, i, s, e in sequential.parse(b, data)]
    assert " ".join(format_parse(tree).split()) == kat["text"]
    assert format_parse(None) == "#none" and format_lex(None) == "#none"
    k, s, e = coracle.Oracle(b).lex(data)
    toks = [(tnames[int(x)], (int(y), int(z))) for x, y, z in zip(k, s, e)]
    text = format_lex(toks)
    assert text.startswith("#some [(#num, (0, 2)),\n       (#literal_5, (2, 3)),")
    assert re.findall(r"\(#(\w+), \((\d+), (\d+)\)\)", text) == \
        [(t, str(x), str(y)) for t, (x, y) in toks]
    assert [nd["span"] for nd in kat["nodes"] if nd["span"]] == [[x, y] for _, (x, y) in toks]


@pytest.mark.parametrize("g,maxlen", [("sexp", 5), ("arithmetic", 4), ("json", 2)])
def test_lexers_agree_exhaustive(bundles, g, maxlen):
    """tokenize (DFA.hs) == literal lexer.fut restatement == C oracle on every string up to
    maxlen over one byte per class (the reference's exhaustive test mode)."""
    from oracle import sequential
    from oracle.c import coracle
    from oracle.lexer_fut import LexerFut
    b = bundles[g]
    lf = LexerFut(b.lexer)
    orc = coracle.Oracle(b)
    reps = [int(np.nonzero(b.lexer.byte_class == c)[0][0]) for c in range(b.lexer.num_classes)]
    for L in range(0, maxlen + 1):
        for tup in itertools.product(reps, repeat=L):
            d = bytes(tup)
            want = sequential.lex(b, d)
            r = lf.lex_int(d)
            got = None if r is None else [tuple(int(v) for v in t) for t in zip(*r)]
            assert got == want, d
            r = orc.lex(d)
            got = None if r is None else [tuple(int(v) for v in t) for t in zip(*r)]
            assert got == want, d


@pytest.mark.parametrize("g", ["json", "arithmetic", "sexp"])
def test_parsers_agree_documents(bundles, g):
    from alpacc_amd import datagen
    from oracle import sequential
    from oracle.c import coracle
    from oracle.lexer_fut import LexerFut
    from oracle.parser_fut import ParserFut
    b = bundles[g]
    gen = {"json": datagen.json_doc, "arithmetic": datagen.arith_doc, "sexp": datagen.sexp_doc}[g]
    orc = coracle.Oracle(b)
    lf, pf = LexerFut(b.lexer), ParserFut(b.parser)
    for n in (2, 3, 17, 300, 4096, 40_000):
        for seed in (1, 2):
            doc = gen(n, seed=seed)
            for d in (doc, datagen.flip_bytes(doc, seed=seed, block=max(8, n // 3))):
                d = d.tobytes()
                want = sequential.parse(b, d)
                r = orc.parse(d)
                assert (None if r is None else _tree_tuples(*r)) == want, (g, n, seed)
                toks = lf.lex_int(d)
                lit = None if toks is None else pf.parse_int(*toks)
                lit = None if lit is None else [(int(a), bool(c), int(x), int(f), int(h))
                                                for a, c, x, f, h in zip(*lit)]
                assert lit == want, (g, n, seed)


def test_structural_rejections(bundles):
    from oracle import sequential
    from oracle.c import coracle
    b = bundles["json"]
    orc = coracle.Oracle(b)
    for d in (b"[1,2", b"[1,,2]", b'{"a" 1}', b"]", b"[]]", b"", b"   ", b"1 2", b'"\\"', b"[1.]"):
        r = orc.parse(d)
        assert (None if r is None else _tree_tuples(*r)) == sequential.parse(b, d), d
    assert sequential.parse(b, b"{}") is not None and sequential.parse(b, b"[1,2]") is not None
    assert sequential.parse(b, b"[1,2") is None and sequential.lex(b, b"[1,2") is not None


@pytest.mark.parametrize("g", ["json", "sexp"])
def test_chunked_mode_small_chunks(bundles, g):
    """lex_int_flag's chunk loop (lexer.fut:144-172): the C and numpy restatements agree at small
    chunk sizes, and a chunk covering the input equals the unchunked result."""
    from alpacc_amd import datagen
    from oracle.c import coracle
    from oracle.lexer_fut import LexerFut
    b = bundles[g]
    gen = datagen.json_doc if g == "json" else datagen.sexp_doc
    orc = coracle.Oracle(b)
    lf = LexerFut(b.lexer)
    for n in (50, 700):
        d = gen(n, seed=9).tobytes()
        for C in (1, 2, 7, 64, 333):
            a = orc.lex(d, chunk=C, wide=True)
            r = lf.lex_int(d, chunk_size=C)
            a = None if a is None else [tuple(int(v) for v in t) for t in zip(*a)]
            r = None if r is None else [tuple(int(v) for v in t) for t in zip(*r)]
            assert a == r, (n, C)
        big = orc.lex(d, chunk=n, wide=True)
        un = orc.lex(d, chunk=0, wide=True)
        assert (big is None) == (un is None)
        if un is not None:
            assert all(np.array_equal(x, y) for x, y in zip(big, un))


def test_wire_formats(bundles):
    from oracle import sequential, wire
    b = bundles["arithmetic"]
    inputs = [b"12+(3*4)", b"(", b"a"]
    blob = wire.encode_inputs(inputs)
    assert wire.decode_inputs(blob) == inputs
    assert blob[:8] == struct.pack(">Q", 3) and blob[8:16] == struct.pack(">Q", 8)
    trees = [sequential.parse(b, s) for s in inputs]
    enc = wire.encode_tree_outputs(trees)
    assert wire.decode_tree_outputs(enc) == trees
    # first node of the README tree: tag 0, parent 0, production 0, zero span
    assert enc[8:9] == b"\x01" and enc[17:18] == b"\x00" and enc[18:50] == struct.pack(">QQQQ", 0, 0, 0, 0)
    lx = wire.encode_lexer_outputs([sequential.lex(b, s) for s in inputs])
    assert lx[:9] == struct.pack(">Q", 3) + b"\x01"


def test_datagen_properties():
    from alpacc_amd import datagen
    for fn in (datagen.json_doc, datagen.arith_doc, datagen.sexp_doc):
        a = fn(10_000, seed=3)
        assert a.shape == (10_000,) and np.array_equal(a, fn(10_000, seed=3))
        assert not np.array_equal(a, fn(10_000, seed=4))
    sizes = datagen.sexp_batch_sizes()
    assert sizes.shape == (4096,) and int(sizes.sum()) == 1 << 31
    assert sizes.min() >= 1 << 10 and sizes.max() <= 16 << 20
    assert abs(np.median(sizes) / (64 << 10) - 1) < 0.25
    j = datagen.json_doc(1 << 20, seed=42)
    ws = np.isin(j, np.frombuffer(b" \n\t\r", dtype=np.uint8)).mean()
    assert 0.03 < ws < 0.2
