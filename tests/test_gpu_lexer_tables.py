"""GPU parity with the lexer's tables read from L2 instead of LDS: a keyword grammar whose DFA
table ((m + 1) x C entries) is larger than lex_emit's LDS copy holds, and json.alp with every
lexer table forced into L2 (lexer_tables_in_lds = 0), single and batched, against the C
restatement of lexer.fut + parser.fut."""
import os

import numpy as np
import pytest

from conftest import GRAMMARS
from test_gpu_fuzz_grammars import _check

pytestmark = pytest.mark.gpu

KEYWORDS = ["alpha", "beta", "gamma", "delta", "epsilon", "zeta", "theta", "iota", "kappa",
            "lambda", "mu", "omicron", "sigma", "upsilon", "omega", "while", "for", "return",
            "break", "continue", "switch", "case", "default", "struct", "union", "typedef",
            "extern", "static", "volatile", "register"]


def keyword_grammar() -> str:
    return ("params {\n  lookback = 1.\n  lookahead = 1.\n}\nident = /[a-z]+/.\n"
            "ignore = /\\s/.\n"
            "S -> " + " | ".join(f'"{k}" S' for k in KEYWORDS) + " | ident S | .\n")


def _words(rng, n):
    out = []
    while sum(len(w) + 1 for w in out) < n:
        r = rng.random()
        if r < 0.5:
            out.append(KEYWORDS[int(rng.integers(0, len(KEYWORDS)))])
        else:  # identifiers, including keyword prefixes and extensions
            k = KEYWORDS[int(rng.integers(0, len(KEYWORDS)))]
            out.append(k[:int(rng.integers(1, len(k) + 1))] + "".join(
                chr(97 + int(x)) for x in rng.integers(0, 26, size=int(rng.integers(0, 3)))))
    return " ".join(out).encode()


def test_keyword_grammar_dfa_in_l2():
    from alpacc_amd.gen.tables import generate
    from alpacc_amd.runtime import Context
    b = generate(keyword_grammar(), name="keywords")
    lx = b.lexer
    assert lx.num_states * lx.num_classes > 4096  # larger than lex_emit's LDS table
    rng = np.random.default_rng(9)
    docs, derived = [], []
    for n in (5, 100, 3000, 70_000, 400_000):
        d = _words(rng, n)
        derived.append(len(docs))
        docs.append(d)
        bad = bytearray(d)
        bad[int(rng.integers(0, len(bad)))] = ord("#")  # no token holds '#'
        docs.append(bytes(bad))
    ctx = Context(b)
    try:
        for d in docs[:4]:  # the lexer alone too
            from oracle.c import coracle
            want = coracle.Oracle(b).lex(d)
            got = ctx.lex_int(d)
            assert (got is None) == (want is None)
            if want is not None:
                assert all(np.array_equal(x, y) for x, y in zip(got, want))
    finally:
        ctx.close()
    _check(b, docs, "keywords", derived=derived)


@pytest.mark.parametrize("g", ["json", "sexp"])
def test_tables_forced_to_l2(g):
    from alpacc_amd import datagen
    from alpacc_amd.gen.tables import generate
    from alpacc_amd.runtime import Context
    from oracle.c import coracle
    b = generate(os.path.join(GRAMMARS, f"{g}.alp"))
    gen = {"json": datagen.json_doc, "sexp": datagen.sexp_doc}[g]
    docs = []
    for n in (2, 2049, 65_537, 1 << 20):
        d = gen(n, seed=800 + n)
        docs.append(d.tobytes())
        docs.append(datagen.flip_bytes(d, seed=n, block=max(16, n // 2)).tobytes())
    orc = coracle.Oracle(b)
    want = [orc.parse(d) for d in docs]
    ctx = Context(b)
    try:
        ctx.set_tuning(lexer_tables_in_lds=0)
        assert ctx.tuning()["lexer_tables_in_lds"] == 0
        for d, w in zip(docs, want):
            got = ctx.parse_int(d)
            assert (got is None) == (w is None)
            if w is not None:
                assert np.array_equal(got[0], w[0]) and np.array_equal(got[1], w[1])
                assert np.array_equal(got[2].reshape(-1), w[2])
        got = ctx.parse_int_batch(docs)
        for r, w in zip(got, want):
            assert (r is None) == (w is None)
            if w is not None:
                assert np.array_equal(r[0], w[0]) and np.array_equal(r[2].reshape(-1), w[2])
    finally:
        ctx.close()
