"""GPU parity of the batched entries (alpacc_lex_batch / alpacc_parse_batch): every document of a
batch must get exactly what lex_int / parse_int give it alone (the C restatement of lexer.fut +
parser.fut as the oracle), including empty, tiny, whitespace-only and invalid documents, documents
at tile boundaries, and many tiny documents sharing tiles and 64-element chunks."""
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


GEN = {"json": "json_doc", "arithmetic": "arith_doc", "sexp": "sexp_doc"}


def _docs(g, seed):
    from alpacc_amd import datagen
    gen = getattr(datagen, GEN[g])
    rng = np.random.default_rng(seed)
    docs = [b"", b" ", b"\n\n", b"(", b"1", b"a"]
    for n in (2, 3, 5, 100, 2047, 2048, 2049, 4095, 4096, 4097, 10_000, 65_537, 100_000):
        d = gen(n, seed=seed + n)
        docs.append(d.tobytes())
        docs.append(datagen.flip_bytes(d, seed=seed + n, block=max(16, n // 3)).tobytes())
    rng.shuffle(docs)
    return docs


def _lex_rows(r):
    return None if r is None else [tuple(int(v) for v in t) for t in zip(*r)]


def _tree(r):
    if r is None:
        return None
    par, node, span = r
    span = np.asarray(span).reshape(-1)
    return (np.asarray(par).tolist(), np.asarray(node).tolist(), span.tolist())


@pytest.mark.parametrize("g", ["json", "arithmetic", "sexp"])
def test_lex_batch_matches_single(ctxs, g):
    from oracle.c import coracle
    b, ctx = ctxs[g]
    orc = coracle.Oracle(b)
    docs = _docs(g, 11)
    print(f"lex batch {g}: {len(docs)} docs", flush=True)
    got = ctx.lex_int_batch(docs)
    for i, d in enumerate(docs):
        assert _lex_rows(got[i]) == _lex_rows(orc.lex(d)), (g, i, len(d))


@pytest.mark.parametrize("g", ["json", "arithmetic", "sexp"])
def test_parse_batch_matches_single(ctxs, g):
    from oracle.c import coracle
    b, ctx = ctxs[g]
    orc = coracle.Oracle(b)
    docs = _docs(g, 12)
    print(f"parse batch {g}: {len(docs)} docs", flush=True)
    got = ctx.parse_int_batch(docs)
    for i, d in enumerate(docs):
        assert _tree(got[i]) == _tree(orc.parse(d)), (g, i, len(d))


@pytest.mark.parametrize("g", ["sexp", "arithmetic"])
def test_many_tiny_documents(ctxs, g):
    """Thousands of documents of 1-300 bytes: several documents per wave tile of the lexer, per
    position tile of the parser and per 64-element chunk of the searches."""
    from alpacc_amd import datagen
    from oracle.c import coracle
    b, ctx = ctxs[g]
    orc = coracle.Oracle(b)
    gen = getattr(datagen, GEN[g])
    rng = np.random.default_rng(5)
    docs = []
    for i in range(3000):
        n = int(rng.integers(2, 300))
        d = gen(n, seed=1000 + i)
        if i % 7 == 3:
            d = datagen.flip_bytes(d, seed=i, block=max(8, n // 2))
        docs.append(d.tobytes())
    print(f"tiny docs {g}: {len(docs)}", flush=True)
    lex = ctx.lex_int_batch(docs)
    par = ctx.parse_int_batch(docs)
    for i, d in enumerate(docs):
        assert _lex_rows(lex[i]) == _lex_rows(orc.lex(d)), (g, i)
        assert _tree(par[i]) == _tree(orc.parse(d)), (g, i)


def test_batch_capacity_retry(ctxs):
    """A too-small output capacity reports the size needed; the retried call succeeds."""
    from alpacc_amd import datagen
    from oracle.c import coracle
    b, ctx = ctxs["json"]
    docs = [datagen.json_doc(5000, seed=s).tobytes() for s in range(3)]
    got = ctx.parse_int_batch(docs, cap=16)
    orc = coracle.Oracle(b)
    for i, d in enumerate(docs):
        assert _tree(got[i]) == _tree(orc.parse(d))


def test_dense_tokens_past_the_tile_estimate():
    """The batched expansion's grid covers an estimate of the key positions (a quarter of the input
    bytes at first); documents of one-byte tokens have four times as many, so a fresh context's
    first call expands most tiles in the striding pass. Both that call and the next (estimate
    grown) must give every document what it gets alone."""
    from alpacc_amd.gen.tables import generate
    from alpacc_amd.runtime import Context
    from oracle.c import coracle
    b = generate(os.path.join(GRAMMARS, "sexp.alp"))
    orc = coracle.Oracle(b)
    docs = []
    for i in range(20):
        k = 20_000 + 997 * i
        d = b"(" * k + b"a" + b")" * k
        if i == 7:
            d = d[:-1]  # unbalanced: rejected
        docs.append(d)
    ctx = Context(b)
    try:
        for call in range(2):
            got = ctx.parse_int_batch(docs)
            for i, d in enumerate(docs):
                assert _tree(got[i]) == _tree(orc.parse(d)), (call, i, len(d))
    finally:
        ctx.close()


@pytest.mark.parametrize("g", ["json", "sexp"])
def test_document_ends_at_every_lane_offset(ctxs, g):
    """Prefixes of one document ending at every offset of a lane's 32 bytes, packed back to back
    (so every document starts at another offset from a 16-byte boundary and ends inside a lane):
    each document's tokens, verdict and tree must equal the C restatement's."""
    from alpacc_amd import datagen
    from oracle.c import coracle
    b, ctx = ctxs[g]
    orc = coracle.Oracle(b)
    base = getattr(datagen, GEN[g])(6000, seed=32).tobytes()
    lengths = list(range(1, 41)) + list(range(2048 - 33, 2048 + 34)) + [4064 + r for r in range(33)]
    docs = [base[:n] for n in lengths]
    print(f"lane-offset ends {g}: {len(docs)} docs", flush=True)
    lex = ctx.lex_int_batch(docs)
    par = ctx.parse_int_batch(docs)
    verdicts = set()
    for i, d in enumerate(docs):
        want = orc.lex(d)
        verdicts.add(want is None)
        assert _lex_rows(lex[i]) == _lex_rows(want), (g, i, len(d))
        assert _tree(par[i]) == _tree(orc.parse(d)), (g, i, len(d))
    if g == "json":  # every prefix of an s-expression lexes
        assert verdicts == {True, False}, "the cuts should give both verdicts"
