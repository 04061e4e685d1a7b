"""GPU parity for lookback / lookahead beyond 1 (key windows of 2 to 6 terminals; pair indexes
held as u8 in LDS and as u32 in global memory) on the LLP paper grammar: alpacc_parse and the
batched entry against the C restatement of lexer.fut + parser.fut, which test_lookahead_oracle.py
pins to the sequential LLP parser for the same settings."""
import numpy as np
import pytest

from lookahead_grammar import QK, sentence, source

pytestmark = pytest.mark.gpu


def _tree(r):
    if r is None:
        return None
    par, node, span = r
    return (np.asarray(par).tolist(), np.asarray(node).tolist(),
            np.asarray(span).reshape(-1).tolist())


@pytest.mark.parametrize("q,k", QK)
def test_lookahead_parse(q, k):
    from alpacc_amd.gen.tables import generate
    from alpacc_amd.runtime import Context
    from oracle.c import coracle
    b = generate(source(q, k))
    orc = coracle.Oracle(b)
    ctx = Context(b)
    try:
        docs = []
        for i, n in enumerate((1, 7, 300, 4_097, 70_000, 600_000)):
            d = sentence(n, seed=1000 * q + 100 * k + i)
            bad = bytearray(d)
            bad[len(d) // 3] = ord("]")
            docs += [d, bytes(bad)]
        docs += [b"", b"a", b"[", b"[a]]", b"a+"]
        print(f"q={q} k={k}: {len(docs)} documents", flush=True)
        for i, d in enumerate(docs):
            assert _tree(ctx.parse_int(d)) == _tree(orc.parse(d)), (q, k, i, len(d))
        got = ctx.parse_int_batch(docs)
        for i, d in enumerate(docs):
            assert _tree(got[i]) == _tree(orc.parse(d)), (q, k, i, "batch")
    finally:
        ctx.close()
