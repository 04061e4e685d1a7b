"""GPU parity on documents rejected by the bracket check alone (brackets_matches, parser.fut:174-184),
against the C restatement of parser.fut: every token stream up to a length (mismatches in cancelled
pairs and among the remaining brackets, test_bracket_corpus.py shows both occur), pairs that
straddle two 256-position groups (not cancelled: only the search sees them), and large documents
with a mismatch at the far end or far apart, through alpacc_parse_tokens, alpacc_parse and the
batched entry."""
import numpy as np
import pytest

import bracket_corpus as bc
from test_bracket_corpus import CASES, bundle_of

pytestmark = pytest.mark.gpu


def _tree(r):
    if r is None:
        return None
    par, node, span = r
    return (np.asarray(par).tolist(), np.asarray(node).tolist(),
            np.asarray(span).reshape(-1).tolist())


@pytest.mark.parametrize("name,max_len", [(c[0], c[1]) for c in CASES])
def test_short_streams(name, max_len):
    from alpacc_amd.runtime import Context
    from oracle.c import coracle
    b = bundle_of(name)
    orc = coracle.Oracle(b)
    ctx = Context(b)
    try:
        accepted = 0
        for kinds in bc.short_streams(b, max_len):
            t = np.arange(kinds.shape[0], dtype=np.uint32)
            want = _tree(orc.parse_tokens(kinds, t.astype(np.int64), t.astype(np.int64) + 1))
            assert _tree(ctx.parse_tokens(kinds, t, t + 1)) == want, kinds.tolist()
            accepted += want is not None
        print(f"{name}: {accepted} accepted", flush=True)
    finally:
        ctx.close()


def _docs_parity(b, docs, label):
    from alpacc_amd.runtime import Context
    from oracle.c import coracle
    orc = coracle.Oracle(b)
    ctx = Context(b)
    try:
        want = [_tree(orc.parse(d)) for d in docs]
        for i, d in enumerate(docs):
            assert _tree(ctx.parse_int(d)) == want[i], (label, i, len(d))
        got = ctx.parse_int_batch(docs)
        for i in range(len(docs)):
            assert _tree(got[i]) == want[i], (label, i, "batch")
        return want
    finally:
        ctx.close()


def test_group_boundaries():
    """Faulty pairs on both sides of the 256- and 1024-position boundaries, between valid
    documents in the batch."""
    b = bundle_of("sexp")
    docs = []
    for k in list(range(248, 258)) + list(range(1016, 1026)) + list(range(4090, 4100)):
        docs += [bc.boundary_doc(k), bc.boundary_doc(k)[:-2]]
    want = _docs_parity(b, docs, "boundaries")
    assert [w is None for w in want] == [True, False] * (len(docs) // 2)


def test_large_documents():
    """300 KB S-expressions: a second expression after the document (a mismatch at the far end),
    before it, and two tokens exchanged far apart (mismatches across tiles and search levels)."""
    from alpacc_amd import datagen
    b = bundle_of("sexp")
    docs, labels = [], []
    for seed in (3, 4):
        d = datagen.sexp_doc(300_000, seed=seed).tobytes()
        docs += [d, d + b" b", b"b " + d]
        for d2, lab in bc.byte_swaps(b, d, seed, 60):
            docs.append(d2)
            labels.append(lab)
    print("swaps:", labels, flush=True)
    want = _docs_parity(b, docs, "large")
    assert want[0] is not None and want[1] is None and want[2] is None
    assert "remaining" in labels
