"""The C restatement of lexer.fut + parser.fut (oracle/c) against the sequential LLP parser
(llpParse + flattenTree, LLP.hs:600-745) for lookback and lookahead from 1 to 3, on the LLP paper
grammar: sentences and one-byte corruptions of them. The GPU parity test for these settings
(test_gpu_lookahead.py) compares the HIP path with the C restatement."""
import numpy as np
import pytest

from lookahead_grammar import QK, sentence, source

TERMINAL = 0x80000000


def tuples(r):
    """(parent, is_terminal, id, start, end) rows of a (parent, node, span) result."""
    if r is None:
        return None
    par, node, span = r
    out = []
    for p, nd, sp in zip(np.asarray(par).tolist(), np.asarray(node).tolist(),
                         np.asarray(span).reshape(-1, 2).tolist()):
        if nd & TERMINAL:
            out.append((p, True, nd & 0x7FFFFFFF, sp[0], sp[1]))
        else:
            out.append((p, False, nd, 0, 0))
    return out


@pytest.mark.parametrize("q,k", QK)
def test_c_oracle_matches_sequential(q, k):
    from alpacc_amd.gen.tables import generate
    from oracle import sequential
    from oracle.c import coracle
    b = generate(source(q, k))
    orc = coracle.Oracle(b)
    rng = np.random.default_rng(q * 10 + k)
    accepted = rejected = 0
    for i in range(12):
        doc = sentence(int(rng.integers(1, 400)), seed=100 * q + 10 * k + i)
        bad = bytearray(doc)
        bad[int(rng.integers(0, len(doc)))] = ord("[]+a "[int(rng.integers(0, 5))])
        for d in (doc, bytes(bad)):
            want = sequential.parse(b, d)
            assert tuples(orc.parse(d)) == want, (q, k, d)
            accepted += want is not None
            rejected += want is None
    assert accepted >= 12 and rejected >= 1, (accepted, rejected)
