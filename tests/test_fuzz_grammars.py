"""Random-grammar fuzzing on the CPU (the role of tests/testfuthark.sh:95-156 with
src/Alpacc/Random.hs): random grammars kept when the LLP(q, k) construction accepts them, for
lookback / lookahead (1,1), (0,1), (2,2) and (3,3); sentences from random derivations and one-token
corruptions; every derived sentence must be accepted, and the C restatement of lexer.fut + parser.fut
must equal the sequential LLP parser (llpParse + flattenTree, LLP.hs:600-745) on every one, verdicts
included."""
import numpy as np
import pytest

from random_grammar import corrupt, grammar, sentence

TERMINAL = 0x80000000
QK = [(1, 1), (0, 1), (2, 2), (3, 3)]


def tuples(r):
    if r is None:
        return None
    par, node, span = r
    out = []
    for p, nd, sp in zip(np.asarray(par).tolist(), np.asarray(node).tolist(),
                         np.asarray(span).reshape(-1, 2).tolist()):
        out.append((p, True, nd & 0x7FFFFFFF, sp[0], sp[1]) if nd & TERMINAL else (p, False, nd, 0, 0))
    return out


def kept_grammars(count: int):
    """The first `count` (seed, q, k, source, productions) that generate."""
    from alpacc_amd.gen.llp import NotLLP
    from alpacc_amd.gen.tables import generate
    out = []
    seed = 0
    while len(out) < count:
        for q, k in QK:
            g = grammar(seed, q, k)
            if g is None:
                continue
            try:
                b = generate(g[0])
            except NotLLP:
                continue
            out.append((seed, q, k, b, g[1]))
        seed += 1
    return out[:count]


@pytest.mark.parametrize("part", range(4))
def test_c_oracle_matches_sequential_on_random_grammars(part):
    from oracle import sequential
    from oracle.c import coracle
    accepted = rejected = 0
    for seed, q, k, b, prods in kept_grammars(40)[part::4]:
        orc = coracle.Oracle(b)
        for target in (1, 4, 16, 60):
            d = sentence(prods, seed * 100 + target, target)
            # a sentence derived from the grammar is in its language: the tables must accept it,
            # except the empty one, which the lexer rejects (lex_int_flag on n = 0 ends in the
            # identity state, which is not accepting, lexer.fut:158-172)
            assert (sequential.parse(b, d) is not None) == (d != b""), (seed, q, k, d)
            for doc in (d, corrupt(d, seed + target)):
                want = sequential.parse(b, doc)
                assert tuples(orc.parse(doc)) == want, (seed, q, k, doc)
                accepted += want is not None
                rejected += want is None
    assert accepted > 0 and rejected > 0, (accepted, rejected)
