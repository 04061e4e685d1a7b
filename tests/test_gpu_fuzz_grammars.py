"""GPU parity on random grammars (src/Alpacc/Random.hs and tests/testfuthark.sh:95-156 in the
reference: random grammars that are LLP(q, k), parsed by every implementation and compared) and on
the reference's own LLP grammars grammars/llp01.alp (lookback 0), paper_grammar.alp and list.alp:
alpacc_parse and the batched entry against the C restatement of lexer.fut + parser.fut, which
test_fuzz_grammars.py pins to the sequential LLP parser on the same grammars."""
import os

import numpy as np
import pytest

from conftest import GRAMMARS
from random_grammar import corrupt, sentence
from test_fuzz_grammars import kept_grammars

pytestmark = pytest.mark.gpu


def _tree(r):
    if r is None:
        return None
    par, node, span = r
    return (np.asarray(par).tolist(), np.asarray(node).tolist(),
            np.asarray(span).reshape(-1).tolist())


def _check(bundle, docs, label, derived=(), invalid=()):
    """GPU (single and batched) against the C restatement on every document; the documents at
    `derived` are sentences of the grammar and must be accepted, those at `invalid` rejected."""
    from alpacc_amd.runtime import Context
    from oracle.c import coracle
    orc = coracle.Oracle(bundle)
    ctx = Context(bundle)
    try:
        want = [_tree(orc.parse(d)) for d in docs]
        for i in derived:
            assert want[i] is not None, (label, i, "derived sentence rejected")
        for i in invalid:
            assert want[i] is None, (label, i, "invalid document accepted")
        for i, d in enumerate(docs):
            assert _tree(ctx.parse_int(d)) == want[i], (label, i, len(d))
        got = ctx.parse_int_batch(docs)
        for i in range(len(docs)):
            assert _tree(got[i]) == want[i], (label, i, "batch")
        return sum(w is not None for w in want)
    finally:
        ctx.close()


@pytest.mark.parametrize("part", range(4))
def test_random_grammars(part):
    accepted = 0
    for seed, q, k, b, prods in kept_grammars(40)[part::4]:
        docs = []
        for target in (1, 4, 16, 60, 1000):
            d = sentence(prods, seed * 100 + target, target)
            docs += [d, corrupt(d, seed + target)]
        # derived sentences must be accepted, the empty one aside (the lexer rejects n = 0)
        accepted += _check(b, docs, (seed, q, k),
                           derived=[i for i in range(0, len(docs), 2) if docs[i]])
    print(f"part {part}: {accepted} accepted", flush=True)
    assert accepted > 0


def test_random_grammars_large():
    """Grammars whose sentences grow: documents of about 150 KB (many tiles and chunks)."""
    done = 0
    for seed, q, k, b, prods in kept_grammars(80):
        if done == 6:
            break
        if len(sentence(prods, seed, 2000)) < 1000:
            continue
        d = sentence(prods, seed + 7, 60_000)
        assert _check(b, [d, corrupt(d, seed)], (seed, q, k, "large"), derived=[0]) >= 1
        print(f"grammar {seed} q={q} k={k}: {len(d)} bytes", flush=True)
        done += 1
    assert done == 6


def _reference_docs(g, rng):
    if g == "llp01":  # S -> "[" S "]" S | .
        def s(budget):
            # balanced brackets, at most `budget` pairs: a random walk that closes at the end
            out, depth = [], 0
            for _ in range(budget):
                if depth and rng.random() < 0.5:
                    out.append("]")
                    depth -= 1
                else:
                    out.append("[")
                    depth += 1
            return "".join(out) + "]" * depth
        return [s(int(rng.integers(1, 60_000))).encode() for _ in range(6)] + [b"[[]", b"]["]
    if g == "paper_grammar":  # E -> T E_, E_ -> "+" T E_ | ., T -> a | "[" E "]"
        def e(depth):
            ts = []
            for _ in range(int(rng.integers(1, 6))):
                ts.append(str(int(rng.integers(0, 999))) if depth > 8 or rng.random() < 0.6
                          else "[" + e(depth + 1) + "]")
            return " + ".join(ts)
        return [e(0).encode() for _ in range(6)] + [b"1+", b"[1"]
    # list: V -> "[" EL0 "]", EL0 -> | atom EL1, EL1 -> | "," atom EL1
    return ([("[" + ",".join(["a"] * int(rng.integers(0, 5000))) + "]").encode() for _ in range(6)]
            + [b"[a,]", b"[,a]"])


@pytest.mark.parametrize("g", ["llp01", "paper_grammar", "list"])
def test_reference_llp_grammars(g):
    from alpacc_amd.gen.tables import generate
    b = generate(os.path.join(GRAMMARS, f"{g}.alp"))
    docs = _reference_docs(g, np.random.default_rng(3))
    assert _check(b, docs, g, derived=range(6), invalid=range(6, 8)) == 6
