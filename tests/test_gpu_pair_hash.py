"""GPU parity with the hashed device pair table: a grammar whose dense table would pass 2^26 key
windows (lookback 5, lookahead 6: 6^11 windows) and the three benchmark grammars with the hashed
table forced, single and batched, against the C restatement of lexer.fut + parser.fut (which
looks pairs up in the reference's own FNV table, not in the device's)."""
import os

import numpy as np
import pytest

from conftest import GRAMMARS
from test_gpu_fuzz_grammars import _check
from test_pair_hash import PAREN

pytestmark = pytest.mark.gpu


def _balanced(rng, n):
    """A random balanced parenthesis string of about n pairs, with random spacing."""
    out, depth, opened = [], 0, 0
    while opened < n or depth:
        if opened < n and (depth == 0 or rng.random() < 0.55):
            out.append("(")
            depth += 1
            opened += 1
        else:
            out.append(")")
            depth -= 1
        if rng.random() < 0.2:
            out.append(" ")
    return "".join(out).encode()


def test_large_window_grammar():
    from alpacc_amd.gen.tables import generate
    b = generate(PAREN, name="paren56")
    assert b.parser.pair_index.shape[0] == 0  # hashed
    rng = np.random.default_rng(3)
    docs, derived, invalid = [], [], []
    for n in (1, 2, 5, 40, 700, 30_000):
        d = _balanced(rng, n)
        derived.append(len(docs))
        docs.append(d)
        bad = bytearray(d)
        bad[int(rng.integers(0, len(bad)))] ^= 0x01  # "(" <-> ")": unbalanced, or a space
        docs.append(bytes(bad))
    docs += [b"(", b")(", b"(()", b"())("]
    invalid += [len(docs) - 4, len(docs) - 3, len(docs) - 2, len(docs) - 1]
    accepted = _check(b, docs, "paren56", derived=derived, invalid=invalid)
    assert accepted >= len(derived)


@pytest.mark.parametrize("g", ["json", "sexp", "arithmetic"])
def test_forced_hash_benchmark_grammars(g):
    from alpacc_amd import datagen
    from alpacc_amd.gen.tables import generate
    b = generate(os.path.join(GRAMMARS, f"{g}.alp"), pair_table="hash")
    gen = {"json": datagen.json_doc, "sexp": datagen.sexp_doc, "arithmetic": datagen.arith_doc}[g]
    docs = []
    for n in (2, 100, 4097, 70_000):
        try:
            d = gen(n, seed=600 + n)
        except RuntimeError:
            continue
        docs.append(d.tobytes())
        docs.append(datagen.flip_bytes(d, seed=n, block=max(16, n // 3)).tobytes())
    _check(b, docs, g)
