"""The artifact's `test` entry on the C ABI (alpacc_test, futhark/test.fut:31-180) produces the
wire format of the oracle's encoders for the same inputs: lexer+parser, lexer-only and parser-only
bundles, including rejected and empty inputs and an input longer than the reference's 16 MiB
chunk."""
import os

import numpy as np
import pytest

from conftest import GRAMMARS

pytestmark = pytest.mark.gpu


def _ctx(g, mode):
    from alpacc_amd.gen.tables import generate
    from alpacc_amd.runtime import Context
    b = generate(os.path.join(GRAMMARS, f"{g}.alp"), mode=mode)
    return b, Context(b)


def test_lexer_parser_entry():
    from alpacc_amd import datagen
    from oracle import sequential, wire
    b, ctx = _ctx("arithmetic", "both")
    try:
        inputs = [b"12+(3*4)", b"(", b"a", b"", b" ", b"1+2*3-4",
                  datagen.arith_doc(5000, seed=3).tobytes()]
        got = ctx.test(wire.encode_inputs(inputs))
        want = wire.encode_tree_outputs([sequential.parse(b, s) for s in inputs])
        assert got == want
    finally:
        ctx.close()


def test_lexer_entry():
    from alpacc_amd import datagen
    from oracle import sequential, wire
    b, ctx = _ctx("json", "lexer")
    try:
        inputs = [b"[1, 2]", b"", b'{"a" ', b"[1.]", datagen.json_doc(20_000, seed=4).tobytes()]
        got = ctx.test(wire.encode_inputs(inputs))
        want = wire.encode_lexer_outputs([sequential.lex(b, s) for s in inputs])
        assert got == want
    finally:
        ctx.close()


def test_parser_only_entry():
    from oracle import sequential, wire
    from oracle.parser_fut import ParserFut
    from alpacc_amd.gen.tables import generate
    lb = generate(os.path.join(GRAMMARS, "arithmetic.alp"), mode="both")
    b, ctx = _ctx("arithmetic", "parser")
    try:
        pf = ParserFut(b.parser)
        seqs = []
        for s in (b"12+(3*4)", b"a*b", b"1+2+3"):
            seqs.append([k for k, _, _ in sequential.lex(lb, s)])
        seqs += [[0, 0], [], [99]]
        got = ctx.test(wire.encode_terminal_inputs(seqs))
        want = []
        for ts in seqs:
            if any(t >= b.parser.num_terminals for t in ts):
                want.append(None)
                continue
            r = pf.pre_productions_int(np.array(ts, dtype=np.int64))
            want.append(None if r is None else [int(x) for x in r])
        assert got == wire.encode_production_outputs(want)
    finally:
        ctx.close()


def test_entry_above_one_chunk():
    """17 MiB: the entry runs the reference's chunk loop (two chunks) like lex_int 16777216."""
    from alpacc_amd import datagen
    from alpacc_amd.runtime import REFERENCE_CHUNK
    from oracle import wire
    from oracle.c import coracle
    b, ctx = _ctx("json", "lexer")
    try:
        d = datagen.json_doc(17 << 20, seed=40).tobytes()
        got = ctx.test(wire.encode_inputs([d]))
        r = coracle.Oracle(b).lex(d, chunk=REFERENCE_CHUNK, wide=True)
        want = wire.encode_lexer_outputs([None if r is None else
                                          [tuple(int(v) for v in t) for t in zip(*r)]])
        assert got == want
    finally:
        ctx.close()
