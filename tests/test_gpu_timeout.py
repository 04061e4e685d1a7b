"""The lexer's bounded look-back wait: when it gives up, every entry must report ALPACC_ERR_TIMEOUT
(-4) instead of tokens or trees built on a wrong prefix. The time-out flag is raised from the host
(alpacc_test_force_timeout) right after the call clears it, so every kernel still runs to its normal
end; then the hook is switched off and the same context must give the oracle's results again."""
import os

import numpy as np
import pytest

from conftest import GRAMMARS

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def json_ctx():
    from alpacc_amd.gen.tables import generate
    from alpacc_amd.runtime import Context
    b = generate(os.path.join(GRAMMARS, "json.alp"))
    c = Context(b)
    yield b, c
    c.close()


def _docs():
    from alpacc_amd import datagen
    return [datagen.json_doc(n, seed=300 + n).tobytes() for n in (100, 5000, 70_000, 300_000)]


def _expect_timeout(fn):
    with pytest.raises(RuntimeError, match=r"alpacc error -4"):
        fn()


def test_forced_timeout_every_entry(json_ctx):
    from oracle import sequential
    b, ctx = json_ctx
    docs = _docs()
    lib = ctx.lib
    print("forced time-out: single, chunked and batched entries", flush=True)
    assert lib.alpacc_test_force_timeout(ctx.handle, 1) == 0
    try:
        _expect_timeout(lambda: ctx.lex_int(docs[2]))
        _expect_timeout(lambda: ctx.parse_int(docs[2]))
        _expect_timeout(lambda: ctx.lex_int(docs[3], chunk_size=100_000))
        _expect_timeout(lambda: ctx.lex_int_batch(docs))
        _expect_timeout(lambda: ctx.parse_int_batch(docs))
        blob = bytearray(len(docs).to_bytes(8, "big"))
        for d in docs:
            blob += len(d).to_bytes(8, "big") + d
        _expect_timeout(lambda: ctx.test(bytes(blob)))
    finally:
        assert lib.alpacc_test_force_timeout(ctx.handle, 0) == 0
    # the hook off again: the same context gives the oracle's results
    for d in docs:
        want = sequential.parse(b, d)
        got = ctx.parse_int(d)
        assert got is not None and want is not None
        assert np.asarray(got[0]).tolist() == [w[0] for w in want]
    got = ctx.parse_int_batch(docs)
    for d, g in zip(docs, got):
        assert np.asarray(g[0]).tolist() == [w[0] for w in sequential.parse(b, d)]
