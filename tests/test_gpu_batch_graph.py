"""The batched lex + parse as one queue with no host synchronisation before its end, replayed as
a HIP graph (graph_mode 1: captured once a call's plan repeats): every replay, with eager calls and
synchronisations between the replays, must give what the directly queued call gives, and both
what the C restatement of lexer.fut + parser.fut gives. Documents change between replays (same
lengths, other bytes), so a graph that kept any data of its capture would fail. A batch with more
brackets than the scratch estimate runs again with the exact size."""
import os

import numpy as np
import pytest

from conftest import GRAMMARS

pytestmark = pytest.mark.gpu


def _tree(r):
    if r is None:
        return None
    par, node, span = r
    return (np.asarray(par).tolist(), np.asarray(node).tolist(),
            np.asarray(span).reshape(-1).tolist())


def _bundle(g):
    from alpacc_amd.gen.tables import generate
    return generate(os.path.join(GRAMMARS, f"{g}.alp"))


def _batches(g, lengths, seeds):
    from alpacc_amd import datagen
    gen = {"json": datagen.json_doc, "sexp": datagen.sexp_doc, "arithmetic": datagen.arith_doc}[g]
    out = []
    for s in seeds:
        docs = []
        for i, n in enumerate(lengths):
            d = gen(n, seed=s * 1000 + i)
            if i % 5 == 2:
                d = datagen.flip_bytes(d, seed=s + i, block=max(16, n // 2))
            docs.append(d.tobytes())
        out.append(docs)
    return out


@pytest.mark.parametrize("g", ["sexp", "json"])
def test_graph_replays_equal_direct_queue(g):
    from alpacc_amd.runtime import Context
    from oracle.c import coracle
    b = _bundle(g)
    orc = coracle.Oracle(b)
    lengths = [3, 100, 2047, 2049, 40_000, 5, 300_000, 64, 1_000, 70_000]
    batches = _batches(g, lengths, seeds=(1, 2))
    want = [[_tree(orc.parse(d)) for d in docs] for docs in batches]
    direct = Context(b)
    direct.set_tuning(graph_mode=0)
    graph = Context(b)
    assert graph.tuning()["graph_mode"] == 1
    cap = 8 * sum(lengths) + 64
    try:
        for rep in range(6):
            k = rep % 2  # the two batches alternate: same plan, other bytes
            print(f"{g}: replay {rep}, batch {k}", flush=True)
            got_d = direct.parse_int_batch(batches[k], cap=cap)
            got_g = graph.parse_int_batch(batches[k], cap=cap)
            for i in range(len(lengths)):
                assert _tree(got_d[i]) == want[k][i], (rep, i)
                assert _tree(got_g[i]) == want[k][i], (rep, i)
            # eager work on the graph's context between replays
            single = graph.parse_int(batches[1 - k][4])
            assert _tree(single) == want[1 - k][4]
        t = graph.tuning()
        assert t["graph_captures"] == 1, t
        assert t["graph_launches"] == 5, t
        assert direct.tuning()["graph_launches"] == 0
    finally:
        direct.close()
        graph.close()


@pytest.mark.parametrize("g", ["arithmetic", "json"])
def test_bracket_estimate_exceeded_reruns(g):
    """With a first bracket estimate far below what the batch needs (1 per KiB of input) the
    first pass writes nothing and the call runs again sized exactly; the estimate grows."""
    from alpacc_amd.runtime import Context
    from oracle.c import coracle
    b = _bundle(g)
    orc = coracle.Oracle(b)
    docs = _batches(g, [10, 5_000, 200_000, 3, 70_000], seeds=(7,))[0]
    if g == "arithmetic":
        for depth in (10, 3000):
            docs.append(("(" * depth + "a+b*(c-d)" + ")" * depth + "+1").encode())
        docs.append(b"(" * 50 + b"1")  # rejected
    ctx = Context(b)
    try:
        ctx.set_tuning(brackets_per_kib=1)
        got = ctx.parse_int_batch(docs)
        nbytes = sum(len(d) for d in docs)
        assert ctx.last_counts()["brackets"] > nbytes / 1024
        assert ctx.tuning()["brackets_per_kib"] > 1
        for i, d in enumerate(docs):
            assert _tree(got[i]) == _tree(orc.parse(d)), i
        got2 = ctx.parse_int_batch(docs)  # now sized from the start
        for i, d in enumerate(docs):
            assert _tree(got2[i]) == _tree(orc.parse(d)), i
    finally:
        ctx.close()
