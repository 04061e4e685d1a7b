"""Parity at the benchmark workloads' own sizes: the HIP path against the C/OpenMP restatement of
futhark/lexer.fut + futhark/parser.fut (oracle/c), bit-exact, on the same seeded documents the
benchmark runs.

* JSON lex+parse of one 256 MiB document (seed 42), unchunked: parent, node and span of every
  node.
* JSON lexer alone (lexer-only numbering, ``alpacc futhark --lexer``) on one 512 MiB document:
  kind, start and end of every token.
* 64 MiB JSON documents with one byte flipped per MiB (four seeds), unchunked: verdict and tree.
* The whole batched workload: 4096 sexp documents totalling 2 GiB (log-normal lengths, median
  64 KiB, 1 KiB..16 MiB) in one ``alpacc_parse_batch`` call, every document's verdict and tree
  against the oracle (four tests of 1024 documents over one batched run).
* Serving: a request mix with sizes log-uniform from 256 B up to 64 MiB (the largest request is
  64 MiB), 10 % invalid, every result against the oracle.
"""
import os

import numpy as np
import pytest

from conftest import GRAMMARS

pytestmark = pytest.mark.gpu

SEED = 42


def _bundle(g, mode="both"):
    from alpacc_amd.gen.tables import generate
    return generate(os.path.join(GRAMMARS, f"{g}.alp"), mode=mode)


def _same_tree(got, want):
    if got is None or want is None:
        return got is None and want is None
    par, node, span = got
    return (np.array_equal(par, want[0]) and np.array_equal(node, want[1])
            and np.array_equal(np.asarray(span).reshape(-1), want[2]))


def _parse_resident(ctx, doc):
    """alpacc_parse of a document resident in HBM into device outputs; the host copy afterwards."""
    from alpacc_amd.runtime import DeviceBuffer
    n = int(doc.shape[0])
    d_in = DeviceBuffer(n, ctx.lib)
    d_in.upload(doc)
    cap = 2 * n + 16
    outs = [DeviceBuffer(4 * cap, ctx.lib), DeviceBuffer(4 * cap, ctx.lib),
            DeviceBuffer(8 * cap, ctx.lib)]
    try:
        rc, nn = ctx.parse_device(d_in.ptr, n, outs[0].ptr, outs[1].ptr, outs[2].ptr, cap)
        assert rc in (0, 1), rc
        if rc != 1:
            return None
        return (outs[0].download(np.uint32, nn), outs[1].download(np.uint32, nn),
                outs[2].download(np.uint32, 2 * nn))
    finally:
        for o in outs:
            o.free()
        d_in.free()


def test_json_lex_parse_256MiB_unchunked():
    from alpacc_amd import datagen
    from alpacc_amd.runtime import Context
    from oracle.c import coracle
    b = _bundle("json")
    doc = datagen.json_doc(256 << 20, seed=SEED)
    ctx = Context(b)
    try:
        got = _parse_resident(ctx, doc)
    finally:
        ctx.close()
    want = coracle.Oracle(b).parse(doc)
    assert want is not None and got is not None
    assert len(got[0]) == len(want[0])
    assert _same_tree(got, want)


def test_json_lexer_only_512MiB():
    from alpacc_amd import datagen
    from alpacc_amd.runtime import Context, DeviceBuffer
    from oracle.c import coracle
    b = _bundle("json", mode="lexer")
    n = 512 << 20
    doc = datagen.json_doc(n, seed=SEED)
    ctx = Context(b)
    d_in = DeviceBuffer(n, ctx.lib)
    outs = [DeviceBuffer(4 * n, ctx.lib) for _ in range(3)]
    try:
        d_in.upload(doc)
        rc, nt = ctx.lex_device(d_in.ptr, n, outs[0].ptr, outs[1].ptr, outs[2].ptr, n)
        got = tuple(o.download(np.uint32, nt) for o in outs)
    finally:
        for o in outs:
            o.free()
        d_in.free()
        ctx.close()
    want = coracle.Oracle(b).lex(doc)
    assert rc == 1 and want is not None
    assert nt == len(want[0])
    for g_, w_ in zip(got, want):
        assert np.array_equal(g_, w_)


@pytest.mark.parametrize("flip_seed", [0, 1, 2, 3])
def test_json_invalid_variants_64MiB_unchunked(flip_seed):
    """One random byte XORed per MiB (1 + rng(255)); the verdict comes from the oracle."""
    from alpacc_amd import datagen
    from alpacc_amd.runtime import Context
    from oracle.c import coracle
    b = _bundle("json")
    doc = datagen.flip_bytes(datagen.json_doc(64 << 20, seed=SEED), seed=flip_seed)
    ctx = Context(b)
    try:
        got = _parse_resident(ctx, doc)
    finally:
        ctx.close()
    want = coracle.Oracle(b).parse(doc)
    assert _same_tree(got, want), flip_seed


@pytest.fixture(scope="module")
def sexp_batch():
    """The batched workload run once: 4096 documents, 2 GiB, one alpacc_parse_batch call."""
    from alpacc_amd import datagen
    from alpacc_amd.runtime import Context, DeviceBuffer
    b = _bundle("sexp")
    total = 1 << 31
    sizes = datagen.sexp_batch_sizes(count=4096, total=total)
    assert int(sizes.sum()) == total and int(sizes.max()) <= 16 << 20
    ctx = Context(b)
    d_in = DeviceBuffer(total, ctx.lib)
    offsets = np.zeros(len(sizes), dtype=np.uint64)
    off = 0
    for i, sz in enumerate(sizes.tolist()):
        d_in.upload(datagen.sexp_doc(int(sz), seed=SEED + 1000 + i), offset=off)
        offsets[i] = off
        off += int(sz)
    lengths = sizes.astype(np.uint64)
    rc, need, _ = ctx.parse_batch_device(d_in.ptr, offsets, lengths, 0, 0, 0, 0)
    assert rc == -3
    cap = int(need)
    outs = [DeviceBuffer(4 * cap, ctx.lib), DeviceBuffer(4 * cap, ctx.lib),
            DeviceBuffer(8 * cap, ctx.lib)]
    rc, got, res = ctx.parse_batch_device(d_in.ptr, offsets, lengths, outs[0].ptr, outs[1].ptr,
                                          outs[2].ptr, cap)
    assert rc == 1 and got == cap
    d_in.free()
    yield b, ctx, sizes, res, outs
    for o in outs:
        o.free()
    ctx.close()


@pytest.mark.parametrize("part", [0, 1, 2, 3])
def test_sexp_batch_2GiB_every_document(sexp_batch, part):
    from alpacc_amd import datagen
    from oracle.c import coracle
    b, ctx, sizes, res, outs = sexp_batch
    orc = coracle.Oracle(b)
    for i in range(part * 1024, (part + 1) * 1024):
        doc = datagen.sexp_doc(int(sizes[i]), seed=SEED + 1000 + i)
        want = orc.parse(doc)
        r = res[i]
        if want is None:
            assert r["status"] == 0, i
            continue
        assert r["status"] == 1, i
        o, c = int(r["offset"]), int(r["count"])
        assert c == len(want[0]), i
        got = (outs[0].download(np.uint32, c, offset=4 * o),
               outs[1].download(np.uint32, c, offset=4 * o),
               outs[2].download(np.uint32, 2 * c, offset=8 * o))
        assert _same_tree(got, want), i


def test_serving_requests_up_to_64MiB():
    from alpacc_amd import datagen, serving
    from oracle.c import coracle
    b = _bundle("json")
    sizes = serving.request_sizes(48, lo=256, hi=64 << 20, seed=11)
    sizes[int(np.argmax(sizes))] = 64 << 20  # the largest request is the largest size allowed
    rng = np.random.default_rng(12)
    datas = []
    for i, n in enumerate(sizes.tolist()):
        d = datagen.json_doc(max(int(n), 2), seed=900 + i)
        if rng.random() < 0.1:
            d = datagen.flip_bytes(d, seed=i, block=min(1 << 20, d.shape[0]))
        datas.append(d)
    arrivals = serving.poisson_arrivals(len(datas), rate=400.0, seed=13)
    srv = serving.Server(b, workers=3)
    try:
        reqs = serving.serve_poisson(srv, datas, arrivals)
    finally:
        srv.close()
    orc = coracle.Oracle(b)
    for r, d in zip(reqs, datas):
        want = orc.parse(d)
        assert (r.status == 1) == (want is not None), r.id
        assert _same_tree(r.tree, want), r.id
