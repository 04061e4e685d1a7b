"""The serving loop (alpacc_amd/serving.py): every request's result in the mix, served by
continuous batching over several worker streams, equals its result alone (parse_int through
alpacc_parse); rejected requests stay rejected. Tolerance 0: the outputs are integers."""
import os

import numpy as np
import pytest

from conftest import GRAMMARS

pytestmark = pytest.mark.gpu


def _same(a, b):
    if a is None or b is None:
        return a is None and b is None
    return all(np.array_equal(np.asarray(x), np.asarray(y)) for x, y in zip(a, b))


def test_mix_equals_alone():
    from alpacc_amd import datagen, serving
    from alpacc_amd.gen.tables import generate
    from alpacc_amd.runtime import Context
    b = generate(os.path.join(GRAMMARS, "json.alp"))
    sizes = serving.request_sizes(240, lo=256, hi=1 << 20, seed=3)
    rng = np.random.default_rng(4)
    datas = []
    for i, n in enumerate(sizes.tolist()):
        d = datagen.json_doc(max(int(n), 2), seed=500 + i)
        if rng.random() < 0.1:
            d = datagen.flip_bytes(d, seed=i, block=max(64, d.shape[0] // 2))
        datas.append(d)
    arrivals = serving.poisson_arrivals(len(datas), rate=4000.0, seed=5)
    print(f"serving {len(datas)} requests", flush=True)
    srv = serving.Server(b, workers=3, max_batch_bytes=2 << 20)
    try:
        reqs = serving.serve_poisson(srv, datas, arrivals)
    finally:
        srv.close()
    assert len(srv.batch_sizes) >= 2 and max(srv.batch_sizes) > 1
    ctx = Context(b)
    try:
        for r, d in zip(reqs, datas):
            alone = ctx.parse_int(d)
            assert _same(r.tree, alone), r.id
            assert (r.status == 1) == (alone is not None)
    finally:
        ctx.close()
    print(serving.latency_stats(reqs), flush=True)
