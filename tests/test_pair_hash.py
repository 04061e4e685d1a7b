"""The hashed device pair table (no GPU): grammars whose dense table would pass 2^26 key windows
get a hashed one, and every key window of the LLP table is found by the device's probe rule with
its dense entry, while windows outside the table are not. The same rule is checked on the three
benchmark grammars with the hashed table forced, against their dense tables."""
import os

import numpy as np
import pytest

from conftest import GRAMMARS

PAREN = ('params {\n  lookback = 5.\n  lookahead = 6.\n}\nignore = /\\s/.\n'
         'S -> "(" S ")" S | .\n')


def probe(px, di):
    """The device's lookup (pk_hash_lookup): the entry id of dense window index di, or None."""
    from alpacc_amd.gen import tables as T
    keys, vals = px.pair_hash_keys, px.pair_hash_vals
    H = keys.shape[0]
    bits = H.bit_length() - 1
    h = T.device_hash_slot(int(di), bits)
    for _ in range(H):
        k = int(keys[h])
        if k == int(di):
            return int(vals[h])
        if k == T.HASH_EMPTY:
            return None
        h = (h + 1) & (H - 1)
    return None


def dense_index(px, key):
    di = 0
    for t in key:
        di = di * px.num_terminals + int(t)
    return di


@pytest.mark.parametrize("g", ["json", "sexp", "arithmetic"])
def test_forced_hash_equals_dense(g):
    from alpacc_amd.gen import tables as T
    path = os.path.join(GRAMMARS, f"{g}.alp")
    d = T.generate(path).parser
    h = T.generate(path, pair_table="hash").parser
    assert h.pair_index.shape[0] == 0 and h.pair_hash_keys.shape[0] >= 16
    assert np.array_equal(d.pair_entry, h.pair_entry)
    present = np.nonzero(d.pair_index != T.MISSING)[0]
    for di in present.tolist():
        assert probe(h, di) == int(d.pair_index[di]), di
    rng = np.random.default_rng(1)
    absent = np.nonzero(d.pair_index == T.MISSING)[0]
    for di in rng.choice(absent, size=min(2000, absent.shape[0]), replace=False).tolist():
        assert probe(h, di) is None


def test_large_window_grammar_is_hashed():
    """lookback 5, lookahead 6 over 6 terminals: 6^11 windows, too many for a dense table."""
    from alpacc_amd.gen import tables as T
    b = T.generate(PAREN, name="paren56")
    px = b.parser
    assert px.num_terminals ** 11 > T.DENSE_MAX
    assert px.pair_index.shape[0] == 0
    H = px.pair_hash_keys.shape[0]
    used = int((px.pair_hash_keys != np.uint64(T.HASH_EMPTY)).sum())
    assert H >= 2 * used and H & (H - 1) == 0
    # every key of the LLP table (padded as the reference pads them) is found with its entry
    keys = [tuple(int(v) for v in k) for k in px.hash_keys[px.hash_used]]
    assert len(keys) == used
    ids = {}
    for key in keys:
        e = probe(px, dense_index(px, key))
        assert e is not None, key
        ids.setdefault(e, 0)
    assert len(ids) == px.pair_entry.shape[0]
    with pytest.raises(ValueError):
        T.generate(PAREN, pair_table="dense")
