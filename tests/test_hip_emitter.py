"""The steps of the Haskell `hip` backend's parser tables (alpacc_amd/gen/hs/Alpacc/Generator/Hip/
Parser.hs), run on the reference's own LLP table layout: the open-addressing array of padded keys
with spans into the flattened brackets and productions (src/Alpacc/Encode.hs:353-375,
HashTable.hs:38-49, restated by alpacc_amd/gen/tables.py). Decoding the left-bracket bit at
bracket width - 1, taking the keys in ascending order, one entry per distinct (brackets,
productions) and the dense or hashed window index must give exactly the device tables that the
Python generator builds from the LLP homomorphisms directly."""
import os

import numpy as np
import pytest

from alpacc_amd.gen.tables import (DENSE_MAX, HASH_EMPTY, LEFT_BIT, MISSING, device_hash_slot,
                                   generate)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GRAMMARS = os.path.join(ROOT, "alpacc_amd", "grammars")


def hip_parser_steps(px, hashed=None):
    """Parser.hs, step for step, on the reference layout held in ``px``. Parser.hs hashes beyond
    DENSE_MAX windows; ``hashed`` forces the choice, so small grammars test the hashing too."""
    nT = 1 + max(px.start_terminal, px.end_terminal)
    nbs = px.bracket_bits
    flat_st = [int(v) for v in px.flat_stacks]
    flat_pr = [int(v) for v in px.flat_productions]
    keyed = {}
    for used, key, (s0, s1, p0, p1) in zip(px.hash_used, px.hash_keys, px.hash_vals):
        if not used:
            continue
        brackets = []
        for b in flat_st[s0:s1]:
            left = (b >> (nbs - 1)) & 1
            s = b & ~(1 << (nbs - 1))
            brackets.append(s | LEFT_BIT if left else s)
        keyed[tuple(int(t) for t in key)] = (tuple(brackets), tuple(flat_pr[p0:p1]))
    entries, stacks, prods, key_entry, seen = [], [], [], [], {}
    for key in sorted(keyed):
        sig = keyed[key]
        if sig not in seen:
            seen[sig] = len(seen)
            entries.append((len(stacks), len(stacks) + len(sig[0]),
                            len(prods), len(prods) + len(sig[1])))
            stacks.extend(sig[0])
            prods.extend(sig[1])
        di = 0
        for t in key:
            di = di * nT + t
        key_entry.append((di, seen[sig]))
    windows = nT ** (px.q + px.k)
    out = {"num_terminals": nT, "pair_entry": entries, "stacks": stacks, "productions": prods}
    if hashed is None:
        hashed = windows > DENSE_MAX
    if hashed:
        size = 16
        while size < 2 * len(key_entry):
            size *= 2
        bits = size.bit_length() - 1
        slots = {}
        for di, e in key_entry:
            h = device_hash_slot(di, bits)
            while h in slots:
                h = (h + 1) & (size - 1)
            slots[h] = (di, e)
        out["hash_keys"] = [slots.get(h, (HASH_EMPTY, MISSING))[0] for h in range(size)]
        out["hash_vals"] = [slots.get(h, (HASH_EMPTY, MISSING))[1] for h in range(size)]
        out["pair_index"] = []
    else:
        index = dict(key_entry)
        out["pair_index"] = [index.get(i, MISSING) for i in range(windows)]
    return out


def _check(px, hashed):
    got = hip_parser_steps(px, hashed)
    assert got["num_terminals"] == px.num_terminals
    assert np.array_equal(np.array(got["pair_entry"], dtype=np.uint32).reshape(-1, 4), px.pair_entry)
    assert got["stacks"] == [int(v) for v in px.stacks]
    assert got["productions"] == [int(v) for v in px.productions]
    if hashed:
        assert got["pair_index"] == [] and px.pair_index.shape[0] == 0
        assert got["hash_keys"] == [int(v) for v in px.pair_hash_keys]
        assert got["hash_vals"] == [int(v) for v in px.pair_hash_vals]
    else:
        assert got["pair_index"] == [int(v) for v in px.pair_index]


@pytest.mark.parametrize("grammar", ["json", "sexp", "arithmetic", "llp01", "paper_grammar"])
def test_reference_layout_gives_device_tables(grammar):
    path = os.path.join(GRAMMARS, grammar + ".alp")
    if not os.path.exists(path):
        pytest.skip(f"{grammar}.alp not shipped")
    _check(generate(path).parser, hashed=False)
    _check(generate(path, pair_table="hash").parser, hashed=True)

