"""The endomorphism count e under the reference's "connected" closure (no GPU).

The reference composes only connected endomorphisms (ParallelLexing.hs:222-264: a worklist over
pairs (e, e') where e' can follow e, seeded by connectedTable, :266-283, which lets byte t' follow t
when some state that t reaches has a move on t'). The restated generator (alpacc_amd/gen/endo.py)
closes over every pair and identifies every all-dead composition with the dead endomorphism. Both
give the same set: a byte string with an unconnected neighbouring pair maps every state to dead,
so every other composition is the endomorphism of a connected string. This test checks it: a
breadth-first search over connected byte strings (extend a string ending in byte t by a byte t'
that connectedTable lets follow t) reaches exactly the non-trivial endomorphisms of the full
closure, for every bundled grammar, so e is not inflated by the full closure.
"""
import os

import numpy as np
import pytest

from conftest import GRAMMARS


def connected_closure(lx):
    d = lx.dfa.dfa
    m = d.n
    # byte endomorphisms as (state map, produce flags), as endo.py builds them; the search
    # composes such pairs
    width = m + 1
    fb, pb = [], []
    for b in range(256):
        f = np.zeros(width, dtype=np.int32)
        p = np.zeros(width, dtype=bool)
        for q in range(m):
            r = d.delta[q].get(b)
            if r is not None:
                f[q + 1] = r + 1
                p[q + 1] = (q, b) in lx.dfa.producing
        fb.append(f)
        pb.append(p)
    # connectedTable: t' may follow t when some state that t reaches has a move on t'
    reach = [set(int(x) for x in fb[b][1:] if x) for b in range(256)]
    follows = [[c for c in range(256) if any(d.delta[s - 1].get(c) is not None for s in reach[b])]
               for b in range(256)]
    # search over (endomorphism, last byte class); bytes with the same move function and the same
    # followers behave alike, so one representative byte per such class suffices
    rep = {}
    for b in range(256):
        if fb[b].any():
            rep.setdefault((fb[b].tobytes(), pb[b].tobytes(), tuple(follows[b])), b)
    reps = sorted(rep.values())
    cls_of = {}
    for b in range(256):
        if fb[b].any():
            cls_of[b] = rep[(fb[b].tobytes(), pb[b].tobytes(), tuple(follows[b]))]
    seen_f = set()
    seen = set()
    frontier = []
    for b in reps:
        key = (fb[b].tobytes() + pb[b].tobytes(), b)
        seen.add(key)
        seen_f.add(key[0])
        frontier.append((fb[b], pb[b], b))
    while frontier:
        nxt = []
        for f, p, b in frontier:
            for c in follows[b]:
                c = cls_of[c]
                nf = fb[c][f]
                if not nf.any():
                    continue
                np_ = pb[c][f]
                kf = nf.tobytes() + np_.tobytes()
                key = (kf, c)
                if key in seen:
                    continue
                seen.add(key)
                seen_f.add(kf)
                nxt.append((nf, np_, c))
        frontier = nxt
    return seen_f


@pytest.mark.parametrize("g", ["json", "sexp", "arithmetic"])
def test_connected_closure_gives_the_same_endomorphisms(g):
    from alpacc_amd.gen.tables import generate
    b = generate(os.path.join(GRAMMARS, f"{g}.alp"))
    lx = b.lexer
    reached = connected_closure(lx)
    et = lx.endo
    d = lx.dfa.dfa
    # the full closure keeps state maps (rows 2..e-1 of functions): compare those, and the count
    # of distinct (map, produce) endomorphisms
    maps_reached = {k[: (d.n + 1) * 4] for k in reached}
    maps_full = {et.functions[i].astype(np.int32).tobytes() for i in range(2, et.num_endos)}
    assert maps_reached == maps_full
    assert len(reached) == et.num_endos - 2, (g, len(reached), et.num_endos)
