"""The endomorphism closure of alpacc_amd/gen/endo.py composes every pair of reachable
endomorphisms. The reference composes only the pairs its "connected" worklist visits
(src/Alpacc/Lexer/ParallelLexing.hs:157-264): it starts from the byte endomorphisms, links e to e'
when some state reached by a byte of e has a move on a byte of e' (connectedTable, :266-290), and
after composing a linked pair gives the result the predecessors of its left operand and the
successors of its right one (preSets / postSets, :207-226). Both closures must reach the same
non-dead endomorphisms, so that the full closure inflates neither e nor the composition table.
The worklist is restated here step by step and run on the same lexer DFA. (json.alp, 29 states:
1225 non-dead endomorphisms either way, with 1.5 M compositions in the worklist; that run takes
about two minutes, so it is left out here.)"""
import os

import numpy as np
import pytest

from conftest import GRAMMARS


def _byte_endo(ld, b):
    d = ld.dfa
    f = np.zeros(d.n + 1, dtype=np.int32)
    p = np.zeros(d.n + 1, dtype=bool)
    for q in range(d.n):
        r = d.delta[q].get(b)
        if r is not None:
            f[q + 1] = r + 1  # 0 is the dead state
            p[q + 1] = (q, b) in ld.producing
    return f, p


def connected_worklist(ld) -> set:
    """Keys (move function, produce flags) of the non-dead endomorphisms the reference's worklist
    reaches, starting from the DFA's alphabet."""
    d = ld.dfa
    alphabet = sorted({b for q in range(d.n) for b in d.delta[q]})
    key = lambda f, p: f.tobytes() + p.tobytes()
    ids, vals, t_to_e = {}, {}, {}
    nxt = 2  # 0 identity, 1 dead
    for b in alphabet:
        f, p = _byte_endo(ld, b)
        k = key(f, p)
        if k not in ids:
            ids[k] = nxt
            vals[nxt] = (f, p)
            nxt += 1
        t_to_e[b] = ids[k]
    # connectedTable, then mapMapSet: on equal endomorphisms Map.mapKeys keeps the set of the
    # greatest byte
    connected = {}
    for t in alphabet:
        reached = {d.delta[q][t] for q in range(d.n) if t in d.delta[q]}
        connected[t_to_e[t]] = {t_to_e[t2] for s in reached for t2 in d.delta[s]}
    inverse = {}
    for e, succ in connected.items():
        for e2 in succ:
            inverse.setdefault(e2, set()).add(e)
    done = set()
    work = {e: set(s) for e, s in connected.items()}
    while True:
        keys = [k for k, s in work.items() if s]
        if not keys:
            break
        e = min(keys)  # popElement: the smallest pair
        e2 = min(work[e])
        work[e].discard(e2)
        if (e, e2) in done:
            continue
        done.add((e, e2))
        (fa, pa), (fb, pb) = vals[e], vals[e2]
        f, p = fb[fa], pb[fa]  # e then e2
        k = key(f, p)
        if k not in ids:
            ids[k] = nxt
            vals[nxt] = (f, p)
            nxt += 1
        e3 = ids[k]
        for x in inverse.get(e, set()) - inverse.get(e3, set()):
            work.setdefault(x, set()).add(e3)
            connected.setdefault(x, set()).add(e3)
            inverse.setdefault(e3, set()).add(x)
        for y in connected.get(e2, set()) - connected.get(e3, set()):
            work.setdefault(e3, set()).add(y)
            connected.setdefault(e3, set()).add(y)
            inverse.setdefault(y, set()).add(e3)
    return {key(f, p) for f, p in vals.values() if f.any()}


@pytest.mark.parametrize("grammar", ["arithmetic", "sexp", "list", "llp01", "paper_grammar",
                                     "lexer"])
def test_full_closure_reaches_what_the_connected_worklist_reaches(grammar):
    from alpacc_amd.gen.endo import DEAD, IDENTITY
    from alpacc_amd.gen.tables import generate
    path = os.path.join(GRAMMARS, grammar + ".alp")
    if not os.path.exists(path):
        pytest.skip(f"{grammar}.alp not shipped")
    lx = generate(path, mode="lexer" if grammar == "lexer" else "both").lexer
    if lx is None:
        pytest.skip(f"{grammar}.alp has no lexer")
    want = connected_worklist(lx.dfa)
    # the full closure's endomorphisms (move function and produce flags of every row but the
    # identity and the dead one), keyed the same way
    endo = lx.endo
    got = set()
    for i in range(endo.num_endos):
        if i in (IDENTITY, DEAD):
            continue
        f, p = endo.functions[i], endo.produces[i]
        got.add(np.asarray(f, dtype=np.int32).tobytes() + np.asarray(p, dtype=bool).tobytes())
    assert len(got) == endo.num_endos - 2
    assert got == want
