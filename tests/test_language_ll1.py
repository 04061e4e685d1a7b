"""Language check of the generated LLP tables of json.alp, sexp.alp and arithmetic.alp against a
textbook LL(1) predictive parser (no GPU).

The three grammars are LL(1). A predictive parser built here from its own FIRST / FOLLOW sets over
the terminal-extended, augmented grammar (Start -> |- S -|, one production t -> t per terminal,
Grammar.hs:194-214, 356-378) yields, for every token string, the verdict and the leftmost
derivation; without its first step (the augmented start production) that derivation is what
llpParse returns (LLP.hs:641-650), and a stack walk over it
gives the preorder parents that flattenTree gives (LLP.hs:716-727). Every token string up to a
bounded length over the grammar's terminals is parsed three ways:
  * the predictive parser (this file);
  * llp_parse over the LLP(q, k) pair table the generator built;
  * the C restatement of futhark/parser.fut over the device tables of the same bundle,
and verdicts, node words (terminal productions as terminal nodes, parser.fut:305-332) and parents
must agree; the same holds for the token streams of seeded documents of the benchmark generators
(thousands of tokens, nested up to 16 / 24 / 12 levels) and one-token corruptions of them. Together
these strings use every pair entry of each table that a key formed by get_key can reach.
"""
import itertools
import os

import numpy as np
import pytest

from conftest import GRAMMARS

TERMINAL = 0x80000000


def first_follow(g):
    """Nullable set, FIRST (of every symbol) and FOLLOW (of every nonterminal), by fixpoint."""
    nts = set(g.nonterminals)
    nullable = set()
    first = {a: set() for a in nts}
    changed = True
    while changed:
        changed = False
        for lhs, rhs in g.productions:
            if lhs not in nullable and all(x in nullable for x in rhs):
                nullable.add(lhs)
                changed = True
            for x in rhs:
                add = first[x] if x in nts else {x}
                if not add <= first[lhs]:
                    first[lhs] |= add
                    changed = True
                if x not in nullable:
                    break
    follow = {a: set() for a in nts}
    changed = True
    while changed:
        changed = False
        for lhs, rhs in g.productions:
            for i, x in enumerate(rhs):
                if x not in nts:
                    continue
                rest_nullable = True
                for y in rhs[i + 1:]:
                    add = first[y] if y in nts else {y}
                    if not add <= follow[x]:
                        follow[x] |= add
                        changed = True
                    if y not in nullable:
                        rest_nullable = False
                        break
                if rest_nullable and not follow[lhs] <= follow[x]:
                    follow[x] |= follow[lhs]
                    changed = True
    return nullable, first, follow


def ll1_table(g):
    nts = set(g.nonterminals)
    nullable, first, follow = first_follow(g)
    table = {}
    for p, (lhs, rhs) in enumerate(g.productions):
        sel = set()
        all_null = True
        for x in rhs:
            sel |= first[x] if x in nts else {x}
            if x not in nullable:
                all_null = False
                break
        if all_null:
            sel |= follow[lhs]
        for t in sel:
            assert (lhs, t) not in table, f"not LL(1) at {(lhs, t)}"
            table[(lhs, t)] = p
    return table


def ll1_parse(g, table, tokens):
    """Leftmost derivation of |- tokens -| from the start symbol, or None."""
    nts = set(g.nonterminals)
    inp = [g.sym.start_terminal] + list(tokens) + [g.sym.end_terminal]
    stack = [g.start]
    i = 0
    out = []
    while stack:
        x = stack.pop()
        if x in nts:
            p = table.get((x, inp[i] if i < len(inp) else None))
            if p is None:
                return None
            out.append(p)
            stack.extend(reversed(g.productions[p][1]))
        else:
            if i >= len(inp) or inp[i] != x:
                return None
            i += 1
    if i != len(inp):
        return None
    # the augmented start production is not part of the output (the README tree's root is the
    # grammar's first production, README.md:81-106)
    assert g.productions[out[0]][0] == g.start
    return out[1:]


def nodes_and_parents(g, deriv):
    """Node words and preorder parents of a leftmost derivation (the root is its own parent)."""
    nts = set(g.nonterminals)
    nodes, parents = [], []
    open_slots = []  # (node index, nonterminal children still to come)
    for i, p in enumerate(deriv):
        t = g.production_to_terminal[p]
        nodes.append(TERMINAL | t if t is not None else p)
        if open_slots:
            par, left = open_slots[-1]
            parents.append(par)
            if left == 1:
                open_slots.pop()
            else:
                open_slots[-1] = (par, left - 1)
        else:
            parents.append(i)
        kids = sum(1 for x in g.productions[p][1] if x in nts)
        if kids:
            open_slots.append((i, kids))
    return nodes, parents


@pytest.mark.parametrize("gname,maxlen", [("json", 5), ("sexp", 8), ("arithmetic", 5)])
def test_llp_tables_equal_ll1_language(gname, maxlen):
    from alpacc_amd.gen.llp import llp_parse
    from alpacc_amd.gen.tables import generate
    from oracle.c import coracle
    b = generate(os.path.join(GRAMMARS, f"{gname}.alp"))
    pt = b.parser
    g = pt.grammar
    table = ll1_table(g)
    orc = coracle.Oracle(b)
    special = {g.sym.start_terminal, g.sym.end_terminal, b.lexer.ignore, b.lexer.dead_terminal}
    terms = [t for t in range(g.sym.num_terminals) if t not in special]
    used_entries = set()
    nt = pt.num_terminals

    def check(toks):
        deriv = ll1_parse(g, table, toks)
        lp = llp_parse(pt.llp, list(toks), g.sym.start_terminal, g.sym.end_terminal)
        assert (lp is None) == (deriv is None), toks
        k = np.array(toks, dtype=np.uint32)
        pos = np.arange(len(toks), dtype=np.int64)
        r = orc.parse_tokens(k, pos, pos + 1)
        assert (r is None) == (deriv is None), toks
        if deriv is None:
            return False
        assert list(lp) == deriv, toks
        nodes, parents = nodes_and_parents(g, deriv)
        assert r[1].tolist() == nodes, toks
        assert r[0].tolist() == parents, toks
        # the pair entries this string used (keys as get_key forms them, parser.fut:117-132)
        s = [g.sym.start_terminal] + list(toks) + [g.sym.end_terminal]
        for i in range(len(s)):
            key = [pt.empty_terminal] * max(0, pt.q - i) + s[max(0, i - pt.q):i + pt.k]
            key += [pt.empty_terminal] * (pt.q + pt.k - len(key))
            di = 0
            for t in key:
                di = di * nt + t
            used_entries.add(int(pt.pair_index[di]))
        return True

    accepted = total = 0
    for L in range(0, maxlen + 1):
        for toks in itertools.product(terms, repeat=L):
            total += 1
            accepted += check(toks)
    assert accepted > 10 and accepted < total
    # long token streams: the seeded documents of the benchmark generators, and one-token
    # corruptions of them
    from alpacc_amd import datagen
    gen = {"json": datagen.json_doc, "sexp": datagen.sexp_doc, "arithmetic": datagen.arith_doc}[gname]
    rng = np.random.default_rng(7)
    for seed in range(6):
        toks = [int(x) for x in orc.lex(gen(20_000, seed=seed))[0]]
        assert check(toks)
        for _ in range(3):
            bad = list(toks)
            bad[int(rng.integers(len(bad)))] = int(rng.choice(terms))
            check(bad)
    # keys are formed at positions 0 .. T + 1 of [|-] ++ tokens ++ [-|] only (to_keys and get_key,
    # parser.fut:117-122, 235-240), so an entry whose every key holds -| among its q lookback
    # symbols is never looked up: the LLP collection has the pair, the runtime cannot reach it
    reachable = set()
    for di in np.nonzero(pt.pair_index != 0xFFFFFFFF)[0]:
        key, x = [], int(di)
        for _ in range(pt.q + pt.k):
            key.append(x % nt)
            x //= nt
        key.reverse()
        if g.sym.end_terminal not in key[:pt.q]:
            reachable.add(int(pt.pair_index[di]))
    assert used_entries == reachable, "pair entries never exercised"
