"""CPU tests of the restated table generator against the reference's golden vectors and
closed-form properties (no GPU)."""
import itertools
import json
import os
import random

import numpy as np
import pytest

from conftest import GRAMMARS, ROOT

GOLDEN = json.load(open(os.path.join(ROOT, "tests", "golden", "haskell_vectors.json")))


# --------------------------------------------------------------------------------------------
# regular expressions (haskell-tests/RegularExpression.hs:37-138)

@pytest.mark.parametrize("case", GOLDEN["regex"], ids=lambda c: c["regex"] or "<empty>")
def test_regex_golden(case):
    from alpacc_amd.gen.automata import regex_dfa
    from alpacc_amd.gen.regex import regex_from_text
    d = regex_dfa(regex_from_text(case["regex"]))
    for s in case["valid"]:
        assert d.matches(s.encode()), s
    for s in case["invalid"]:
        assert not d.matches(s.encode()), s


def test_regex_dialect_closed_forms():
    from alpacc_amd.gen.automata import regex_dfa
    from alpacc_amd.gen.regex import regex_from_text
    # ranges of alphanumerics and bytes; escapes; + and ?; whitespace between tokens is skipped
    d = regex_dfa(regex_from_text(r"[a-c\x30-\x32]+ x?"))
    for b in range(256):
        assert d.matches(bytes([b])) == (b in b"abc012")
    assert d.matches(b"a0x") and d.matches(b"cc") and not d.matches(b"x")
    d = regex_dfa(regex_from_text(r"\s|\n|\t|\r|\\|\/|\[|\]|\(|\)|\*|\+|\||\?"))
    assert {b for b in range(256) if d.matches(bytes([b]))} == set(b" \n\t\r\\/[]()*+|?")
    d = regex_dfa(regex_from_text(r"é"))      # UTF-8 encoded code point
    assert d.matches("é".encode()) and not d.matches(b"\xe9")


# --------------------------------------------------------------------------------------------
# numbering (src/Alpacc/Encode.hs:122-162, Grammar.hs:194-214, 356-378, Analyzer.hs:84-147)

def test_numbering_json():
    from alpacc_amd.gen.grammar import build_grammar, load_cfg, terminal_names_both, lexer_spec, lexer_only_ids
    g = build_grammar(load_cfg(os.path.join(GRAMMARS, "json.alp")))
    names = terminal_names_both(g)
    assert names[:2] == ["string", "number"] and names[11] == "ignore" and names[12] == "empty_12"
    assert (g.sym.start_terminal, g.sym.end_terminal, g.sym.empty) == (13, 14, 12)
    assert g.production_names[:8] == ["Object_production", "Array_production", "String", "Number",
                                      "Null", "True", "False", "Object_head"]
    assert g.production_names[30] == "start_30"
    assert g.arities[:18] == [1, 1, 1, 1, 1, 1, 1, 3, 0, 2, 0, 3, 3, 3, 0, 2, 0, 3]
    assert g.production_to_terminal[18:30] == [0, 1, 11, 2, 3, 4, 5, 6, 7, 8, 9, 10]
    ids, dead, ign, lnames = lexer_only_ids(lexer_spec(load_cfg(os.path.join(GRAMMARS, "json.alp"))))
    assert lnames[:2] == ["number", "string"] and dead == 12 and ign == 11


def test_numbering_arithmetic_and_sexp():
    from alpacc_amd.gen.grammar import build_grammar, load_cfg, terminal_names_both
    g = build_grammar(load_cfg(os.path.join(GRAMMARS, "arithmetic.alp")))
    assert terminal_names_both(g)[:10] == ["num", "var", "literal_2", "literal_3", "literal_4",
                                           "literal_5", "literal_6", "literal_7", "ignore", "empty_9"]
    assert g.production_names[:11] == ["E0_0", "Add", "Sub", "Empty", "E1_4", "Mul", "Div", "Empty",
                                       "Parentheses", "Num", "Var"]
    assert g.production_to_terminal[11:20] == [0, 1, 8, 2, 3, 4, 5, 6, 7]
    s = build_grammar(load_cfg(os.path.join(GRAMMARS, "sexp.alp")))
    assert terminal_names_both(s) == ["atom", "literal_1", "literal_2", "ignore", "empty_4"]
    assert s.production_names == ["E_0", "E_1", "E0_2", "E0_3", "ignore_4", "atom_5", "Literal_6",
                                  "Literal_7", "start_8"]
    assert s.arities == [1, 3, 0, 2, 0, 0, 0, 0, 1]


# --------------------------------------------------------------------------------------------
# LL / LLP (haskell-tests/LL.hs:75-131, LLP.hs:83-97, 488-542)

class _Sym:
    pass


def plain_grammar(terms, nts, prods, start, augment=True):
    """Integer grammar for the LL/LLP builders; augmented with Start -> |- S -| when asked."""
    names = list(terms) + (["|-", "-|"] if augment else [])
    nT = len(names)
    ntid = {n: nT + i for i, n in enumerate(nts)}
    names += list(nts)
    tid = {t: i for i, t in enumerate(terms)}

    def sid(x):
        return tid[x] if x in tid else ntid[x]

    P = [(ntid[a], tuple(sid(s) for s in rhs)) for a, rhs in prods]
    start_id = ntid[start]
    if augment:
        names.append("<start>")
        start_id = len(names) - 1
        P.append((start_id, (nT - 2, ntid[start], nT - 1)))
    sym = _Sym()
    sym.num_terminals = nT
    sym.start_terminal = nT - 2 if augment else None
    sym.end_terminal = nT - 1 if augment else None
    g = _Sym()
    g.sym = sym
    g.productions = P
    g.nonterminals = list(range(nT, len(names)))
    g.start = start_id
    g.names = names
    return g, tid


PAPER = (["[", "]", "+", "a"], ["E", "E'", "T"],
         [("E", ["T", "E'"]), ("E'", ["+", "T", "E'"]), ("E'", []), ("T", ["a"]),
          ("T", ["[", "E", "]"])], "E")


@pytest.mark.parametrize("q", [1, 2, 3])
@pytest.mark.parametrize("k", [1, 2, 3])
def test_llp_parse_golden(q, k):
    from alpacc_amd.gen.llp import LLPBuilder, llp_parse
    g, tid = plain_grammar(*PAPER)
    tb = LLPBuilder(g, q, k).build()
    v = GOLDEN["llp_parse"]
    got = llp_parse(tb, [tid[c] for c in v["input"]], g.sym.start_terminal, g.sym.end_terminal)
    assert got == v["expected"]


def _earley(prods, start, terms_seq):
    """Independent CFG recogniser (Earley) used to decide derivability."""
    chart = [set() for _ in range(len(terms_seq) + 1)]
    chart[0].add(("S'", (start,), 0, 0))
    rules = {}
    for a, rhs in prods:
        rules.setdefault(a, []).append(tuple(rhs))
    for i in range(len(terms_seq) + 1):
        agenda = list(chart[i])
        while agenda:
            lhs, rhs, dot, org = agenda.pop()
            if dot < len(rhs):
                sym = rhs[dot]
                if sym in rules:
                    for r in rules[sym]:
                        it = (sym, r, 0, i)
                        if it not in chart[i]:
                            chart[i].add(it)
                            agenda.append(it)
                    # nullable completion
                    for it2 in list(chart[i]):
                        if it2[0] == sym and it2[2] == len(it2[1]) and it2[3] == i:
                            adv = (lhs, rhs, dot + 1, org)
                            if adv not in chart[i]:
                                chart[i].add(adv)
                                agenda.append(adv)
                elif i < len(terms_seq) and terms_seq[i] == sym:
                    chart[i + 1].add((lhs, rhs, dot + 1, org))
            else:
                for it in list(chart[org]):
                    if it[2] < len(it[1]) and it[1][it[2]] == lhs:
                        adv = (it[0], it[1], it[2] + 1, it[3])
                        if adv not in chart[i]:
                            chart[i].add(adv)
                            agenda.append(adv)
    return any(it[0] == "S'" and it[2] == 1 for it in chart[len(terms_seq)])


@pytest.mark.parametrize("q,k", [(1, 1), (2, 2), (1, 3)])
def test_llp_accepts_exactly_the_language(q, k):
    """Every string of length <= 6 over the paper grammar's terminals: LLP acceptance equals
    derivability (stronger than haskell-tests/LLP.hs:516-542, which samples both sides)."""
    from alpacc_amd.gen.llp import LLPBuilder, llp_parse
    terms, nts, prods, start = PAPER
    g, tid = plain_grammar(*PAPER)
    tb = LLPBuilder(g, q, k).build()
    for L in range(0, 7):
        for s in itertools.product(terms, repeat=L):
            want = _earley(prods, start, list(s))
            got = llp_parse(tb, [tid[c] for c in s], g.sym.start_terminal, g.sym.end_terminal)
            assert (got is not None) == want, s


def test_psls_golden():
    """Admissible pairs -> pushed symbols at q = k = 1 (haskell-tests/LLP.hs:83-97)."""
    from alpacc_amd.gen.llp import LLPBuilder
    g, tid = plain_grammar(*PAPER)
    b = LLPBuilder(g, 1, 1)
    coll = b.collection_table()
    name = {v: k for k, v in tid.items()}
    name[g.sym.start_terminal] = "|-"
    name[g.sym.end_terminal] = "-|"
    for i, n in zip(g.nonterminals, PAPER[1] + ["<start>"]):
        name[i] = n

    def enc(sym_list):
        return [["nt" if s >= g.sym.num_terminals else "t", name[s]] for s in sym_list]

    got = []
    for (x, y), (alpha, omega, pi) in coll.items():
        if y:     # the D0 pair after -| has an empty prefix; it is not part of PSLS
            got.append({"suffix": enc(x), "prefix": enc(y), "alpha": enc(alpha)})
    key = lambda d: json.dumps(d, sort_keys=True)
    assert sorted(got, key=key) == sorted(GOLDEN["psls_paper_q1_k1"], key=key)


LL_T = (["a", "b", "c", "$"], ["T'", "R", "T"],
        [("T'", ["T", "$"]), ("T", ["R"]), ("T", ["a", "T", "c"]), ("R", []), ("R", ["b", "R"])], "T'")
LL_FOLLOW_T = (["a", "b", "c", "$"], ["T'", "R", "T"],
               [("T'", ["T", "$"]), ("T", ["R"]), ("T", ["a", "T", "c"]), ("R", []),
                ("R", ["R", "b", "R"])], "T'")
BOOK = (["$", "a", "b"], ["N'", "N", "A", "B", "C"],
        [("N'", ["N", "$"]), ("N", ["A", "B"]), ("N", ["B", "A"]), ("A", ["a"]), ("A", ["C", "A", "C"]),
         ("B", ["b"]), ("B", ["C", "B", "C"]), ("C", ["a"]), ("C", ["b"])], "N'")


def _sets(fs, names):
    return sorted(sorted([names[t] for t in x] for x in s) for s in fs)


def test_first_follow_golden():
    from alpacc_amd.gen.llp import FirstFn, follow_sets
    # FIRST(1) of the right-hand sides of T -> R | a T c, R -> eps | b R (haskell-tests/LL.hs:75-84)
    terms, nts, prods, start = (["a", "b", "c"], ["R", "T"],
                                [("T", ["R"]), ("T", ["a", "T", "c"]), ("R", []), ("R", ["b", "R"])], "T")
    g, tid = plain_grammar(terms, nts, prods, start, augment=False)
    inv = {v: k for k, v in tid.items()}
    f = FirstFn(1, g.productions, g.nonterminals, lambda s: s < g.sym.num_terminals)
    got = [sorted([inv[t] for t in x] for x in f(rhs)) for _, rhs in g.productions]
    assert got == GOLDEN["ll"]["first_small"]
    for spec, key in ((LL_FOLLOW_T, "follow_small"), (BOOK, "follow_large")):
        g, tid = plain_grammar(*spec, augment=False)
        inv = {v: k for k, v in tid.items()}
        fol = follow_sets(1, g.start, g.productions, g.nonterminals, lambda s: s < g.sym.num_terminals)
        got = [sorted([inv[t] for t in x] for x in fol[nt]) for nt in g.nonterminals]
        assert got == GOLDEN["ll"][key]


@pytest.mark.parametrize("k", range(1, 11))
def test_ll_parse_golden(k):
    """LL(k) predictive parse of aabbbcc$ (haskell-tests/LL.hs:124-131)."""
    from alpacc_amd.gen.llp import LLPBuilder
    g, tid = plain_grammar(*LL_T, augment=False)
    b = LLPBuilder.__new__(LLPBuilder)
    b.g, b.q, b.k = g, 0, k
    from alpacc_amd.gen.llp import FirstFn, follow_sets
    b.is_t = lambda s: s < g.sym.num_terminals
    b.prods = g.productions
    b.first = FirstFn(k, g.productions, g.nonterminals, b.is_t)
    b.follow = follow_sets(k, g.start, g.productions, g.nonterminals, b.is_t)
    table = b._ll_table()
    inp = [tid[c] for c in GOLDEN["ll"]["ll_parse"]["input"]]
    stack = [g.start]
    out = []
    pos = 0
    while stack:
        top = stack.pop(0)
        if b.is_t(top):
            assert inp[pos] == top
            pos += 1
            continue
        idx = table[(top, tuple(inp[pos:pos + k]))]
        out.append(idx)
        stack = list(g.productions[idx][1]) + stack
    assert out == GOLDEN["ll"]["ll_parse"]["expected"] and pos == len(inp)


# --------------------------------------------------------------------------------------------
# endomorphism tables (closed forms)

@pytest.mark.parametrize("name", ["json", "arithmetic", "sexp", "lexer"])
def test_endomorphism_laws(name):
    from alpacc_amd.gen.tables import generate
    b = generate(os.path.join(GRAMMARS, f"{name}.alp"), mode="lexer" if name == "lexer" else "both")
    lt = b.lexer
    e = lt.endo.num_endos
    comp = lt.endo_comp.astype(np.int64)          # [a, b] = a then b
    F = lt.endo.functions                          # [e, m+1]
    rng = random.Random(1)
    for _ in range(3000):
        a, bb, c = (rng.randrange(e) for _ in range(3))
        assert comp[comp[a, bb], c] == comp[a, comp[bb, c]]             # associativity
    assert np.array_equal(comp[0, :], np.arange(e)) and np.array_equal(comp[:, 0], np.arange(e))
    assert np.all(comp[1, :] == 1) and np.all(comp[:, 1] == 1)          # dead absorbs
    for a in range(2, e):                          # table agrees with function composition
        for bb in rng.sample(range(2, e), min(e - 2, 8)):
            want = F[bb][F[a]]
            got = F[comp[a, bb]]
            assert np.array_equal(want, got) or (not want.any() and comp[a, bb] == 1)
    # endo_step is the (endo x class) restriction of the composition table
    for c in range(lt.num_classes):
        byte = int(np.nonzero(lt.byte_class == c)[0][0])
        assert np.array_equal(lt.endo_step[:, c], comp[:, lt.endo.byte_endo[byte]])


@pytest.mark.parametrize("name", ["json", "arithmetic", "sexp"])
def test_scan_of_byte_endomorphisms_runs_the_dfa(name):
    """The composed endomorphism of a prefix applied to the initial state is the DFA run."""
    from alpacc_amd import datagen
    from alpacc_amd.gen.tables import generate
    b = generate(os.path.join(GRAMMARS, f"{name}.alp"))
    lt = b.lexer
    gen = {"json": datagen.json_doc, "arithmetic": datagen.arith_doc, "sexp": datagen.sexp_doc}[name]
    doc = gen(3000, seed=5)
    comp = lt.endo_comp
    s = lt.init_state
    acc = 0
    for x in doc.tolist():
        acc = comp[acc, lt.endo.byte_endo[x]]
        s = lt.dfa_step[s * lt.num_classes + lt.byte_class[x]] & 0xFF
        assert lt.endo_apply[acc, lt.init_state] == s
