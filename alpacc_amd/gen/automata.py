"""Byte-level automata for the lexer: regex -> NFA -> DFA, and the union lexer DFA.

Semantics follow src/Alpacc/Lexer/{NFA,DFA,FSA}.hs:
  * each token's regex becomes an automaton over bytes (NFA.hs:225-255);
  * the lexer automaton is the subset construction of the union of the token automata, started
    from a fresh state with an epsilon move to every token's start (DFA.hs:333-406), restricted to
    states from which an accepting state is reachable (``removeUselessStates``, DFA.hs:180-194);
  * an accepting DFA state is labelled with the highest-priority (lowest order) token whose
    accepting state it contains (DFA.hs:371-384);
  * "producing" transitions are added from every accepting state q on every byte t that q has no
    move on but the initial state has; they jump to the initial state's t-successor
    (``tokenProducingTransitions``, DFA.hs:272-310).

The labelled DFA is additionally minimised by partition refinement over (token label, moves).
Lexing only observes, for each reachable state, its token label and which continuations exist, so
merging states that agree on both for every suffix leaves every lexer output unchanged; it only
shrinks the internal state numbering (which the reference also treats as internal).
"""
from __future__ import annotations

from dataclasses import dataclass, field


@dataclass
class NFA:
    n: int = 0
    eps: dict = field(default_factory=dict)      # state -> set(state)
    trans: dict = field(default_factory=dict)    # state -> {byte: set(state)}

    def new(self) -> int:
        s = self.n
        self.n += 1
        return s

    def add_eps(self, a: int, b: int) -> None:
        self.eps.setdefault(a, set()).add(b)

    def add(self, a: int, byte: int, b: int) -> None:
        self.trans.setdefault(a, {}).setdefault(byte, set()).add(b)

    def add_bytes(self, a: int, bs: bytes, b: int) -> None:
        # newTransitions (NFA.hs:215-223): a chain of fresh states for multi-byte literals.
        if len(bs) == 0:
            self.add_eps(a, b)
            return
        cur = a
        for j, byte in enumerate(bs):
            nxt = b if j == len(bs) - 1 else self.new()
            self.add(cur, byte, nxt)
            cur = nxt


def build_nfa(nfa: NFA, r, s: int, t: int) -> None:
    """Thompson-style construction between s and t (auxRegExToNFA, NFA.hs:225-245)."""
    tag = r[0]
    if tag == "eps":
        nfa.add_eps(s, t)
    elif tag == "lit":
        nfa.add_bytes(s, r[1], t)
    elif tag == "range":
        for bs in r[1]:
            nfa.add_bytes(s, bs, t)
    elif tag == "cat":
        m = nfa.new()
        build_nfa(nfa, r[1], s, m)
        build_nfa(nfa, r[2], m, t)
    elif tag == "star":
        m = nfa.new()
        nfa.add_eps(s, m)
        nfa.add_eps(m, t)
        build_nfa(nfa, r[1], m, m)
    elif tag == "alt":
        build_nfa(nfa, r[1], s, t)
        build_nfa(nfa, r[2], s, t)
    else:
        raise ValueError(tag)


def _closure(nfa: NFA, states) -> frozenset:
    stack = list(states)
    seen = set(states)
    while stack:
        q = stack.pop()
        for r in nfa.eps.get(q, ()):
            if r not in seen:
                seen.add(r)
                stack.append(r)
    return frozenset(seen)


@dataclass
class DFA:
    """A deterministic automaton over bytes. States are 0..n-1; ``delta[q]`` maps byte -> state."""
    n: int
    initial: int
    delta: list            # list[dict[int, int]]
    accepting: set
    label: list            # token (or None) per state

    def run(self, data: bytes):
        q = self.initial
        for b in data:
            q = self.delta[q].get(b)
            if q is None:
                return None
        return q

    def matches(self, data: bytes) -> bool:
        q = self.run(data)
        return q is not None and q in self.accepting


def subset_construction(nfa: NFA, start: int, accept_label: dict, priority: dict) -> DFA:
    """Subset construction (DFA.hs:123-151) with useless states removed (DFA.hs:180-194).

    ``accept_label`` maps accepting NFA states to their token; ``priority`` maps tokens to order.
    """
    init = _closure(nfa, [start])
    index = {init: 0}
    sets = [init]
    delta: list[dict] = [{}]
    work = [init]
    while work:
        cur = work.pop()
        ci = index[cur]
        moves: dict[int, set] = {}
        for q in cur:
            for byte, dst in nfa.trans.get(q, {}).items():
                moves.setdefault(byte, set()).update(dst)
        for byte, dst in moves.items():
            nxt = _closure(nfa, dst)
            if nxt not in index:
                index[nxt] = len(sets)
                sets.append(nxt)
                delta.append({})
                work.append(nxt)
            delta[ci][byte] = index[nxt]
    n = len(sets)
    labels = []
    accepting = set()
    for i, st in enumerate(sets):
        toks = [accept_label[q] for q in st if q in accept_label]
        if toks:
            accepting.add(i)
            labels.append(min(toks, key=lambda t: priority[t]))
        else:
            labels.append(None)
    # removeUselessStates: keep states that can reach an accepting state.
    rev: list[set] = [set() for _ in range(n)]
    for q in range(n):
        for r in delta[q].values():
            rev[r].add(q)
    useful = set(accepting)
    stack = list(accepting)
    while stack:
        q = stack.pop()
        for p in rev[q]:
            if p not in useful:
                useful.add(p)
                stack.append(p)
    if 0 not in useful:
        raise ValueError("the lexer accepts no input")
    keep = sorted(useful)
    remap = {q: i for i, q in enumerate(keep)}
    new_delta = []
    for q in keep:
        new_delta.append({b: remap[r] for b, r in delta[q].items() if r in useful})
    return DFA(n=len(keep), initial=remap[0], delta=new_delta,
               accepting={remap[q] for q in accepting}, label=[labels[q] for q in keep])


def minimize(d: DFA) -> DFA:
    """Moore partition refinement over (label, moves); missing moves form their own class."""
    block = {}
    cls = []
    for q in range(d.n):
        key = (q in d.accepting, d.label[q])
        cls.append(block.setdefault(key, len(block)))
    alphabet = sorted({b for m in d.delta for b in m})
    while True:
        sig = {}
        new_cls = []
        for q in range(d.n):
            key = (cls[q],) + tuple(cls[d.delta[q][b]] if b in d.delta[q] else -1 for b in alphabet)
            new_cls.append(sig.setdefault(key, len(sig)))
        if len(sig) == len(set(cls)):
            cls = new_cls
            break
        cls = new_cls
    # Renumber so that the initial state's class comes first, others in first-seen order.
    order = {}
    order[cls[d.initial]] = 0
    for q in range(d.n):
        if cls[q] not in order:
            order[cls[q]] = len(order)
    m = len(order)
    delta = [dict() for _ in range(m)]
    label = [None] * m
    accepting = set()
    for q in range(d.n):
        c = order[cls[q]]
        label[c] = d.label[q]
        if q in d.accepting:
            accepting.add(c)
        for b, r in d.delta[q].items():
            delta[c][b] = order[cls[r]]
    return DFA(n=m, initial=0, delta=delta, accepting=accepting, label=label)


def regex_dfa(r) -> DFA:
    """A DFA of a single regex (fromRegExToDFA, DFA.hs:153-154); label "match" on acceptance."""
    nfa = NFA()
    s = nfa.new()
    t = nfa.new()
    build_nfa(nfa, r, s, t)
    return subset_construction(nfa, s, {t: "match"}, {"match": 0})


@dataclass
class LexerDFA:
    """The lexer automaton: a DFA plus the set of producing (state, byte) moves."""
    dfa: DFA
    producing: set          # set of (state, byte)


def lexer_dfa(token_regexes: list, priority: dict) -> LexerDFA:
    """Union lexer DFA over all tokens (lexerDFA, DFA.hs:333-406) with producing moves added.

    ``token_regexes`` is a list of (token, regex AST); ``priority`` maps token -> order.
    """
    nfa = NFA()
    start = nfa.new()
    accept_label = {}
    for tok, r in token_regexes:
        s = nfa.new()
        t = nfa.new()
        build_nfa(nfa, r, s, t)
        nfa.add_eps(start, s)
        accept_label[t] = tok
    d = minimize(subset_construction(nfa, start, accept_label, priority))
    # tokenProducingTransitions (DFA.hs:272-297)
    producing = set()
    init_moves = d.delta[d.initial]
    for q in sorted(d.accepting):
        for b, r in init_moves.items():
            if b not in d.delta[q]:
                d.delta[q][b] = r
                producing.add((q, b))
    return LexerDFA(dfa=d, producing=producing)


def tokenize(ld: LexerDFA, data: bytes, ignore):
    """Sequential lexer (``tokenize``, DFA.hs:415-447). Returns [(token, (i, j))] or None."""
    d = ld.dfa
    n = len(data)
    if n == 0:
        return None if d.initial not in d.accepting else []
    out = []
    s = d.initial
    i = 0
    for j in range(n):
        s2 = d.delta[s].get(data[j])
        if s2 is None:
            return None
        if j == n - 1:
            if s2 not in d.accepting:
                return None
            k = d.label[s2]
            if k != ignore:
                out.append((k, (i, j + 1)))
            return out
        if (s2, data[j + 1]) in ld.producing:
            k = d.label[s2]
            if k is None:
                return None
            if k != ignore:
                out.append((k, (i, j + 1)))
            i = j + 1
        s = s2
    return out
