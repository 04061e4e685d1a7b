"""LL(k) and LLP(q,k) table construction over integer symbols.

Restates src/Alpacc/LL.hs and src/Alpacc/LLP.hs (Vagner & Melichar, "Parallel LL parsing", Acta
Informatica 44(1), 2007, Algorithm 8):
  * FIRST_k by least fixed point over productions (``firstMap``, LL.hs:343-353); the FIRST_k of a
    string is the k-truncated product of its symbols' sets (``alphaBetaProducts``, LL.hs:262-309);
  * FOLLOW_k on the grammar extended with ``S' -> S End^k`` (``followMap'``, LL.hs:457-500);
  * LAST_q / BEFORE_q as FIRST_q / FOLLOW_q of the grammar with every right-hand side reversed
    (LL.hs:482-491, 543-552);
  * the LL(k) table: production i is predicted for (A, y) for y in FIRST_k(rhs) (x)_k FOLLOW_k(A);
    any conflict is an error (``llTableM``, LL.hs:589-612);
  * the LLP collection: item sets explored breadth-first from D0, moving the dot one symbol left
    per step (``newLlpItemsMemo``) and closing under ``extraLlpItemsMemo``; every item whose
    dotted prefix ends in a terminal contributes the admissible pair (suffix, prefix) with its
    shortest-prefix LL parse configuration (LLP.hs:82-447); two different configurations for one
    pair mean the grammar is not LLP(q,k);
  * start pairs ([], y) for y in FIRST_k(|- S -|) with configuration ([Start], [S, -|], [])
    (``allStarts``, LLP.hs:487-508) and the bracket homomorphism: pop ``alpha`` as right brackets,
    push ``reverse omega`` as left brackets (LLP.hs:541-572).
"""
from __future__ import annotations

from collections import defaultdict
from dataclasses import dataclass


class NotLLP(ValueError):
    pass


def trunc_product(k: int, a, b) -> frozenset:
    """``truncatedProduct`` (LL.hs:204-210)."""
    out = set()
    for x in a:
        if len(x) >= k:
            out.add(x[:k])
            continue
        for y in b:
            out.add((x + y)[:k])
    return frozenset(out)


class FirstFn:
    """FIRST_k of symbol strings under a fixed grammar, memoised."""

    def __init__(self, k: int, productions, nonterminals, is_terminal):
        self.k = k
        self.is_terminal = is_terminal
        self.memo: dict = {}
        fm = {a: frozenset() for a in nonterminals}
        while True:
            new = dict(fm)
            for lhs, rhs in productions:
                s = self._string(rhs, fm)
                if not s <= new[lhs]:
                    new[lhs] = new[lhs] | s
            if new == fm:
                break
            fm = new
        self.first_map = fm

    def _string(self, w, fm) -> frozenset:
        cur = frozenset({()})
        for s in w:
            if self.is_terminal(s):
                sset = frozenset({(s,)[: self.k]})
            else:
                sset = fm[s]
            cur = trunc_product(self.k, cur, sset)
            if not cur:
                break
        return cur

    def __call__(self, w) -> frozenset:
        w = tuple(w)
        r = self.memo.get(w)
        if r is None:
            r = self._string(w, self.first_map)
            self.memo[w] = r
        return r


END = -1          # the End terminal of the follow-set extension (LL.hs:371-375)


def follow_sets(k: int, start, productions, nonterminals, is_terminal) -> dict:
    """FOLLOW_k (``followMap``, LL.hs:457-500): extended grammar S' -> S End^k, End dropped after."""
    ext_start = object()
    ext_prods = [(ext_start, (start,) + (END,) * k)] + list(productions)
    ext_nts = [ext_start] + list(nonterminals)

    def is_t(s):
        return s == END or (s is not ext_start and is_terminal(s))

    first = FirstFn(k, ext_prods, ext_nts, is_t)
    fol = {a: frozenset() for a in ext_nts}
    fol[start] = frozenset({(END,) * k})
    right = []
    for aj, syms in ext_prods:
        for idx, s in enumerate(syms):
            if not is_t(s):
                right.append((aj, s, tuple(syms[idx + 1:])))
    while True:
        new = dict(fol)
        for aj, ai, w in right:
            sub = trunc_product(k, first(w), fol[aj])
            if not sub <= new[ai]:
                new[ai] = new[ai] | sub
        if new == fol:
            break
        fol = new
    return {a: frozenset(tuple(t for t in x if t != END) for x in fol[a]) for a in nonterminals}


@dataclass
class LLPTable:
    q: int
    k: int
    # (suffix tuple, prefix tuple) -> (alpha, omega, pi), before padding
    table: dict
    # (suffix tuple, prefix tuple) -> (brackets list of (is_left, symbol), productions list)
    homomorphisms: dict


class LLPBuilder:
    def __init__(self, g, q: int, k: int):
        if q < 0:
            raise ValueError("Error: Lookback must be non-negative.")
        if k < 1:
            raise ValueError("Error: Lookahead must be positive.")
        self.g = g
        self.q = q
        self.k = k
        nT = g.sym.num_terminals
        self.is_t = lambda s: s < nT
        self.prods = g.productions
        self.by_lhs = defaultdict(list)
        for i, (lhs, rhs) in enumerate(self.prods):
            self.by_lhs[lhs].append(rhs)
        nts = g.nonterminals
        self.first = FirstFn(k, self.prods, nts, self.is_t)
        self.follow = follow_sets(k, g.start, self.prods, nts, self.is_t)
        rprods = [(a, tuple(reversed(r))) for a, r in self.prods]
        self._rfirst = FirstFn(q, rprods, nts, self.is_t)
        rfollow = follow_sets(q, g.start, rprods, nts, self.is_t)
        self.before = {a: frozenset(tuple(reversed(x)) for x in s) for a, s in rfollow.items()}
        self._last_memo: dict = {}
        self.ll_table = self._ll_table()
        self._cfg_memo: dict = {}

    def last(self, w) -> frozenset:
        w = tuple(w)
        r = self._last_memo.get(w)
        if r is None:
            r = frozenset(tuple(reversed(x)) for x in self._rfirst(tuple(reversed(w))))
            self._last_memo[w] = r
        return r

    def _ll_table(self) -> dict:
        table = defaultdict(set)
        for i, (nt, a) in enumerate(self.prods):
            for y in trunc_product(self.k, self.first(a), self.follow[nt]):
                table[(nt, y)].add(i)
        conflicts = {key: v for key, v in table.items() if len(v) > 1}
        if conflicts:
            raise NotLLP(f"LL({self.k}) Table Conflicts: {len(conflicts)}")
        return {key: min(v) for key, v in table.items()}

    def parse_config(self, y, alpha):
        """``parseLlpConfig`` (LLP.hs:454-478)."""
        inp = tuple(y)
        stack = list(alpha)
        parsed = []
        k = self.k
        while True:
            if inp and stack and self.is_t(stack[0]):
                if inp[0] == stack[0]:
                    return (tuple(alpha), tuple(stack[1:]), tuple(parsed))
                return None
            if stack and not self.is_t(stack[0]):
                idx = self.ll_table.get((stack[0], inp[:k]))
                if idx is None:
                    return None
                stack = list(self.prods[idx][1]) + stack[1:]
                parsed.append(idx)
                continue
            if not inp and not stack:
                return (tuple(alpha), (), tuple(parsed))
            return None

    def shortest_prefix(self, v, x_delta):
        """``solveShortestsPrefix`` (LLP.hs:98-112)."""
        key = (v, x_delta)
        r = self._cfg_memo.get(key, 0)
        if r != 0:
            return r
        r = None
        for L in range(1, len(x_delta) + 1):
            r = self.parse_config(v, x_delta[:L])
            if r is not None:
                break
        self._cfg_memo[key] = r
        return r

    # Items: (dp, suffix, prefix, config) with dp = (nt, alpha, beta).
    def init_d(self) -> frozenset:
        starts = self.by_lhs[self.g.start]
        if len(starts) != 1:
            return frozenset()
        p = starts[0]
        return frozenset(((self.g.start, p, ()), l, (), ((), (), ()))
                         for l in self.last(p))

    def new_items(self, item) -> set:
        (y, alpha_x, beta), _u, v_i, delta = item
        if not alpha_x:
            return set()
        alpha = alpha_x[:-1]
        uj = set()
        for b in self.before[y]:
            uj |= self.last(b + alpha)
        x = alpha_x[-1:]
        vj = self.first(x + v_i)
        x_delta = x + delta[0]
        dp = (y, alpha, x + beta)
        out = set()
        for u in uj:
            for v in vj:
                cfg = self.shortest_prefix(v, x_delta)
                if cfg is None:
                    raise NotLLP("no LL configuration for an LLP item")
                out.add((dp, u, v, cfg))
        return out

    def extra_items(self, item) -> set:
        (_, alpha_y, _), _u, v, gamma = item
        if not alpha_y or self.is_t(alpha_y[-1]):
            return set()
        y = alpha_y[-1]
        out = set()
        for delta in self.by_lhs[y]:
            for b in self.before[y]:
                for u in self.last(b + delta):
                    out.add(((y, delta, ()), u, v, gamma))
        return out

    def close(self, items: set) -> frozenset:
        items = set(items)
        frontier = list(items)
        while frontier:
            nxt = []
            for it in frontier:
                for e in self.extra_items(it):
                    if e not in items:
                        items.add(e)
                        nxt.append(e)
            frontier = nxt
        return frozenset(items)

    def successors(self, items: frozenset) -> list:
        groups = defaultdict(list)
        for it in items:
            alpha_x = it[0][1]
            if alpha_x:
                groups[alpha_x[-1]].append(it)
        out = []
        for sym in sorted(groups):
            new = set()
            for it in groups[sym]:
                new |= self.new_items(it)
            out.append(self.close(new))
        return out

    def collection_table(self) -> dict:
        """``llpCollectionMemo`` (LLP.hs:377-447): the admissible-pair table."""
        table: dict = {}
        seen = set()
        queue = [self.init_d()]
        head = 0
        while head < len(queue):
            top = queue[head]
            head += 1
            if top in seen:
                continue
            seen.add(top)
            for (dp, x, y, gamma) in top:
                alpha_z = dp[1]
                if alpha_z and self.is_t(alpha_z[-1]):
                    key = (x, y)
                    old = table.get(key)
                    if old is not None and old != gamma:
                        raise NotLLP(f"The grammar is not LLP({self.q}, {self.k}): pair {key} "
                                     f"has two configurations")
                    table[key] = gamma
            queue.extend(self.successors(top))
        return table

    def starts(self) -> dict:
        """``allStarts`` (LLP.hs:487-508)."""
        s = self.by_lhs[self.g.start][0]
        tail = tuple(s[1:])
        return {((), z): ((self.g.start,), tail, ()) for z in self.first(s)}

    def build(self) -> LLPTable:
        coll = self.collection_table()
        starts = self.starts()
        table = dict(starts)
        table.update(coll)     # Map.union keeps the collection's entry on a shared key
        hom = {}
        for key, (alpha, omega, pi) in table.items():
            if key in coll:
                br = [(False, a) for a in alpha] + [(True, w) for w in reversed(omega)]
            else:
                br = [(True, w) for w in reversed(omega)]
            hom[key] = (br, list(pi))
        return LLPTable(q=self.q, k=self.k, table=table, homomorphisms=hom)


def llp_parse(tb: LLPTable, string, start_terminal, end_terminal):
    """Sequential LLP parse by gluing pair configurations (``llpParse``, LLP.hs:600-650).

    Returns the production sequence (a left parse) or None.
    """
    q, k = tb.q, tb.k
    s = [start_terminal] + list(string) + [end_terminal]
    pairs = []
    for i in range(len(s)):
        back = tuple(s[max(0, i - q):i])
        fwd = tuple(s[i:i + k])
        v = tb.table.get((back, fwd))
        if v is None:
            return None
        pairs.append(v)
    if not pairs:
        return None
    alpha_l, omega_l, pi_l = list(pairs[0][0]), list(pairs[0][1]), list(pairs[0][2])
    for alpha_r, omega_r, pi_r in pairs[1:]:
        alpha_r, omega_r = list(alpha_r), list(omega_r)
        pi = pi_l + list(pi_r)
        if omega_l == alpha_r:
            omega_l = omega_r
        elif alpha_r[:len(omega_l)] == omega_l:
            alpha_l = alpha_l + alpha_r[len(omega_l):]
            omega_l = omega_r
        elif omega_l[:len(alpha_r)] == alpha_r:
            omega_l = omega_r + omega_l[len(alpha_r):]
        else:
            return None
        pi_l = pi
    return pi_l
