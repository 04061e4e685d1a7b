"""Parsing grammar, symbol numbering and production metadata derived from a ``.alp`` file.

Restates the output-visible numbering rules of the reference:
  * grammar terminals = nub(declared terminal rules in file order ++ terminals used in the
    nonterminal rules, in ``Ord T`` order) (CFG.hs:115-139);
  * ``extendByTerminals``: every grammar symbol becomes a nonterminal and each terminal t gets a
    terminal production ``t -> t``; nonterminal order = terminal-nonterminals, then the declared
    nonterminals (Grammar.hs:356-378); an ``Unused`` terminal is appended (Grammar.hs:338-354);
    the grammar is augmented with ``Start -> |- S -|`` (Grammar.hs:194-214);
  * symbol IDs: non-special terminals in order, then ignore, Unused, |-, -|, then the
    nonterminals in order (``encodeSymbols``, Encode.hs:122-146);
  * lexer-only terminal IDs number the *sorted* terminal list (``encodeTerminals``,
    Encode.hs:148-162; Analyzer.hs:106-131);
  * lexer priority: implicit literal terminals (``Ord T`` order) first, then the declared terminal
    rules in file order (``cfgToDFALexerSpec``, CFG.hs:146-170);
  * production names (Analyzer.hs:140-147, Encode.hs:344-351), terminal names (Analyzer.hs:97-104),
    arities (Analyzer.hs:133-138) and production-to-terminal (Analyzer.hs:84-95).
"""
from __future__ import annotations

from dataclasses import dataclass

from .alp import CFG, NT, T, parse_alp
from .regex import literal_regex, produces_epsilon

IGNORE = T(0, "ignore")


@dataclass
class Symbols:
    """Integer symbol space of the augmented parsing grammar."""
    names: list                # printable name of each symbol id
    num_terminals: int         # ids [0, num_terminals) are terminals
    ignore: int | None         # id of the ignore terminal (if the grammar declares one)
    empty: int                 # id of Unused (the empty / dead terminal)
    start_terminal: int        # |-
    end_terminal: int          # -|
    start_nt: int              # Start
    t_id: dict                 # T -> terminal id
    nt_of_t: dict              # T -> id of its terminal-nonterminal
    nt_id: dict                # NT -> nonterminal id

    def is_terminal(self, s: int) -> bool:
        return s < self.num_terminals


@dataclass
class Grammar:
    """Augmented, terminal-extended parsing grammar over integer symbols."""
    sym: Symbols
    start: int
    terminals: list
    nonterminals: list
    productions: list          # list of (lhs, rhs tuple)
    production_names: list
    arities: list
    production_to_terminal: list   # terminal id or None per production
    q: int
    k: int

    def reversed(self) -> "Grammar":
        g = Grammar(**self.__dict__)
        g.productions = [(a, tuple(reversed(r))) for a, r in self.productions]
        return g


def grammar_terminals(cfg: CFG) -> list:
    used = sorted({s for r in cfg.nt_rules for s in r.rhs if isinstance(s, T)})
    out = []
    for t in [r.t for r in cfg.t_rules] + used:
        if t not in out:
            out.append(t)
    return out


def grammar_nonterminals(cfg: CFG) -> list:
    out = []
    for r in cfg.nt_rules:
        if r.nt not in out:
            out.append(r.nt)
    return out


def _check_grammar(cfg: CFG, nts: list) -> None:
    if not cfg.nt_rules:
        raise ValueError("CFG has no production rules.")
    prods = [(r.nt, tuple(r.rhs)) for r in cfg.nt_rules]
    if len(set(prods)) != len(prods):
        raise ValueError("The given grammar contains duplicate productions.")
    # closure of productive nonterminals (Grammar.hs:297-305)
    productive = set()
    changed = True
    while changed:
        changed = False
        for nt, rhs in prods:
            if nt not in productive and all(isinstance(s, T) or s in productive for s in rhs):
                productive.add(nt)
                changed = True
    bad = [n for n in nts if n not in productive]
    if bad:
        raise ValueError(f"The given grammar contains nonproductive productions due to {bad}.")


def build_grammar(cfg: CFG) -> Grammar:
    ts = grammar_terminals(cfg)
    nts = grammar_nonterminals(cfg)
    _check_grammar(cfg, nts)
    # encodeSymbols: fts ++ ignore ++ unused ++ |- ++ -|
    fts = [t for t in ts if t != IGNORE]
    names = []
    t_id = {}
    for t in fts:
        t_id[t] = len(names)
        names.append(repr(t))
    ignore = None
    if IGNORE in ts:
        ignore = len(names)
        t_id[IGNORE] = ignore
        names.append("ignore")
    empty = len(names)
    names.append("<unused>")
    start_terminal = len(names)
    names.append("|-")
    end_terminal = len(names)
    names.append("-|")
    num_terminals = len(names)
    nt_of_t = {}
    for t in ts:
        nt_of_t[t] = len(names)
        names.append(f"<{t!r}>")
    nt_id = {}
    for n in nts:
        nt_id[n] = len(names)
        names.append(n.text)
    start_nt = len(names)
    names.append("<start>")
    sym = Symbols(names=names, num_terminals=num_terminals, ignore=ignore, empty=empty,
                  start_terminal=start_terminal, end_terminal=end_terminal, start_nt=start_nt,
                  t_id=t_id, nt_of_t=nt_of_t, nt_id=nt_id)

    def as_nt(s):
        return nt_of_t[s] if isinstance(s, T) else nt_id[s]

    productions = []
    pnames = []
    for i, r in enumerate(cfg.nt_rules):
        productions.append((nt_id[r.nt], tuple(as_nt(s) for s in r.rhs)))
        pnames.append(r.name if r.name is not None else f"{r.nt.text}_{i}")
    for t in ts:
        i = len(productions)
        productions.append((nt_of_t[t], (t_id[t],)))
        pnames.append(f"Literal_{i}" if t.is_lit else f"{t.text}_{i}")
    i = len(productions)
    productions.append((start_nt, (start_terminal, nt_id[nts[0]], end_terminal)))
    pnames.append(f"start_{i}")
    arities = [sum(1 for s in rhs if s >= num_terminals) for _, rhs in productions]
    p2t = []
    inv_nt_of_t = {v: k for k, v in nt_of_t.items()}
    for lhs, _ in productions:
        p2t.append(t_id[inv_nt_of_t[lhs]] if lhs in inv_nt_of_t else None)
    return Grammar(sym=sym, start=start_nt, terminals=list(range(num_terminals)),
                   nonterminals=list(range(num_terminals, len(names))), productions=productions,
                   production_names=pnames, arities=arities, production_to_terminal=p2t,
                   q=cfg.lookback, k=cfg.lookahead)


def terminal_names_both(g: Grammar) -> list:
    """Names of terminal IDs 0..empty for lexer+parser / parser modes (Analyzer.hs:97-104)."""
    out = []
    inv = {v: k for k, v in g.sym.t_id.items()}
    for i in range(g.sym.empty + 1):
        if i == g.sym.empty:
            out.append(f"empty_{i}")
        else:
            t = inv[i]
            out.append(f"literal_{i}" if t.is_lit else t.text)
    return out


def lexer_spec(cfg: CFG) -> list:
    """Token regexes in priority order (``cfgToDFALexerSpec``, CFG.hs:146-170)."""
    declared = [r.t for r in cfg.t_rules]
    used = sorted({s for r in cfg.nt_rules for s in r.rhs if isinstance(s, T)})
    implicit = [t for t in used if t not in declared]
    rules = []
    for t in implicit:
        if not t.is_lit:
            raise ValueError(f"Can not create literal from: {t.text}")
        rules.append((t, literal_regex(t.text)))
    rules += [(r.t, r.regex) for r in cfg.t_rules]
    if not rules:
        raise ValueError("CFG has no lexical rules.")
    for t, r in rules:
        if produces_epsilon(r):
            raise ValueError(f"Error: {t!r} may not produce empty strings.")
    # Map.fromList keeps the last regex and order of a duplicated terminal.
    order = {}
    regex = {}
    for i, (t, r) in enumerate(rules):
        order[t] = i
        regex[t] = r
    return [(t, regex[t], order[t]) for t in regex]


def lexer_only_ids(spec: list) -> tuple[dict, int, int | None, list]:
    """Terminal IDs of a lexer-only artifact (``encodeTerminals``, Encode.hs:148-162).

    Returns (T -> id, dead id, ignore id or None, names).
    """
    ts = sorted(t for t, _, _ in spec)
    fts = [t for t in ts if t != IGNORE]
    ids = {}
    names = []
    for t in fts:
        ids[t] = len(names)
        names.append(f"literal_{len(names)}" if t.is_lit else t.text)
    ignore = None
    if IGNORE in ts:
        ignore = len(names)
        ids[IGNORE] = ignore
        names.append("ignore")
    dead = len(names)
    names.append(f"empty_{dead}")
    return ids, dead, ignore, names


def load_cfg(path_or_text: str) -> CFG:
    if "\n" not in path_or_text and path_or_text.endswith(".alp"):
        with open(path_or_text, "r", encoding="utf-8") as f:
            return parse_alp(f.read())
    return parse_alp(path_or_text)
