"""Parser for alpacc's ``.alp`` grammar files (src/Alpacc/CFG.hs:172-235).

File layout: optional ``params { lookback = q. lookahead = k. }`` (defaults q=0, k=1; CFG.hs:68-106),
then terminal rules ``name = /regex/.`` (CFG.hs:210-217), then nonterminal rules
``Name [Label] -> syms | syms | ... .`` where every alternative shares the optional label
(CFG.hs:219-224). Terminals are lowercase identifiers or string literals ``"..."``; nonterminals
are identifiers starting with an uppercase letter. ``#`` starts a line comment (CFG.hs:174-175).
"""
from __future__ import annotations

from dataclasses import dataclass, field

from .regex import parse_regex


@dataclass(frozen=True, order=True)
class T:
    """A terminal: ``kind`` 0 = named terminal (``T``), 1 = literal (``TLit``).

    Ordering matches the derived ``Ord T`` of Grammar.hs:90: every ``T`` sorts before every
    ``TLit``; within a constructor, by text.
    """
    kind: int
    text: str

    @property
    def is_lit(self) -> bool:
        return self.kind == 1

    def __repr__(self) -> str:
        return f'"{self.text}"' if self.is_lit else self.text


@dataclass(frozen=True, order=True)
class NT:
    text: str

    def __repr__(self) -> str:
        return self.text


@dataclass
class TRule:
    t: T
    regex: tuple


@dataclass
class NTRule:
    nt: NT
    name: str | None
    rhs: list            # list of T | NT


@dataclass
class CFG:
    lookback: int = 0
    lookahead: int = 1
    t_rules: list = field(default_factory=list)
    nt_rules: list = field(default_factory=list)


class AlpError(ValueError):
    pass


class _Src:
    def __init__(self, text: str):
        self.s = text
        self.i = 0

    def peek(self, k=0):
        j = self.i + k
        return self.s[j] if j < len(self.s) else ""

    def space(self):
        while self.i < len(self.s):
            c = self.s[self.i]
            if c.isspace():
                self.i += 1
            elif c == "#":
                while self.i < len(self.s) and self.s[self.i] != "\n":
                    self.i += 1
            else:
                break

    def err(self, msg):
        line = self.s.count("\n", 0, self.i) + 1
        return AlpError(f"line {line}: {msg}")

    def lit(self, tok: str) -> bool:
        if self.s.startswith(tok, self.i):
            self.i += len(tok)
            self.space()
            return True
        return False

    def expect(self, tok: str):
        if not self.lit(tok):
            raise self.err(f"expected {tok!r}")

    def ident(self, first) -> str | None:
        c = self.peek()
        if c and c.isascii() and first(c):
            j = self.i + 1
            while j < len(self.s) and (self.s[j] == "_" or (self.s[j].isascii() and self.s[j].isalnum())):
                j += 1
            out = self.s[self.i:j]
            self.i = j
            self.space()
            return out
        return None

    def decimal(self) -> int:
        j = self.i
        while j < len(self.s) and self.s[j].isdigit():
            j += 1
        if j == self.i:
            raise self.err("expected a number")
        v = int(self.s[self.i:j])
        self.i = j
        self.space()
        return v


def _lower(c):
    return "a" <= c <= "z"


def _upper(c):
    return "A" <= c <= "Z"


def parse_alp(text: str) -> CFG:
    src = _Src(text)
    src.space()
    cfg = CFG()
    # pParams
    save = src.i
    if src.lit("params"):
        src.expect("{")
        while True:
            if src.lit("lookback"):
                src.expect("=")
                cfg.lookback = src.decimal()
                src.expect(".")
            elif src.lit("lookahead"):
                src.expect("=")
                cfg.lookahead = src.decimal()
                src.expect(".")
            else:
                break
        src.expect("}")
    else:
        src.i = save
    # many (try pTRule)
    while True:
        save = src.i
        name = src.ident(_lower)
        if name is None or not src.lit("="):
            src.i = save
            break
        if not src.lit("/"):
            src.i = save
            break
        regex, end = parse_regex(src.s, src.i)
        src.i = end
        if not src.lit("/") or not src.lit("."):
            src.i = save
            break
        cfg.t_rules.append(TRule(T(0, name), regex))
    # many pNTRule
    while src.i < len(src.s):
        nt = src.ident(_upper)
        if nt is None:
            raise src.err("expected a nonterminal rule")
        label = None
        if src.peek() == "[":
            src.i += 1
            label = src.ident(_upper)
            if label is None:
                raise src.err("production name must start with an uppercase letter")
            if src.peek() != "]":
                raise src.err("expected ]")
            src.i += 1
            src.space()
        src.expect("->")
        alts = [[]]
        while True:
            c = src.peek()
            if c == '"':
                j = src.i + 1
                while j < len(src.s) and src.s[j] != '"' and src.s[j].isprintable():
                    j += 1
                if j == src.i + 1 or j >= len(src.s) or src.s[j] != '"':
                    raise src.err("bad string literal")
                alts[-1].append(T(1, src.s[src.i + 1:j]))
                src.i = j + 1
                src.space()
                continue
            t = src.ident(_lower)
            if t is not None:
                alts[-1].append(T(0, t))
                continue
            n = src.ident(_upper)
            if n is not None:
                alts[-1].append(NT(n))
                continue
            if src.lit("|"):
                alts.append([])
                continue
            break
        src.expect(".")
        for rhs in alts:
            cfg.nt_rules.append(NTRule(NT(nt), label, rhs))
    return cfg
