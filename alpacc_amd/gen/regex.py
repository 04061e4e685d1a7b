"""Regular-expression dialect of alpacc grammar files, parsed into a byte-level AST.

The dialect follows src/Alpacc/Lexer/RegularExpression.hs:
  * alternation ``R|S`` (left-associative chain, :169-170), concatenation by juxtaposition
    (``foldl Concat Epsilon``, :166-167), postfix ``*``, ``+`` (= R R*) and ``?`` (= eps | R),
    which may be stacked (:199-215);
  * groups ``( ... )``; ranges ``[a-z\\x00-\\x1F...]`` whose members are alphanumeric characters or
    ``\\xHH`` bytes, single or as ``lo-hi`` spans (:175-197);
  * escapes ``\\\\ \\/ \\[ \\] \\( \\) \\* \\+ \\| \\s \\t \\n \\r \\?`` (:87-103), ``\\uHEX`` (:118-123) and
    ``\\xHH`` (:125-130); any other printable, non-space character outside
    ``\\ / [ ] ( ) * | + ?`` is a literal (:132-139);
  * whitespace after every token is skipped (``lexeme``, :81-85).

Characters are UTF-8 encoded, so a literal is a byte string (``charToBytes``, :78-79).

AST nodes are tuples: ("eps",), ("lit", bytes), ("range", (bytes, ...)), ("star", r),
("alt", a, b), ("cat", a, b).
"""
from __future__ import annotations

_ESCAPES = {
    "\\": "\\", "/": "/", "[": "[", "]": "]", "(": "(", ")": ")", "*": "*",
    "+": "+", "|": "|", "s": " ", "t": "\t", "n": "\n", "r": "\r", "?": "?",
}
_SPECIAL = set("\\/[]()*|+?")
_HEX = set("0123456789abcdefABCDEF")

EPS = ("eps",)


class RegexError(ValueError):
    pass


def _is_print(c: str) -> bool:
    # Haskell's Data.Char.isPrint is false for control and unassigned characters; for the
    # byte/ASCII range used by grammars this is the same as str.isprintable().
    return c.isprintable()


def _is_space(c: str) -> bool:
    return c.isspace()


class _P:
    def __init__(self, text: str, pos: int = 0):
        self.s = text
        self.i = pos

    def peek(self, k: int = 0) -> str:
        j = self.i + k
        return self.s[j] if j < len(self.s) else ""

    def skip_space(self) -> None:
        while self.i < len(self.s) and _is_space(self.s[self.i]):
            self.i += 1

    def error(self, msg: str) -> RegexError:
        return RegexError(f"regex error at offset {self.i}: {msg} in {self.s!r}")

    # pChar (:141-148): escape | \u | \x | printable literal; returns the UTF-8 bytes.
    def char(self) -> bytes | None:
        c = self.peek()
        if c == "\\":
            n = self.peek(1)
            if n in _ESCAPES:
                self.i += 2
                self.skip_space()
                return _ESCAPES[n].encode("utf-8")
            if n == "u" and self.peek(2) in _HEX and self.peek(2) != "":
                j = self.i + 2
                while j < len(self.s) and self.s[j] in _HEX:
                    j += 1
                cp = int(self.s[self.i + 2:j], 16)
                self.i = j
                self.skip_space()
                return chr(cp).encode("utf-8", "surrogatepass")
            if n == "x" and self.peek(2) in _HEX and self.peek(3) in _HEX and self.peek(2) and self.peek(3):
                b = int(self.s[self.i + 2:self.i + 4], 16)
                self.i += 4
                self.skip_space()
                return bytes([b])
            raise self.error("bad escape")
        if c and c not in _SPECIAL and _is_print(c) and not _is_space(c):
            self.i += 1
            self.skip_space()
            return c.encode("utf-8")
        return None

    def byte(self) -> int | None:
        if self.peek() == "\\" and self.peek(1) == "x" and self.peek(2) in _HEX and self.peek(3) in _HEX \
                and self.peek(2) and self.peek(3):
            b = int(self.s[self.i + 2:self.i + 4], 16)
            self.i += 4
            self.skip_space()
            return b
        return None

    def alnum(self) -> str | None:
        c = self.peek()
        if c and c.isalnum():
            self.i += 1
            self.skip_space()
            return c
        return None

    # pRange (:175-197)
    def range_(self):
        assert self.peek() == "["
        self.i += 1
        self.skip_space()
        members: list[bytes] = []
        while True:
            save = self.i
            a = self.alnum()
            if a is not None:
                if self.peek() == "-":
                    save2 = self.i
                    self.i += 1
                    self.skip_space()
                    b = self.alnum()
                    if b is not None:
                        members.extend(chr(c).encode("utf-8", "surrogatepass")
                                       for c in range(ord(a), ord(b) + 1))
                        continue
                    self.i = save2
                members.append(a.encode("utf-8"))
                continue
            self.i = save
            x = self.byte()
            if x is not None:
                if self.peek() == "-":
                    save2 = self.i
                    self.i += 1
                    self.skip_space()
                    y = self.byte()
                    if y is not None:
                        members.extend(bytes([v]) for v in range(x, y + 1))
                        continue
                    self.i = save2
                members.append(bytes([x]))
                continue
            break
        if not members:
            raise self.error("empty range")
        if self.peek() != "]":
            raise self.error("expected ]")
        self.i += 1
        self.skip_space()
        return ("range", tuple(members))

    # pTerm (:199-215)
    def term(self):
        c = self.peek()
        if c == "[":
            t = self.range_()
        elif c == "(":
            self.i += 1
            self.skip_space()
            t = self.alter()
            if self.peek() != ")":
                raise self.error("expected )")
            self.i += 1
            self.skip_space()
        else:
            b = self.char()
            if b is None:
                return None
            t = ("lit", b)
        while True:
            c = self.peek()
            if c == "*":
                t = ("star", t)
            elif c == "+":
                t = ("cat", t, ("star", t))
            elif c == "?":
                t = ("alt", EPS, t)
            else:
                return t
            self.i += 1
            self.skip_space()

    # pConcat (:166-167)
    def concat(self):
        r = EPS
        while True:
            t = self.term()
            if t is None:
                return r
            r = ("cat", r, t)

    # pAlter (:169-170)
    def alter(self):
        r = self.concat()
        while self.peek() == "|":
            self.i += 1
            self.skip_space()
            r = ("alt", r, self.concat())
        return r


def parse_regex(text: str, pos: int = 0) -> tuple[tuple, int]:
    """Parse a regex starting at ``pos``; returns (ast, position after the regex)."""
    p = _P(text, pos)
    p.skip_space()
    r = p.alter()
    return r, p.i


def regex_from_text(text: str):
    r, end = parse_regex(text)
    if end != len(text):
        raise RegexError(f"trailing input at {end} in {text!r}")
    return r


def literal_regex(s: str):
    """Regex of an implicit literal terminal ``"..."`` (src/Alpacc/CFG.hs:151-153)."""
    parts = [("lit", c.encode("utf-8")) for c in s]
    r = parts[0]
    for p in parts[1:]:
        r = ("cat", r, p)
    return r


def produces_epsilon(r) -> bool:
    """RegularExpression.hs:55-61."""
    tag = r[0]
    if tag == "eps" or tag == "star":
        return True
    if tag in ("lit", "range"):
        return False
    if tag == "alt":
        return produces_epsilon(r[1]) or produces_epsilon(r[2])
    return produces_epsilon(r[1]) and produces_epsilon(r[2])
