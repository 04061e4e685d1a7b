"""Token streams rejected by the bracket check of brackets_matches (parser.fut:174-184) and nothing
else: every key has a pair entry and the depths are valid, but some right bracket differs from the
left bracket at its previous smaller-or-equal depth.

The device cancels a position's trailing left brackets against the next position's leading right
brackets when it expands the bracket stream, then the remaining trailing lefts against the
remaining leading rights of the next position that still has brackets (positions inside one
256-position group of a 1024-position tile), checks those pairs there, and leaves the rest to the
bracket search. ``classify`` names where a mismatch sits, so that the tests can show every place
is exercised: "adjacent" (a first-level pair mismatches), "second" (a second-level pair) or
"remaining" (only the search sees it). It classifies
with the literal restatement of parser.fut; the verdicts themselves are compared with the C
restatement.

Mutations: one token swapped with another, or one token's kind replaced, in the token stream of a
valid document; ``byte_swaps`` does the swap on the document's bytes (kept when the lexer gives
exactly the swapped kinds), for the entries that take bytes.
"""
from __future__ import annotations

import numpy as np

from oracle.lexer_fut import LexerFut
from oracle.parser_fut import ParserFut

GROUP = 256  # positions of one cancellation group (a wavefront's 64 lanes x 4 positions)

# Two bracket kinds: a list closed by the other kind's bracket passes every key (each key sees one
# bracket) and fails only in the bracket check.
TWO_BRACKETS = """params {
  lookback = 1.
  lookahead = 1.
}

ignore = /\\s+/.
atom = /[a-z]+/.

E -> atom | "(" L ")" | "[" L "]".
L -> | E L.
"""


def token_kinds(bundle) -> list:
    """The terminals the lexer can emit (every terminal but the ignored and the empty one)."""
    ign = bundle.lexer.ignore
    return [i for i in range(len(bundle.terminal_names))
            if i != ign and i != bundle.parser.empty_terminal]


def short_streams(bundle, max_len: int):
    """Every token kind stream of length 0 .. max_len."""
    ks = token_kinds(bundle)
    yield np.zeros(0, dtype=np.uint32)
    for n in range(1, max_len + 1):
        for idx in np.ndindex(*([len(ks)] * n)):
            yield np.array([ks[i] for i in idx], dtype=np.uint32)


def boundary_doc(k: int) -> bytes:
    """S-expression ``( a a ... a ) b`` (k atoms): two top-level expressions, so the bracket of
    ``)`` and the leading bracket of ``b`` mismatch. Their positions are k + 2 and k + 3: with
    k = 252 (or 1020, ...) the pair straddles two 256-position groups and is not cancelled."""
    return ("( " + "a " * k + ") b").encode()


def classify(bundle, kinds) -> str:
    """"key", "depth", "adjacent", "second", "remaining" or "accept" for a token kind stream."""
    ps = ParserFut(bundle.parser)
    spans = ps.lookup(ps.keys(np.asarray(kinds, dtype=np.int64)))
    if (spans == -1).any():
        return "key"
    pt = bundle.parser
    top = 1 << (pt.bracket_bits - 1)
    seqs = [[int(x) for x in pt.flat_stacks[s0:s1]] for s0, s1 in spans[:, :2]]
    flat = np.array([x for s in seqs for x in s], dtype=np.int64)
    if flat.shape[0]:
        sc = np.cumsum(np.where((flat & top) != 0, 1, -1))
        if (sc < 0).any() or sc[-1] != 0:
            return "depth"
    if ps.brackets_matches(flat):
        return "accept"
    n = len(seqs)
    lr = [0] * n
    tl = [0] * n
    for p, s in enumerate(seqs):
        while lr[p] < len(s) and not (s[lr[p]] & top):
            lr[p] += 1
        while tl[p] < len(s) and s[len(s) - 1 - tl[p]] & top:
            tl[p] += 1
    # first level: a position's trailing lefts against the next position's leading rights
    cr = [0] * n
    for p in range(n - 1):
        if p // GROUP != (p + 1) // GROUP:
            continue
        cr[p] = min(tl[p], lr[p + 1])
        a, b = seqs[p], seqs[p + 1]
        for j in range(cr[p]):
            if (a[len(a) - 1 - j] & ~top) != (b[j] & ~top):
                return "adjacent"
    # second level: the remaining trailing lefts against the remaining leading rights of the next
    # position of the group that still has brackets
    cl = [0] + cr[:-1]
    left = [len(seqs[p]) - cl[p] - cr[p] for p in range(n)]
    for p in range(n):
        q = p + 1
        while q < n and q // GROUP == p // GROUP and left[q] == 0:
            q += 1
        if q >= n or q // GROUP != p // GROUP:
            continue
        c2 = min(tl[p] - cr[p], lr[q] - cl[q])
        a, b = seqs[p], seqs[q]
        for j in range(c2):
            if (a[len(a) - 1 - cr[p] - j] & ~top) != (b[cl[q] + j] & ~top):
                return "second"
    return "remaining"


def cancel_positions(bundle, kinds) -> np.ndarray:
    """Tokens next to a position pair whose brackets cancel (token index = position - 1)."""
    ps = ParserFut(bundle.parser)
    spans = ps.lookup(ps.keys(np.asarray(kinds, dtype=np.int64)))
    pt = bundle.parser
    top = 1 << (pt.bracket_bits - 1)
    out = []
    for p in range(spans.shape[0] - 1):
        a = pt.flat_stacks[spans[p, 0]:spans[p, 1]]
        b = pt.flat_stacks[spans[p + 1, 0]:spans[p + 1, 1]]
        if a.shape[0] and b.shape[0] and (a[-1] & top) and not (b[0] & top):
            out.append(p)
    return np.array(out, dtype=np.int64)


def token_mutations(bundle, doc: bytes, seed: int, trials: int, far: bool = False,
                    near_cancel: bool = False):
    """(kinds, starts, ends, label) of mutated token streams of ``doc`` whose only fault is a
    bracket mismatch. far=True swaps tokens at least 2048 positions apart; near_cancel=True
    mutates tokens within three of a position pair whose brackets cancel."""
    lx = LexerFut(bundle.lexer)
    kinds, starts, ends = (np.asarray(x) for x in lx.lex_int(doc))
    kinds = kinds.astype(np.int64)
    n = kinds.shape[0]
    rng = np.random.Generator(np.random.PCG64(seed))
    nk = int(kinds.max()) + 1
    near = cancel_positions(bundle, kinds) if near_cancel else None
    out = []
    for _ in range(trials):
        k2 = kinds.copy()
        if near is not None and near.shape[0]:
            i = int(np.clip(near[rng.integers(near.shape[0])] + rng.integers(-3, 4), 0, n - 1))
            if rng.integers(2) == 0:
                j = int(np.clip(i + rng.integers(1, 4), 0, n - 1))
                k2[i], k2[j] = k2[j], k2[i]
            else:
                k2[i] = int(rng.integers(nk))
        elif rng.integers(2) == 0 or far:
            i = int(rng.integers(n))
            j = int(rng.integers(n))
            if far and abs(i - j) < 2048:
                continue
            k2[i], k2[j] = k2[j], k2[i]
        else:
            k2[int(rng.integers(n))] = int(rng.integers(nk))
        lab = classify(bundle, k2)
        if lab in ("adjacent", "second", "remaining"):
            out.append((k2.astype(np.uint32), starts.astype(np.uint32), ends.astype(np.uint32), lab))
    return out


def byte_swaps(bundle, doc: bytes, seed: int, trials: int):
    """(document, label) with the texts of two tokens exchanged, rejected by a bracket mismatch."""
    lx = LexerFut(bundle.lexer)
    kinds, starts, ends = (np.asarray(x) for x in lx.lex_int(doc))
    n = kinds.shape[0]
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(trials):
        i, j = sorted(int(x) for x in rng.integers(n, size=2))
        if i == j or kinds[i] == kinds[j]:
            continue
        si, ei, sj, ej = int(starts[i]), int(ends[i]), int(starts[j]), int(ends[j])
        d2 = doc[:si] + doc[sj:ej] + doc[ei:sj] + doc[si:ei] + doc[ej:]
        k2 = kinds.astype(np.int64).copy()
        k2[i], k2[j] = k2[j], k2[i]
        r = lx.lex_int(d2)
        if r is None or not np.array_equal(np.asarray(r[0]).astype(np.int64), k2):
            continue
        lab = classify(bundle, k2)
        if lab in ("adjacent", "second", "remaining"):
            out.append((d2, lab))
    return out
