"""Literal CPU restatement of futhark/lexer.fut (test infrastructure, not product code).

Follows ``mk_lexer`` line by line:
  * ``compose a b = compositions[idx(b)*e + idx(a)]`` (lexer.fut:67-71);
  * ``traverse``: element 0 is pre-composed with the carried state, then an inclusive scan with
    ``compose`` from the identity (lexer.fut:73-81);
  * ``lex_step``: token-end flags, ``(+)`` scan of the flags, take-right scan of token starts,
    scatter of ``(terminal, (start, end))`` and the re-map of the whole destination array
    (lexer.fut:91-142);
  * ``lex_int_flag``: the chunk loop ``for offset in 0..C..<n`` over ``str[offset : min(offset+C+1, n)]``,
    the final token and the accept test (lexer.fut:144-172).

``chunk_size = 0`` means one chunk over the whole input (the unchunked semantics, equal to the
sequential ``tokenize`` oracle). ``chunk_size = 16777216`` is the value the generated entries pass
(src/Alpacc/Generator/Futhark/Generator.hs:25,71). In the chunked mode a scatter writes the
non-token elements of chunk j > 0 to index ``prev_size - 1``; Futhark leaves the winner of
duplicate scatter indices unspecified, and this restatement uses the order of the sequential C
backend (the highest element index wins).
"""
from __future__ import annotations

import numpy as np

HIGHEST = np.iinfo(np.int64).max


class LexerFut:
    def __init__(self, lexer_tables):
        et = lexer_tables.endo
        self.e = et.num_endos
        self.comp = et.compositions.astype(np.int64)          # [b*e + a]
        self.trans = et.transitions_to_states.astype(np.int64)
        self.accept = et.accept.astype(bool)
        self.identity = int(et.identity)
        self.index_mask = et.index_mask
        self.t_off = et.terminal_offset
        self.t_mask = et.terminal_mask
        self.p_off = et.produce_offset
        self.ignore = int(lexer_tables.ignore)

    # lexer.fut:44-65
    def to_index(self, a):
        return a & self.index_mask

    def to_terminal(self, a):
        return (a & self.t_mask) >> self.t_off

    def is_produce(self, a):
        return ((a >> self.p_off) & 1).astype(bool)

    def is_ignore(self, t):
        if self.ignore < 0:
            return np.zeros_like(t, dtype=bool)
        return t == self.ignore

    def compose(self, a, b):
        return self.comp[self.to_index(b) * self.e + self.to_index(a)]

    def traverse(self, prev_state: int, s: np.ndarray) -> np.ndarray:
        n = s.shape[0]
        x = self.trans[s.astype(np.int64)]
        if n:
            x[0] = self.compose(np.int64(prev_state), x[0])
        # inclusive scan with a non-commutative operator (Hillis-Steele doubling)
        d = 1
        while d < n:
            y = x.copy()
            y[d:] = self.compose(x[:-d], x[d:])
            x = y
            d *= 2
        return x

    def lex_step(self, offset, prev_state, prev_start, prev_size, dest: list, s: np.ndarray):
        n = s.shape[0]
        states = self.traverse(prev_state, s)
        term = self.to_terminal(states)
        not_ign = ~self.is_ignore(term)
        prod = self.is_produce(states)
        flags = np.zeros(n, dtype=bool)
        if n > 1:
            flags[:-1] = prod[1:] & not_ign[:-1]
        is_ = np.cumsum(flags.astype(np.int64))
        offsets = np.where(flags, is_ - 1, -1)
        cand = np.where(prod & not_ign, np.arange(n, dtype=np.int64), HIGHEST)
        if n and cand[0] == HIGHEST:
            cand[0] = prev_start - offset          # take_right (prev_start - offset) highest
        # take-right scan: the last non-highest value
        idx = np.where(cand != HIGHEST, np.arange(n), -1)
        last = np.maximum.accumulate(idx) if n else idx
        starts = np.where(last >= 0, cand[np.maximum(last, 0)], HIGHEST)
        ends = np.arange(n, dtype=np.int64)
        size = int(is_[-1]) if n else 0
        need = prev_size + size + 1
        while len(dest) < need:
            dest.append((0, 0, 0))
        # scatter in element order (duplicate indices: the highest element index wins)
        tgt = offsets + prev_size
        for i in np.nonzero(tgt >= 0)[0]:
            dest[int(tgt[i])] = (int(term[i]), int(starts[i]), int(ends[i]))
        # the re-map of the whole destination array (lexer.fut:131-133)
        for j in range(len(dest)):
            t, a, b = dest[j]
            dest[j] = (t, a + offset, 1 + b + offset)
        last_state = int(states[-1])
        last_start = int(starts[-1]) if n else HIGHEST
        new_start = offset + last_start if last_start != HIGHEST else prev_start
        return dest, last_state, new_start, prev_size + size

    def lex_int(self, data, chunk_size: int = 0):
        """Returns None on rejection, else (kinds, starts, ends) as int64 arrays."""
        s = np.frombuffer(bytes(data), dtype=np.uint8) if not isinstance(data, np.ndarray) else data
        n = s.shape[0]
        C = n if chunk_size == 0 else chunk_size
        dest = [(0, 0, 0)]
        state, start, size = self.identity, 0, 0
        if n:
            for offset in range(0, n, max(C, 1)):
                m = min(offset + C + 1, n)
                dest, state, start, size = self.lex_step(offset, state, start, size, dest, s[offset:m])
        last_t = int(self.to_terminal(np.int64(state)))
        out = dest[:size]
        if not (self.ignore >= 0 and last_t == self.ignore):
            out = out + [(last_t, start, n)]
        if not self.accept[int(self.to_index(np.int64(state)))]:
            return None
        arr = np.array(out, dtype=np.int64).reshape(-1, 3)
        return arr[:, 0].copy(), arr[:, 1].copy(), arr[:, 2].copy()
