"""Endomorphism tables for the data-parallel lexer.

Restates src/Alpacc/Lexer/{DFAParallelLexer,ParallelLexing,Encode}.hs:
  * every byte b induces an endomorphism on the lexer DFA's states plus a dead state 0: the move
    function ``f_b`` (missing moves go to 0) and, per source state, whether the move is a producing
    one (``endomorphismTable``, DFAParallelLexer.hs:39-68);
  * composing ``a`` then ``b`` gives ``c[i] = b[a[i]]`` and ``c'[i] = b'[a[i]]``
    (Semigroup instance, DFAParallelLexer.hs:70-80);
  * endomorphism 0 is the identity and 1 the dead endomorphism; the byte endomorphisms and their
    compositions get 2, 3, ... (ParallelLexing.hs:27-54, 375);
  * each endomorphism carries the data of the state it reaches from the DFA's initial state: the
    token label, whether that state accepts, and whether the last move was producing
    (``toEndoData``, ParallelLexing.hs:67-83);
  * the composition table is closed under composition; a composition the reference never reaches
    is dead (``listCompositions``, ParallelLexing.hs:117-124);
  * the state word packs (index, terminal, produce) from bit 0 upwards with field widths
    ``findBitSize`` of (#endomorphisms, #terminals, 1) (Lexer/Encode.hs:137-151, 221-243).

The closure here composes every pair of reachable endomorphisms (not only the pairs the reference's
"connected" worklist visits, ParallelLexing.hs:222-264), and every composition that maps all states
to the dead state is identified with the dead endomorphism. Both differences only change internal
endomorphism numbering: the value a scan produces at every position (token, accept, produce of the
state reached from the initial state) is the same. The worklist, run on the same DFA, reaches the
same non-dead endomorphisms (tests/test_endo_closure.py), so the full closure adds none.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .automata import LexerDFA

IDENTITY = 0
DEAD = 1


def find_bit_size(x: int) -> int:
    """``findBitSize`` (Lexer/Encode.hs:137-140): bits to represent x-1, at least 1 for x <= 2."""
    v = max(1, x - 1)
    return v.bit_length()


@dataclass
class EndoTables:
    num_endos: int                 # e (state_size)
    num_terminals: int
    index_bits: int
    terminal_bits: int
    state_bits: int                # index_bits + terminal_bits + 1
    compositions: np.ndarray       # uint32 [e*e]; entry [b*e + a] = encoded(a then b)
    transitions_to_states: np.ndarray   # uint32 [256]: encoded endomorphism of each byte
    accept: np.ndarray             # bool [e]
    identity: int                  # encoded identity word
    dead_terminal: int
    ignore_terminal: int           # -1 when the grammar has no ignore token
    # decoded per-endomorphism data (for tests / oracles)
    endo_terminal: np.ndarray      # int [e]
    endo_produce: np.ndarray       # bool [e]
    # raw functions (rows of dead + DFA states), for validation
    functions: np.ndarray          # int32 [e, m+1]; rows 0/1 are identity/dead
    byte_endo: np.ndarray          # int32 [256] endomorphism index of each byte
    produces: np.ndarray = None    # bool [e, m+1]: whether the move from each state produces

    def encode(self, idx: int) -> int:
        return (int(idx)
                | (int(self.endo_terminal[idx]) << self.index_bits)
                | (int(self.endo_produce[idx]) << (self.index_bits + self.terminal_bits)))

    @property
    def index_mask(self) -> int:
        return (1 << self.index_bits) - 1

    @property
    def terminal_offset(self) -> int:
        return self.index_bits

    @property
    def terminal_mask(self) -> int:
        return ((1 << self.terminal_bits) - 1) << self.index_bits

    @property
    def produce_offset(self) -> int:
        return self.index_bits + self.terminal_bits


def build_endo_tables(ld: LexerDFA, terminal_id: dict, num_terminals: int, dead_terminal: int,
                      ignore_terminal: int) -> EndoTables:
    """Closure of the byte endomorphisms of ``ld`` under composition and their encoding.

    ``terminal_id`` maps DFA token labels to output terminal IDs.
    """
    d = ld.dfa
    m = d.n
    # DFA states are shifted by one so that 0 is the dead state (DFAParallelLexer.hs:33-37).
    width = m + 1
    init = d.initial + 1
    accept_of = np.zeros(width, dtype=bool)
    term_of = np.full(width, dead_terminal, dtype=np.int64)
    for q in range(m):
        if q in d.accepting:
            accept_of[q + 1] = True
        lab = d.label[q]
        if lab is not None:
            term_of[q + 1] = terminal_id[lab]

    f_rows = []
    p_rows = []
    byte_key = []
    for b in range(256):
        f = np.zeros(width, dtype=np.int32)
        p = np.zeros(width, dtype=bool)
        for q in range(m):
            r = d.delta[q].get(b)
            if r is not None:
                f[q + 1] = r + 1
                p[q + 1] = (q, b) in ld.producing
        f_rows.append(f)
        p_rows.append(p)
        byte_key.append(f.tobytes() + p.tobytes())

    # Endomorphism rows: 0 identity, 1 dead, then distinct byte endomorphisms in byte order.
    ident_f = np.arange(width, dtype=np.int32)
    ident_p = np.zeros(width, dtype=bool)
    dead_f = np.zeros(width, dtype=np.int32)
    dead_p = np.zeros(width, dtype=bool)
    F = [ident_f, dead_f]
    P = [ident_p, dead_p]
    key_index = {}
    byte_endo = np.zeros(256, dtype=np.int32)
    for b in range(256):
        f, p = f_rows[b], p_rows[b]
        if not f.any():
            byte_endo[b] = DEAD
            continue
        k = byte_key[b]
        if k not in key_index:
            key_index[k] = len(F)
            F.append(f)
            P.append(p)
        byte_endo[b] = key_index[k]

    # Closure under composition (vectorised): c = a then b  ->  c_f = b_f[a_f], c_p = b_p[a_f].
    # Each round composes every pair with at least one operand added in the previous round.
    done = 2
    while done < len(F):
        Fa = np.stack(F)
        Pa = np.stack(P)
        e = len(F)
        for a in range(2, e):
            bs = np.arange(2, e) if a >= done else np.arange(done, e)
            # a then b, for every b in bs
            cf = Fa[bs][:, Fa[a]]
            cp = Pa[bs][:, Fa[a]]
            # b then a, for every old b (b new is covered by the row above when a is new)
            if a >= done:
                olds = np.arange(2, done)
                cf = np.concatenate([cf, Fa[a][Fa[olds]]])
                cp = np.concatenate([cp, Pa[a][Fa[olds]]])
            for j in range(cf.shape[0]):
                f = cf[j]
                if not f.any():
                    continue
                k = f.tobytes() + cp[j].tobytes()
                if k not in key_index:
                    key_index[k] = len(F)
                    F.append(f.copy())
                    P.append(cp[j].copy())
        done = e
        if len(F) > 1 << 16:
            raise ValueError("too many endomorphisms for a data-parallel lexer")

    Fa = np.stack(F)
    Pa = np.stack(P)
    e = len(F)
    # Full composition table: comp[b, a] = index of (a then b).
    lookup = {F[i].tobytes() + P[i].tobytes(): i for i in range(2, e)}
    comp = np.full((e, e), DEAD, dtype=np.int64)
    comp[IDENTITY, :] = np.arange(e)
    comp[:, IDENTITY] = np.arange(e)
    for a in range(2, e):
        cf = Fa[2:][:, Fa[a]]
        cp = Pa[2:][:, Fa[a]]
        for j in range(e - 2):
            f = cf[j]
            if not f.any():
                continue
            comp[2 + j, a] = lookup[f.tobytes() + cp[j].tobytes()]
    comp[IDENTITY, DEAD] = DEAD
    comp[DEAD, IDENTITY] = DEAD

    endo_terminal = np.array([term_of[Fa[i][init]] for i in range(e)], dtype=np.int64)
    endo_produce = np.array([bool(Pa[i][init]) for i in range(e)], dtype=bool)
    endo_accept = np.array([bool(accept_of[Fa[i][init]]) for i in range(e)], dtype=bool)
    # identity and dead carry no token, do not accept and do not produce (ParallelLexing.hs:32-51)
    for s in (IDENTITY, DEAD):
        endo_terminal[s] = dead_terminal
        endo_produce[s] = False
        endo_accept[s] = False

    ib = find_bit_size(e)
    tb = find_bit_size(num_terminals)
    tables = EndoTables(
        num_endos=e, num_terminals=num_terminals, index_bits=ib, terminal_bits=tb,
        state_bits=ib + tb + 1,
        compositions=np.zeros(e * e, dtype=np.uint32),
        transitions_to_states=np.zeros(256, dtype=np.uint32),
        accept=endo_accept, identity=0, dead_terminal=dead_terminal,
        ignore_terminal=ignore_terminal, endo_terminal=endo_terminal, endo_produce=endo_produce,
        functions=Fa, byte_endo=byte_endo, produces=Pa)
    enc = np.array([tables.encode(i) for i in range(e)], dtype=np.uint32)
    tables.compositions = enc[comp.reshape(-1)].astype(np.uint32)
    tables.transitions_to_states = enc[byte_endo].astype(np.uint32)
    tables.identity = int(enc[IDENTITY])
    return tables
