"""Generated table bundle for one grammar: what the reference's ``Analyzer`` record carries
(src/Alpacc/Generator/Analyzer.hs:33-82, built by ``mkLexer``/``mkParser``/``mkLexerParser``
:106-250), plus the device layouts the HIP backend consumes.

Device layouts (all little-endian, row-major):
  * ``byte_class``  u8[256]        byte -> byte class (bytes with the same endomorphism share one)
  * ``endo_step``   u16[e*C]       endomorphism index of (endo a, then one byte of class c)
  * ``endo_comp``   u16[e*e]       endomorphism index of (a then b), row a
  * ``endo_apply``  u8[e*(m+1)]    DFA state reached by applying endo a to state s (0 = dead)
  * ``dfa_step``    u16[(m+1)*C]   next state | produce << 8 | label(next) << 9
  * ``state_info``  u8[m+1]        accept flag (bit 0)
  * ``pair_index``  u32[NT^(q+k)]  dense LLP pair table: entry id or 0xFFFFFFFF (up to 2^26 keys)
  * ``pair_hash_keys`` u64[H], ``pair_hash_vals`` u32[H]: the pair table hashed instead (more
                   than 2^26 possible windows): the window's dense index at its multiplicative
                   hash slot (linear probing, H a power of two >= 2 * keys), entry id beside it
  * ``pair_entry``  u32[E*4]       (stack start, stack end, production start, production end)
  * ``stacks``      u16[]          brackets: symbol id | 0x8000 for a left (push) bracket
  * ``productions`` u16[]          production ids
"""
from __future__ import annotations

import hashlib
import json
from dataclasses import dataclass, field

import numpy as np

from .alp import CFG
from .automata import lexer_dfa, LexerDFA
from .endo import EndoTables, build_endo_tables, DEAD, IDENTITY
from .grammar import (Grammar, build_grammar, lexer_spec, lexer_only_ids, load_cfg,
                      terminal_names_both, IGNORE)
from .llp import LLPBuilder, LLPTable

MISSING = 0xFFFFFFFF
LEFT_BIT = 0x8000
DENSE_MAX = 1 << 26                 # largest dense pair index; beyond it the table is hashed
HASH_EMPTY = 0xFFFFFFFFFFFFFFFF     # empty slot of the hashed pair table
HASH_MUL = 0x9E3779B97F4A7C15       # multiplicative hash of a key's dense index


@dataclass
class LexerTables:
    terminal_names: list
    ignore: int                  # -1 when no ignore terminal
    dead_terminal: int
    dfa: LexerDFA
    endo: EndoTables
    token_ids: dict              # T -> terminal id
    # device layouts
    byte_class: np.ndarray
    endo_step: np.ndarray
    endo_comp: np.ndarray
    endo_apply: np.ndarray
    dfa_step: np.ndarray
    state_info: np.ndarray
    init_state: int              # DFA initial state in the shifted numbering (0 = dead)
    num_classes: int
    num_states: int              # m + 1 (with the dead state)


@dataclass
class ParserTables:
    q: int
    k: int
    num_terminals: int           # terminal symbol ids are [0, num_terminals)
    start_terminal: int
    end_terminal: int
    empty_terminal: int
    production_names: list
    arities: np.ndarray          # int32 per production
    production_to_terminal: np.ndarray   # int32 per production, -1 for none
    llp: LLPTable
    grammar: Grammar
    # device layouts
    pair_index: np.ndarray           # dense (empty when hashed)
    pair_hash_keys: np.ndarray       # hashed: u64 dense key index per slot (HASH_EMPTY: free)
    pair_hash_vals: np.ndarray       # hashed: u32 entry id per slot
    pair_entry: np.ndarray
    stacks: np.ndarray
    productions: np.ndarray
    # literal open-addressing table (Encode.hs:353-375, HashTable.hs:38-49), for the oracle
    hash_keys: np.ndarray        # int64 [size, q+k]
    hash_used: np.ndarray        # bool [size]
    hash_vals: np.ndarray        # int64 [size, 4]
    max_iters: int
    flat_stacks: np.ndarray      # int64 brackets in Map key order (reference layout)
    flat_productions: np.ndarray
    bracket_bits: int


@dataclass
class Bundle:
    name: str
    mode: str                    # "lexer", "parser" or "both"
    terminal_names: list
    lexer: LexerTables | None = None
    parser: ParserTables | None = None
    meta: dict = field(default_factory=dict)


def _int_type_bits(v: int) -> int:
    for b in (8, 16, 32, 64):
        if v < (1 << b):
            return b
    raise ValueError("value too large")


def _build_lexer(cfg: CFG, token_ids: dict, names: list, dead: int, ignore) -> LexerTables:
    spec = lexer_spec(cfg)
    prio = {t: o for t, _, o in spec}
    ld = lexer_dfa([(t, r) for t, r, _ in spec], prio)
    ign = -1 if ignore is None else ignore
    endo = build_endo_tables(ld, token_ids, len(names), dead, ign)
    e = endo.num_endos
    d = ld.dfa
    m1 = d.n + 1
    if m1 > 256:
        raise ValueError("lexer DFA has more than 255 states")
    # byte classes: bytes with identical endomorphisms
    cls_of_endo = {}
    byte_class = np.zeros(256, dtype=np.uint8)
    class_endo = []
    for b in range(256):
        be = int(endo.byte_endo[b])
        if be not in cls_of_endo:
            cls_of_endo[be] = len(class_endo)
            class_endo.append(be)
        byte_class[b] = cls_of_endo[be]
    C = len(class_endo)
    comp_ab = endo.compositions.reshape(e, e)      # [b, a] encoded words
    idx_mask = endo.index_mask
    comp_idx = (comp_ab & idx_mask).astype(np.uint16)  # [b, a] index
    endo_comp = np.ascontiguousarray(comp_idx.T)       # [a, b]
    endo_step = np.ascontiguousarray(endo_comp[:, class_endo])   # [a, c]
    endo_apply = endo.functions.astype(np.uint8)       # [e, m1]
    # identity/dead rows: identity maps s->s, dead maps all -> 0 (already so in functions)
    term_of_state = np.full(m1, dead, dtype=np.int64)
    acc = np.zeros(m1, dtype=np.uint8)
    for q in range(d.n):
        if d.label[q] is not None:
            term_of_state[q + 1] = token_ids[d.label[q]]
        if q in d.accepting:
            acc[q + 1] = 1
    if term_of_state.max() >= 128:
        raise ValueError("terminal ids must be < 128 for the device lexer")
    dfa_step = np.zeros((m1, C), dtype=np.uint16)
    for s in range(m1):
        for c, be in enumerate(class_endo):
            nxt = int(endo.functions[be][s])
            prod = bool(ld.producing and s > 0 and (s - 1, _any_byte_of_class(byte_class, c)) in ld.producing)
            dfa_step[s, c] = nxt | (int(prod) << 8) | (int(term_of_state[nxt]) << 9)
    return LexerTables(terminal_names=names, ignore=ign, dead_terminal=dead, dfa=ld, endo=endo,
                       token_ids=token_ids, byte_class=byte_class, endo_step=endo_step,
                       endo_comp=endo_comp, endo_apply=np.ascontiguousarray(endo_apply),
                       dfa_step=dfa_step.reshape(-1), state_info=acc, init_state=d.initial + 1,
                       num_classes=C, num_states=m1)


def _any_byte_of_class(byte_class: np.ndarray, c: int) -> int:
    return int(np.nonzero(byte_class == c)[0][0])


def _fnv(key) -> int:
    """``hash`` (Encode.hs:326-334 / parser.fut:110-115): FNV-1a style over terminal ids."""
    h = 14695981039346656037
    for a in key:
        h ^= int(a) & 0xFFFFFFFFFFFFFFFF
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def device_hash_slot(di: int, bits: int) -> int:
    """Home slot of dense key index di in a hashed pair table of 2**bits slots."""
    return ((di * HASH_MUL) & 0xFFFFFFFFFFFFFFFF) >> (64 - bits)


def _device_hash(entries) -> tuple[np.ndarray, np.ndarray]:
    """Open addressing with linear probing over (dense key index, entry id) pairs, at least twice
    as many slots as keys (a power of two); the empty key is HASH_EMPTY."""
    size = 16
    while size < 2 * len(entries):
        size *= 2
    bits = size.bit_length() - 1
    keys = np.full(size, HASH_EMPTY, dtype=np.uint64)
    vals = np.full(size, MISSING, dtype=np.uint32)
    for di, v in entries:
        h = device_hash_slot(di, bits)
        while keys[h] != np.uint64(HASH_EMPTY):
            h = (h + 1) & (size - 1)
        keys[h] = di
        vals[h] = v
    return keys, vals


def _build_parser(g: Grammar, pair_table: str = "auto") -> ParserTables:
    q, k = g.q, g.k
    tb = LLPBuilder(g, q, k).build()
    empty = g.sym.empty
    nT = g.sym.num_terminals
    # padLLPTableKeys (Encode.hs:259-282): left-pad suffix to q, right-pad prefix to k with empty
    padded = {}
    for (x, y), (br, pi) in tb.homomorphisms.items():
        xs = tuple(x[:q])
        xs = (empty,) * (q - len(xs)) + xs
        ys = tuple(y[:k])
        ys = ys + (empty,) * (k - len(ys))
        padded[xs + ys] = (br, pi)
    keys_sorted = sorted(padded)
    # reference flattening in Map key order (flattenTuple, Encode.hs:301-324)
    nsym = len(g.sym.names)
    bracket_bits = _int_type_bits(2 * (nsym - 1))
    flat_st, flat_pr, spans = [], [], []
    for key in keys_sorted:
        br, pi = padded[key]
        s0 = len(flat_st)
        flat_st.extend(((1 << (bracket_bits - 1)) | s) if left else s for left, s in br)
        p0 = len(flat_pr)
        flat_pr.extend(pi)
        spans.append((s0, len(flat_st), p0, len(flat_pr)))
    # open addressing (HashTable.hs:38-49): size 2*|m|, inserted by foldr (last key first)
    size = 2 * len(keys_sorted)
    used = np.zeros(size, dtype=bool)
    hkeys = np.zeros((size, q + k), dtype=np.int64)
    hvals = np.zeros((size, 4), dtype=np.int64)
    max_iters = 0
    if size:
        hkeys[:] = keys_sorted[0]
        hvals[:] = spans[0]
    for key, span in reversed(list(zip(keys_sorted, spans))):
        h = _fnv(key)
        p = 0
        while True:
            idx = ((p + h) & 0xFFFFFFFFFFFFFFFF) % size
            if not used[idx]:
                used[idx] = True
                hkeys[idx] = key
                hvals[idx] = span
                max_iters = max(max_iters, p + 1)
                break
            p += 1
    # device pair table: dense over all nT^(q+k) windows, or hashed by the window's dense index
    # when that would be too large (the reference always hashes: HashTable.hs:38-49)
    dense = nT ** (q + k)
    if pair_table not in ("auto", "dense", "hash"):
        raise ValueError(pair_table)
    hashed = pair_table == "hash" or (pair_table == "auto" and dense > DENSE_MAX)
    if not hashed and dense > DENSE_MAX:
        raise ValueError(f"dense pair table too large ({dense} entries)")
    if hashed and dense >= HASH_EMPTY:
        raise ValueError("pair table keys do not fit 64 bits")
    pair_index = np.full(0 if hashed else dense, MISSING, dtype=np.uint32)
    hentries = []
    entries = []
    stacks = []
    prods = []
    uniq = {}
    for key in keys_sorted:
        br, pi = padded[key]
        sig = (tuple(br), tuple(pi))
        if sig not in uniq:
            s0 = len(stacks)
            stacks.extend((LEFT_BIT | s) if left else s for left, s in br)
            p0 = len(prods)
            prods.extend(pi)
            uniq[sig] = len(entries)
            entries.append((s0, len(stacks), p0, len(prods)))
        di = 0
        for t in key:
            di = di * nT + t
        if hashed:
            hentries.append((di, uniq[sig]))
        else:
            pair_index[di] = uniq[sig]
    if hashed:
        pair_hash_keys, pair_hash_vals = _device_hash(hentries)
    else:
        pair_hash_keys = np.zeros(0, dtype=np.uint64)
        pair_hash_vals = np.zeros(0, dtype=np.uint32)
    if nsym >= LEFT_BIT:
        raise ValueError("too many grammar symbols for 15-bit brackets")
    p2t = np.array([-1 if t is None else t for t in g.production_to_terminal], dtype=np.int32)
    return ParserTables(
        q=q, k=k, num_terminals=nT, start_terminal=g.sym.start_terminal,
        end_terminal=g.sym.end_terminal, empty_terminal=empty,
        production_names=list(g.production_names), arities=np.array(g.arities, dtype=np.int32),
        production_to_terminal=p2t, llp=tb, grammar=g, pair_index=pair_index,
        pair_hash_keys=pair_hash_keys, pair_hash_vals=pair_hash_vals,
        pair_entry=np.array(entries, dtype=np.uint32).reshape(-1, 4),
        stacks=np.array(stacks, dtype=np.uint16), productions=np.array(prods, dtype=np.uint16),
        hash_keys=hkeys, hash_used=used, hash_vals=hvals, max_iters=max_iters,
        flat_stacks=np.array(flat_st, dtype=np.int64), flat_productions=np.array(flat_pr, dtype=np.int64),
        bracket_bits=bracket_bits)


def generate(source: str, mode: str = "both", name: str | None = None,
             pair_table: str = "auto") -> Bundle:
    """Build the bundle of a grammar (``mkLexer``/``mkParser``/``mkLexerParser``).

    ``source`` is a path to a ``.alp`` file or its text; ``mode`` is "lexer", "parser" or "both".
    ``pair_table``: the device pair table is dense ("dense"), hashed ("hash"), or dense up to
    DENSE_MAX windows and hashed beyond ("auto").
    """
    cfg = load_cfg(source)
    if name is None:
        name = source.rsplit("/", 1)[-1].removesuffix(".alp") if source.endswith(".alp") else "grammar"
    if mode == "lexer":
        spec = lexer_spec(cfg)
        ids, dead, ign, names = lexer_only_ids(spec)
        lex = _build_lexer(cfg, ids, names, dead, ign)
        return Bundle(name=name, mode=mode, terminal_names=names, lexer=lex)
    g = build_grammar(cfg)
    names = terminal_names_both(g)
    par = _build_parser(g, pair_table)
    lex = None
    if mode == "both":
        ids = dict(g.sym.t_id)
        lex = _build_lexer(cfg, ids, names, g.sym.empty, g.sym.ignore)
    elif mode != "parser":
        raise ValueError(mode)
    return Bundle(name=name, mode=mode, terminal_names=names, lexer=lex, parser=par)


_CACHE: dict = {}


def cached(source: str, mode: str = "both") -> Bundle:
    with open(source, "rb") if source.endswith(".alp") else _nullctx(source.encode()) as f:
        digest = hashlib.sha1(f.read()).hexdigest()
    key = (digest, mode)
    if key not in _CACHE:
        _CACHE[key] = generate(source, mode)
    return _CACHE[key]


class _nullctx:
    def __init__(self, data):
        self.data = data

    def __enter__(self):
        return self

    def __exit__(self, *a):
        return False

    def read(self):
        return self.data
