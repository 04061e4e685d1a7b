"""An independent pin of the chunk loop of lex_int_flag (futhark/lexer.fut:144-172) for inputs longer
than one chunk, where no reference output exists to compare against (no CPU tests need a GPU).

The expected output is derived here from first principles, not by simulating the loop:

* ``traverse`` (lexer.fut:73-81) composes element 0 of every chunk with the state the previous chunk
  ended in, so the states of the whole run are exactly the states of ONE unchunked run over the
  virtual string V = chunk_0 ++ chunk_1 ++ ... (chunk j = str[j*C : min(j*C + C + 1, n)], so the
  byte shared by neighbouring chunks appears twice). One sequential ``tokenize`` of V (DFA.hs:415-447,
  ignore tokens kept) therefore gives every token boundary, and a walk of the same DFA gives the
  terminal of every state: the label of the DFA state reached, or the dead terminal where that
  state is not accepting (inside a token that has not matched yet).
* ``lex_step`` flags the last element of a non-ignore token unless it is the chunk's last element
  (``i != n - 1``, lexer.fut:104-107); its start is the last non-ignore token start at or before it,
  carried across chunks (lexer.fut:112-119, 137-140), in original coordinates.
* The map at lexer.fut:131-133 shifts every slot after every chunk, so a slot written in chunk j
  carries start + sum_{j' > j} j'*C and end + sum_{j' > j} (1 + j'*C). The sums are taken
  term by term below (the device and both restatements use closed forms of them).
* Every element of chunk j >= 1 that is not flagged is scattered to prev_size - 1 (offset -1,
  lexer.fut:108-111, 129-130); of these the chunk's last element is taken as the one kept (the
  sequential order of the C backend: the highest index wins; Futhark leaves it unspecified).
* After the loop the final token (terminal of the last state, (carried start, n)) is appended
  unshifted unless it is the ignore token, and the verdict is the last state's acceptance
  (lexer.fut:158-172).

The derivation is compared with the literal numpy restatement (oracle/lexer_fut.py) on every string
up to length 5 over one byte per byte class at chunk sizes 1, 2 and 3, and with it and the C
restatement on seeded documents, valid and byte-flipped, at chunk sizes from 1 to 4099 bytes.
"""
import itertools
import os

import numpy as np
import pytest

from conftest import GRAMMARS


@pytest.fixture(scope="module")
def bundles():
    from alpacc_amd.gen.tables import generate
    return {g: generate(os.path.join(GRAMMARS, f"{g}.alp")) for g in ("json", "arithmetic", "sexp")}


def chunked_from_virtual(bundle, data: bytes, C: int):
    """lex_int_flag C data, derived from one tokenize of the virtual string; None on rejection,
    else a list of (kind, start, end)."""
    from alpacc_amd.gen.automata import tokenize
    lt = bundle.lexer
    n = len(data)
    if n == 0:  # no chunk runs: the identity state decides
        toks = tokenize(lt.dfa, b"", None)
        return None if toks is None else []
    K = -(-n // C)
    offs = [j * C for j in range(K)]
    lens = [min(C + 1, n - o) for o in offs]
    V = b"".join(data[o:o + L] for o, L in zip(offs, lens))
    nV = len(V)
    assert nV == n + K - 1
    toks = tokenize(lt.dfa, V, None)
    if toks is None:
        return None
    ign = lt.ignore
    # the terminal of the state after every position: the label of the DFA state reached (the
    # producing transitions are edges of the DFA), the dead terminal where that state is not
    # accepting (inside a token that has not matched yet)
    d = lt.dfa.dfa
    kind_at = []
    q = d.initial
    for ch in V:
        q = d.delta[q][ch]  # defined: tokenize accepted V
        kind_at.append(lt.dead_terminal if d.label[q] is None else lt.token_ids[d.label[q]])
    # a token start counts when its first state's terminal is not the ignore terminal
    start_at = [0] * nV       # last such start at or before v (V coordinates), else 0
    cur = 0
    for label, (s, e) in toks:
        if kind_at[s] != ign:
            cur = s
        for v in range(s, e):
            start_at[v] = cur
    ends_at = {e - 1: lt.token_ids[label] for label, (s, e) in toks}

    def orig(v):  # the original position of virtual position v (chunks before the last: C + 1)
        return v - v // (C + 1)

    def shift_start(j):
        return sum(offs[jp] for jp in range(j + 1, K))

    def shift_end(j):
        return sum(1 + offs[jp] for jp in range(j + 1, K))

    dest = []
    base = 0
    for j in range(K):
        prev_size = len(dest)
        last = base + lens[j] - 1  # the chunk's last element
        for v in range(base, last):
            k = ends_at.get(v)
            if k is not None and k != ign:
                assert kind_at[v] == k  # a token ends in an accepting state of its own kind
                dest.append([k, orig(start_at[v]) + shift_start(j), orig(v) + 1 + shift_end(j)])
        if j >= 1 and prev_size >= 1:
            dest[prev_size - 1] = [kind_at[last], orig(start_at[last]) + shift_start(j),
                                   orig(last) + 1 + shift_end(j)]
        base += lens[j]
    if kind_at[nV - 1] != ign:
        dest.append([kind_at[nV - 1], orig(start_at[nV - 1]), n])
    return [tuple(x) for x in dest]


def _rows(r):
    return None if r is None else [tuple(int(v) for v in t) for t in zip(*r)]


def test_virtual_derivation_is_unchunked_for_one_chunk(bundles):
    """With one chunk the derivation is plain tokenize (the reference's own test oracle)."""
    from alpacc_amd import datagen
    from oracle import sequential
    b = bundles["json"]
    for seed in range(4):
        d = datagen.json_doc(3000, seed=seed).tobytes()
        for C in (3000, 5000):
            assert chunked_from_virtual(b, d, C) == sequential.lex(b, d)


@pytest.mark.parametrize("g,maxlen", [("sexp", 5), ("arithmetic", 5), ("json", 4)])
def test_literal_chunk_loop_exhaustive(bundles, g, maxlen):
    from oracle.lexer_fut import LexerFut
    b = bundles[g]
    lf = LexerFut(b.lexer)
    reps = [int(np.nonzero(b.lexer.byte_class == c)[0][0]) for c in range(b.lexer.num_classes)]
    if g == "json":  # one byte per class of a smaller set that still forms every token kind
        reps = sorted(set(reps) & set(b'"a1[]{},: \\.-t'))
    multi = 0
    for L in range(1, maxlen + 1):
        for tup in itertools.product(reps, repeat=L):
            d = bytes(tup)
            for C in (1, 2, 3):
                want = chunked_from_virtual(b, d, C)
                got = _rows(lf.lex_int(d, chunk_size=C))
                assert got == want, (d, C)
                multi += want is not None and L > C + 1 and len(want) > 1
    assert multi > 100  # accepted inputs of several chunks and tokens were compared


@pytest.mark.parametrize("g", ["json", "sexp", "arithmetic"])
def test_literal_chunk_loop_documents(bundles, g):
    from alpacc_amd import datagen
    from oracle.c import coracle
    from oracle.lexer_fut import LexerFut
    b = bundles[g]
    gen = {"json": datagen.json_doc, "arithmetic": datagen.arith_doc, "sexp": datagen.sexp_doc}[g]
    lf = LexerFut(b.lexer)
    orc = coracle.Oracle(b)
    accepted = 0
    for n, seed in ((600, 1), (2500, 2), (9000, 3)):
        doc = gen(n, seed=seed)
        for d in (doc.tobytes(), datagen.flip_bytes(doc, seed=seed, block=max(64, n // 4)).tobytes()):
            for C in (1, 5, 64, 255, 1000, 4099):
                want = chunked_from_virtual(b, d, C)
                assert _rows(lf.lex_int(d, chunk_size=C)) == want, (n, seed, C)
                assert _rows(orc.lex(d, chunk=C, wide=True)) == want, (n, seed, C)
                accepted += want is not None
    assert accepted > 0


def test_derivation_catches_planted_mistakes(bundles):
    """Mistakes of the kind the closed forms could hide, planted in a copy of the literal
    restatement on the CPU: each one must disagree with the derivation on some input."""
    from alpacc_amd import datagen
    from oracle.lexer_fut import LexerFut
    b = bundles["sexp"]
    docs = [datagen.sexp_doc(n, seed=s).tobytes() for n, s in ((300, 5), (800, 6))]

    class EndOffByOne(LexerFut):  # the re-map without the 1 + on the ends
        def lex_step(self, offset, *a):
            dest, st, start, size = super().lex_step(offset, *a)
            return [(t, s, e - 1) for t, s, e in dest], st, start, size

    class NoClobber(LexerFut):  # non-token elements not scattered to prev_size - 1
        def lex_step(self, offset, prev_state, prev_start, prev_size, dest, s):
            keep = dest[prev_size - 1] if prev_size >= 1 else None
            dest, st, start, size = super().lex_step(offset, prev_state, prev_start, prev_size,
                                                     dest, s)
            if keep is not None:
                t, a, e = keep
                dest[prev_size - 1] = (t, a + offset, e + 1 + offset)
            return dest, st, start, size

    class NoShiftOfOldSlots(LexerFut):  # the re-map applied only to the chunk's own slots
        def lex_step(self, offset, prev_state, prev_start, prev_size, dest, s):
            before = list(dest[:max(prev_size - 1, 0)])
            dest, st, start, size = super().lex_step(offset, prev_state, prev_start, prev_size,
                                                     dest, s)
            dest[:len(before)] = before
            return dest, st, start, size

    for bad in (EndOffByOne, NoClobber, NoShiftOfOldSlots):
        lf = bad(b.lexer)
        assert any(_rows(lf.lex_int(d, chunk_size=C)) != chunked_from_virtual(b, d, C)
                   for d in docs for C in (7, 64, 100)), bad.__name__
