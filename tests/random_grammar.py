"""Seeded random grammars and sentences for parser fuzzing (the role of the reference's
src/Alpacc/Random.hs with tests/testfuthark.sh:95-156: random grammars, kept when they are LLP(q, k),
parsed by every implementation and compared). Terminals are single-character literals, so the
lexer stays trivial and the fuzzing exercises the pair tables, brackets and parents."""
import numpy as np

LETTERS = "abcdefgh"


def grammar(seed: int, q: int, k: int):
    """(alp source, productions) of a random grammar, or None when it has an unproductive
    nonterminal. productions: {nonterminal: [list of symbols]}; terminals are quoted literals."""
    rng = np.random.default_rng(seed)
    nt = [f"N{i}" for i in range(int(rng.integers(2, 5)))]
    terms = [LETTERS[i] for i in range(int(rng.integers(2, 6)))]
    prods = {}
    for a in nt:
        alts = []
        for _ in range(int(rng.integers(1, 4))):
            rhs = []
            for _ in range(int(rng.integers(0, 4))):
                if rng.random() < 0.6:
                    rhs.append('"' + terms[int(rng.integers(0, len(terms)))] + '"')
                else:
                    rhs.append(nt[int(rng.integers(0, len(nt)))])
            if rhs not in alts:
                alts.append(rhs)
        prods[a] = alts
    # every nonterminal must derive some terminal string
    productive = set()
    changed = True
    while changed:
        changed = False
        for a, alts in prods.items():
            if a not in productive and any(all(s.startswith('"') or s in productive for s in r)
                                           for r in alts):
                productive.add(a)
                changed = True
    if len(productive) != len(nt):
        return None
    lines = ["params {", f"  lookback = {q}.", f"  lookahead = {k}.", "}", "ignore = /\\s/."]
    for a in nt:
        lines.append(f"{a} -> " + " | ".join(" ".join(r) for r in prods[a]) + ".")
    return "\n".join(lines) + "\n", prods


def _heights(prods):
    """Fewest levels each nonterminal needs to derive a terminal string."""
    height = {a: None for a in prods}
    changed = True
    while changed:
        changed = False
        for a, alts in prods.items():
            for r in alts:
                hs = [0 if s.startswith('"') else height[s] for s in r]
                if all(h is not None for h in hs):
                    h = 1 + max(hs, default=0)
                    if height[a] is None or h < height[a]:
                        height[a] = h
                        changed = True
    return height


def sentence(prods, seed: int, target: int) -> bytes:
    """A random leftmost derivation from N0: random alternatives until about `target` symbols are
    out or pending (only alternatives with nonterminals where there are any; a step cap guards
    unit cycles),
    then the alternatives that end soonest, so the derivation terminates. Iterative: no recursion
    limit."""
    rng = np.random.default_rng(seed)
    height = _heights(prods)

    def h_of(r):
        return 1 + max([0 if s.startswith('"') else height[s] for s in r], default=0)

    out = []
    stack = ["N0"]
    steps = 0
    while stack:
        s = stack.pop()
        if s.startswith('"'):
            out.append(s[1:-1])
            continue
        steps += 1
        alts = prods[s]
        if len(out) + len(stack) < target and steps < 20 * target + 100:
            rec = [r for r in alts if any(not x.startswith('"') for x in r)]
            if rec:
                alts = rec
        else:
            best = min(h_of(r) for r in alts)
            alts = [r for r in alts if h_of(r) == best]
        r = alts[int(rng.integers(0, len(alts)))]
        stack.extend(reversed(r))
    return " ".join(out).encode()


def corrupt(doc: bytes, seed: int) -> bytes:
    """Delete one token or swap two neighbouring ones."""
    toks = doc.split()
    if len(toks) < 2:
        return doc + b" a"
    rng = np.random.default_rng(seed)
    i = int(rng.integers(0, len(toks) - 1))
    if rng.random() < 0.5:
        del toks[i]
    else:
        toks[i], toks[i + 1] = toks[i + 1], toks[i]
    return b" ".join(toks)
