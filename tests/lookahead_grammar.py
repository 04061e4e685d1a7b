"""The LLP paper grammar (haskell-tests/LLP.hs:83-97: E -> T E', E' -> + T E' | eps, T -> a | [E])
as an .alp source with a chosen lookback q and lookahead k, and a seeded generator of its
sentences. Shared by the CPU pinning test of the C oracle and the GPU parity test."""
import numpy as np

ALP = """params {
  lookback = %d.
  lookahead = %d.
}
ignore = /\\s/.
E -> T Ep.
Ep [Plus] -> "+" T Ep.
Ep [Done] -> .
T [Atom] -> "a".
T [Group] -> "[" E "]".
"""

QK = [(1, 1), (2, 2), (3, 2), (2, 3), (3, 3), (1, 3), (3, 1)]


def source(q: int, k: int) -> str:
    return ALP % (q, k)


def sentence(n: int, seed: int) -> bytes:
    """A sentence of about n bytes: terms joined by '+', brackets nested up to depth 6, spaces."""
    rng = np.random.default_rng(seed)

    def term(depth):
        if depth <= 0 or rng.random() < 0.55:
            return "a"
        return "[" + expr(depth - 1, int(rng.integers(1, 5))) + "]"

    def expr(depth, count):
        return "+".join(term(depth) for _ in range(count))

    parts, size = [], 0
    while size < max(1, n):
        t = term(6)
        if rng.random() < 0.1:
            t = " " + t + " "
        parts.append(t)
        size += len(t) + 1
    return "+".join(parts).encode()
