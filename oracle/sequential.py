"""Independent sequential oracles (test infrastructure, not product code).

  * ``lex``   - ``tokenize`` (src/Alpacc/Lexer/DFA.hs:415-447) on the lexer DFA, dropping ignore
                tokens; this is what the reference's test generator writes as expected output
                (src/Alpacc/Test/LexerParser.hs:131-153).
  * ``parse`` - ``llpParse`` (src/Alpacc/LLP.hs:600-650) by gluing pair configurations, then
                ``llpParseTree``/``flattenTree`` (LLP.hs:652-745): a preorder tree whose root is its
                own parent, terminal productions replaced by terminal nodes carrying the spans.
"""
from __future__ import annotations

import sys

from alpacc_amd.gen.automata import tokenize
from alpacc_amd.gen.llp import llp_parse


def lex(bundle, data: bytes):
    lt = bundle.lexer
    inv = {v: k for k, v in lt.token_ids.items()}
    ignore_tok = inv.get(lt.ignore) if lt.ignore >= 0 else None
    toks = tokenize(lt.dfa, bytes(data), ignore_tok)
    if toks is None:
        return None
    return [(lt.token_ids[t], i, j) for t, (i, j) in toks]


def build_tree(g, derivation, spans):
    """``mkTree`` (LLP.hs:680-714) + ``flattenTree`` (LLP.hs:716-727), iteratively.

    Returns a list of (parent, is_terminal, id, start, end) or None when the derivation does not
    form a tree (the reference raises an internal error there).
    """
    prods = g.productions
    p2t = g.production_to_terminal
    out = []
    d_pos = 0
    m_pos = 0
    if not derivation:
        return None
    # explicit stack of (production node index, remaining children)
    stack = []

    def new_node(parent):
        nonlocal d_pos, m_pos
        if d_pos >= len(derivation):
            raise ValueError("derivation too short")
        d = derivation[d_pos]
        d_pos += 1
        idx = len(out)
        if p2t[d] is not None:
            if m_pos >= len(spans):
                raise ValueError("too few tokens")
            s, e = spans[m_pos]
            m_pos += 1
            out.append((parent, True, p2t[d], s, e))
            return None
        out.append((parent, False, d, 0, 0))
        return (idx, len(prods[d][1]))

    try:
        top = new_node(0)
        if top is not None:
            stack.append(list(top))
        while stack:
            node = stack[-1]
            if node[1] == 0:
                stack.pop()
                continue
            node[1] -= 1
            child = new_node(node[0])
            if child is not None:
                stack.append(list(child))
    except ValueError:
        return None
    return out


def parse(bundle, data: bytes):
    toks = lex(bundle, data)
    if toks is None:
        return None
    pt = bundle.parser
    deriv = llp_parse(pt.llp, [t for t, _, _ in toks], pt.start_terminal, pt.end_terminal)
    if deriv is None:
        return None
    return build_tree(pt.grammar, deriv, [(i, j) for _, i, j in toks])


sys.setrecursionlimit(max(sys.getrecursionlimit(), 10000))
