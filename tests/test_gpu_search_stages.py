"""GPU parity on documents built to reach every stage of the previous-smaller-or-equal search
(find_previous, futhark/parser.fut:92-108) that bracket matching and parents run on: matches inside
a 64-element chunk, in the previous chunk, in the window of the 64 chunks before that, and beyond
it (more than 64 * 64 elements back, through the min hierarchy). Every document is checked
against the C restatement of lexer.fut + parser.fut, and the CPU confirms beforehand that the
document's brackets and parents really reach the far stage."""
import os

import numpy as np
import pytest

from conftest import GRAMMARS

pytestmark = pytest.mark.gpu

FAR = 64 * 64 + 64  # further back than the previous chunk and the 64-chunk window


def _pse(x):
    """Previous smaller-or-equal index of every element (-1: none), by a stack."""
    out = np.full(len(x), -1, dtype=np.int64)
    st = []
    for i, v in enumerate(x.tolist()):
        while st and x[st[-1]] > v:
            st.pop()
        out[i] = st[-1] if st else -1
        st.append(i)
    return out


def _reach(bundle, orc, doc):
    """Longest bracket-match and parent distances of a document (element counts)."""
    from oracle.parser_fut import ParserFut
    kinds, _, _ = orc.lex(doc)
    pf = ParserFut(bundle.parser)
    spans = pf.lookup(pf.keys(np.asarray(kinds, dtype=np.int64)))
    br = pf.segmented_copy(pf.pt.flat_stacks, spans[:, 0], spans[:, 1])
    left = (br & pf.top) != 0
    depth = np.cumsum(np.where(left, 1, -1)) - left
    j = _pse(depth)
    right = np.nonzero(~left)[0]
    bdist = int((right - j[right]).max()) if len(right) else 0
    par, _, _ = orc.parse(doc)
    par = np.asarray(par, dtype=np.int64)
    pdist = int((np.arange(len(par)) - par)[1:].max()) if len(par) > 1 else 0
    return bdist, pdist


def _tree(r):
    if r is None:
        return None
    par, node, span = r
    return (np.asarray(par).tolist(), np.asarray(node).tolist(),
            np.asarray(span).reshape(-1).tolist())


def _docs():
    wide = "[" + ",".join(["1"] * 12000) + "]"
    deep = "[" * 2500 + "1" + "]" * 2500
    return {
        "json": [wide, deep, '{"a":' + wide + ',"b":' + deep + ',"c":[1,2]}',
                 "[" + ",".join(["[" + ",".join(["2"] * 70) + "]"] * 120) + "]"],
        "arithmetic": ["(" * 2500 + "1" + ")" * 2500, "+".join(["1"] * 12000),
                       "(" + "+".join(["(" + "*".join(["3"] * 90) + ")"] * 100) + ")*2"],
        "sexp": ["(" + " ".join(["a"] * 12000) + ")", "(" * 2500 + "a" + ")" * 2500],
    }


@pytest.mark.parametrize("g", ["json", "arithmetic", "sexp"])
def test_search_stages(g):
    from alpacc_amd.gen.tables import generate
    from alpacc_amd.runtime import Context
    from oracle.c import coracle
    bundle = generate(os.path.join(GRAMMARS, f"{g}.alp"))
    orc = coracle.Oracle(bundle)
    ctx = Context(bundle)
    try:
        far_b = far_p = False
        for i, text in enumerate(_docs()[g]):
            doc = text.encode()
            bdist, pdist = _reach(bundle, orc, doc)
            far_b |= bdist > FAR
            far_p |= pdist > FAR
            print(f"{g} doc {i}: {len(doc)} bytes, longest match {bdist}, longest parent {pdist}",
                  flush=True)
            want = orc.parse(doc)
            assert want is not None, (g, i)
            got = ctx.parse_int(doc)
            assert _tree(got) == _tree(want), (g, i)
            # one byte changed: the verdicts must still agree
            bad = bytearray(doc)
            bad[len(bad) // 2] = ord(")") if g != "json" else ord("]")
            w2, g2 = orc.parse(bytes(bad)), ctx.parse_int(bytes(bad))
            assert _tree(g2) == _tree(w2), (g, i, "flipped")
        assert far_b and far_p, (g, far_b, far_p)
    finally:
        ctx.close()
