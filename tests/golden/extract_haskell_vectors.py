"""Write the golden vectors held by the reference's Haskell test-suite as a JSON fixture.

Reads (text only, nothing is executed):
  * haskell-tests/RegularExpression.hs:37-123 - regexes with their valid / invalid strings;
  * haskell-tests/LLP.hs:83-97  - the PSLS table (admissible pair -> pushed symbols) of the paper
                                  grammar at q = k = 1;
  * haskell-tests/LLP.hs:495-499 - the LLP(q, k) parse of "a+[a+a]";
  * haskell-tests/LL.hs:75-131   - FIRST(1)/FOLLOW(1) sets and the LL(k) parse of "aabbbcc$";
  * README.md:81-106             - the worked lex+parse example (parents and nodes).
and writes tests/golden/haskell_vectors.json. Run it where /root/reference is present.
"""
from __future__ import annotations

import ast
import json
import os
import re
import sys

REF = os.environ.get("ALPACC_REFERENCE", "/root/reference")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "haskell_vectors.json")


def hs_strings(s: str) -> list:
    return [ast.literal_eval('"' + m + '"') for m in re.findall(r'"((?:[^"\\]|\\.)*)"', s)]


def regex_cases() -> list:
    src = open(os.path.join(REF, "haskell-tests/RegularExpression.hs")).read()
    out = []
    for m in re.finditer(r'regularExpressionMatchCase\s+isMatch\s+("(?:[^"\\]|\\.)*")\s+(\[[^\]]*\])\s+(\[[^\]]*\])', src):
        regex = ast.literal_eval(m.group(1))
        out.append({"regex": regex, "valid": hs_strings(m.group(2)), "invalid": hs_strings(m.group(3))})
    return out


def sym(tok: str):
    tok = tok.strip()
    m = re.fullmatch(r'augT\'?\s+"([^"]*)"', tok)
    if m:
        return ["t", m.group(1)]
    m = re.fullmatch(r'augNT\'\s+"([^"]*)"', tok)
    if m:
        return ["nt", m.group(1)]
    if tok in ("LeftTurnstile", "leftTurnstile'"):
        return ["t", "-|"]
    if tok in ("RightTurnstile", "rightTurnstile'"):
        return ["t", "|-"]
    raise ValueError(tok)


def psls_table() -> list:
    src = open(os.path.join(REF, "haskell-tests/LLP.hs")).read()
    block = src[src.index("pslsTable ="):src.index("collection ::")]
    out = []
    for line in block.splitlines():
        line = line.strip()
        if not line.startswith("(") and not line.startswith("[ ("):
            continue
        line = line.lstrip("[ ").rstrip(",")
        m = re.match(r'\(\(\[(.*?)\],\s*\[(.*?)\]\),\s*Set\.fromList\s*\[\[(.*)\]\]\)$', line)
        if not m:
            continue
        out.append({"suffix": [sym(x) for x in m.group(1).split(",")],
                    "prefix": [sym(x) for x in m.group(2).split(",")],
                    "alpha": [sym(x) for x in m.group(3).split(",")]})
    return out


def llp_parse_vector() -> dict:
    src = open(os.path.join(REF, "haskell-tests/LLP.hs")).read()
    m = re.search(r'input = map List\.singleton "([^"]*)"\s+result = parser input\s+expected = Just (\[[^\]]*\])', src)
    return {"input": m.group(1), "expected": json.loads(m.group(2)), "q": [1, 2, 3], "k": [1, 2, 3]}


def balanced(s: str, i: int) -> tuple[str, int]:
    """The bracketed text starting at s[i] == '[' and the index after it."""
    assert s[i] == "["
    depth = 0
    j = i
    in_str = False
    while True:
        c = s[j]
        if in_str:
            if c == "\\":
                j += 1
            elif c == '"':
                in_str = False
        elif c == '"':
            in_str = True
        elif c == "[":
            depth += 1
        elif c == "]":
            depth -= 1
            if depth == 0:
                return s[i:j + 1], j + 1
        j += 1


def ll_vectors() -> dict:
    src = open(os.path.join(REF, "haskell-tests/LL.hs")).read()

    def sets(name):
        blk = src[src.index(name):]
        blk = blk[blk.index("expected ="):]
        blk = blk[:blk.index("\n\n")]
        res = []
        pos = 0
        while True:
            k = blk.find("Set.fromList", pos)
            if k < 0:
                break
            txt, pos = balanced(blk, blk.index("[", k))
            res.append(sorted(json.loads(txt)))
        return res

    m = re.search(r'input = List\.singleton <\$> "([^"]*)"\s+result = llkParse\' \(input, \[Nonterminal \$ start extendedGrammar\], \[\]\)\s+expected = Just \(\[\], \[\], (\[[^\]]*\])\)', src)
    return {"first_small": sets("firstSmallTestCase"), "follow_small": sets("followSmallTestCase"),
            "follow_large": sets("followLargeTestCase"),
            "ll_parse": {"input": m.group(1), "expected": json.loads(m.group(2)), "k": list(range(1, 11))}}


def readme_tree() -> dict:
    src = open(os.path.join(REF, "README.md")).read()
    blk = src[src.index("#some [(0, #production (#E0_0))"):]
    blk = blk[:blk.index("```")]
    nodes = []
    for m in re.finditer(r'\((\d+), #(production|terminal) \(#(\w+)\)(?: \((\d+), (\d+)\))?\)', blk):
        nodes.append({"parent": int(m.group(1)), "kind": m.group(2), "name": m.group(3),
                      "span": [int(m.group(4)), int(m.group(5))] if m.group(4) else None})
    # the printed block itself, whitespace collapsed: the README's form of a parse result
    text = " ".join(blk.split())
    return {"input": "12+(3*4)", "grammar": "arithmetic", "nodes": nodes, "text": text}


def main():
    data = {
        "source": "reference haskell-tests/*.hs and README.md (golden values only)",
        "regex": regex_cases(),
        "psls_paper_q1_k1": psls_table(),
        "llp_parse": llp_parse_vector(),
        "ll": ll_vectors(),
        "readme_tree": readme_tree(),
    }
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
    print(f"wrote {OUT}: {len(data['regex'])} regex cases, {len(data['psls_paper_q1_k1'])} psls "
          f"entries, {len(data['readme_tree']['nodes'])} README nodes", file=sys.stderr)


if __name__ == "__main__":
    main()
