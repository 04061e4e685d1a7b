"""The bracket-mismatch corpus (bracket_corpus.py) on the CPU: its labels agree with the C
restatement of parser.fut, and it holds mismatches of every kind: in a first-level or a
second-level cancelled pair and among the remaining brackets, including one straddling two
256-position groups."""
import os
from collections import Counter

import numpy as np
import pytest

import bracket_corpus as bc
from conftest import GRAMMARS

# grammar, longest stream, labels that must occur (arithmetic's keys and depths catch every short
# fault: 37050 key misses and 200 depth faults among its 37448 streams up to length 5)
CASES = (("sexp", 7, ("accept", "adjacent", "second", "remaining")),
         ("two_brackets", 5, ("accept", "adjacent", "second", "remaining")),
         ("arithmetic", 4, ("accept", "key", "depth")))


def bundle_of(name):
    from alpacc_amd.gen.tables import generate
    if name == "two_brackets":
        return generate(bc.TWO_BRACKETS)
    return generate(os.path.join(GRAMMARS, f"{name}.alp"))


@pytest.mark.parametrize("name,max_len,labels", CASES)
def test_short_streams_labels(name, max_len, labels):
    from oracle.c import coracle
    b = bundle_of(name)
    orc = coracle.Oracle(b)
    seen = Counter()
    for kinds in bc.short_streams(b, max_len):
        lab = bc.classify(b, kinds)
        t = np.arange(kinds.shape[0], dtype=np.int64)
        assert (orc.parse_tokens(kinds, t, t + 1) is not None) == (lab == "accept"), (kinds, lab)
        seen[lab] += 1
    print(name, dict(seen))
    for lab in labels:
        assert seen[lab] > 0, lab


def test_group_boundary_docs():
    from oracle.c import coracle
    b = bundle_of("sexp")
    orc = coracle.Oracle(b)
    for k, want in ((251, "adjacent"), (252, "remaining"), (253, "adjacent"),
                    (1020, "remaining"), (1021, "adjacent")):
        d = bc.boundary_doc(k)
        assert orc.parse(d) is None
        assert bc.classify(b, orc.lex(d)[0]) == want, k
    assert orc.parse(bc.boundary_doc(5)[:-2]) is not None  # without the second expression
