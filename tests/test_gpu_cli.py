"""The per-grammar executables (alpacc_main.c over the C ABI, tables from the emitted header) print
the same tokens and trees as the oracle and exit 0 on acceptance, -1 on rejection, like the
executable of the reference's CUDA backend (cuda/lexer.cu:421-425, 532)."""
import os
import subprocess

import pytest

from conftest import GRAMMARS, ROOT

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "alpacc_amd", "bin")


def _run(g, args, data):
    exe = os.path.join(BIN, f"alpacc_{g}")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (alpacc_amd.build.build_cli)")
    print(f"cli {g} {args} n={len(data)}", flush=True)
    r = subprocess.run([exe, *args], input=data, capture_output=True, timeout=60)
    assert r.returncode in (0, 255), r.stderr.decode()
    return r.returncode, r.stdout.decode()


@pytest.fixture(scope="module")
def bundles():
    from alpacc_amd.gen.tables import generate
    return {g: generate(os.path.join(GRAMMARS, f"{g}.alp")) for g in ("json", "arithmetic", "sexp")}


def _want_lex(b, data):
    from oracle import sequential
    toks = sequential.lex(b, data)
    return None if toks is None else "".join(f"{s} {e} {k}\n" for k, s, e in toks)


def _want_parse(b, data, names=False):
    from oracle import sequential
    tree = sequential.parse(b, data)
    if tree is None:
        return None
    out = []
    for p, term, ident, s, e in tree:
        if term:
            out.append(f"{p} T {b.terminal_names[ident] if names else ident} {s} {e}\n")
        else:
            out.append(f"{p} P {b.parser.production_names[ident] if names else ident}\n")
    return "".join(out)


def test_readme_example(bundles):
    b = bundles["arithmetic"]
    data = b"12+(3*4)"
    rc, out = _run("arithmetic", ["lex"], data)
    assert rc == 0 and out == _want_lex(b, data)
    rc, out = _run("arithmetic", ["parse", "--names"], data)
    assert rc == 0 and out == _want_parse(b, data, names=True)
    assert out.splitlines()[0] == "0 P E0_0"


@pytest.mark.parametrize("g", ["json", "arithmetic", "sexp"])
def test_readme_form(bundles, g):
    """--readme prints what format_lex / format_parse print for the oracle's result (README.md:84-106
    for the arithmetic example, whose block is a fixture), "#none" when rejected."""
    import json
    from alpacc_amd import datagen
    from alpacc_amd.runtime import format_lex, format_parse
    from oracle import sequential
    b = bundles[g]
    tn, pn = b.terminal_names, b.parser.production_names
    gen = {"json": datagen.json_doc, "arithmetic": datagen.arith_doc, "sexp": datagen.sexp_doc}[g]
    docs = [gen(3000, seed=11).tobytes(), b"((" if g != "arithmetic" else b"1+", b""]
    if g == "arithmetic":
        kat = json.load(open(os.path.join(ROOT, "tests", "golden",
                                          "haskell_vectors.json")))["readme_tree"]
        rc, out = _run(g, ["parse", "--readme"], kat["input"].encode())
        assert rc == 0 and " ".join(out.split()) == kat["text"]
        docs.append(kat["input"].encode())
    for d in docs:
        toks = sequential.lex(b, d)
        want = format_lex(None if toks is None else [(tn[k], (s, e)) for k, s, e in toks])
        rc, out = _run(g, ["lex", "--readme"], d)
        assert (rc == 0) == (toks is not None) and out == want + "\n", (g, len(d))
        tree = sequential.parse(b, d)
        want = format_parse(None if tree is None else [
            (p, ("terminal", tn[i], (s, e)) if t else ("production", pn[i]))
            for p, t, i, s, e in tree])
        rc, out = _run(g, ["parse", "--readme"], d)
        assert (rc == 0) == (tree is not None) and out == want + "\n", (g, len(d))


@pytest.mark.parametrize("g", ["json", "arithmetic", "sexp"])
def test_documents_and_rejections(bundles, g):
    from alpacc_amd import datagen
    b = bundles[g]
    gen = {"json": datagen.json_doc, "arithmetic": datagen.arith_doc, "sexp": datagen.sexp_doc}[g]
    doc = gen(100_000, seed=77)
    for d in (doc.tobytes(), datagen.flip_bytes(doc, seed=5, block=20_000).tobytes(), b"", b"(("):
        want = _want_lex(b, d)
        rc, out = _run(g, ["lex"], d)
        assert (rc == 0) == (want is not None), (g, len(d))
        if want is not None:
            assert out == want
        want = _want_parse(b, d)
        rc, out = _run(g, ["parse"], d)
        assert (rc == 0) == (want is not None), (g, len(d))
        if want is not None:
            assert out == want


def test_test_mode(bundles):
    """`alpacc_<grammar> test` is the artifact's test entry: wire format in, wire format out."""
    from oracle import sequential, wire
    b = bundles["arithmetic"]
    inputs = [b"12+(3*4)", b"(", b""]
    exe = os.path.join(BIN, "alpacc_arithmetic")
    r = subprocess.run([exe, "test"], input=wire.encode_inputs(inputs), capture_output=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr.decode()
    assert r.stdout == wire.encode_tree_outputs([sequential.parse(b, s) for s in inputs])
