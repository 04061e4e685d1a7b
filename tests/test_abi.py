"""The C-ABI library loads without a GPU and exports every entry point include/alpacc_hip.h declares;
the table struct layout matches the header (no GPU calls)."""
import ctypes
import os
import re

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "alpacc_hip.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(alpacc_[a-z_]+)\s*\(", src)))


def test_library_exports_header_symbols():
    from alpacc_amd.runtime import load_library, EXPORTS
    lib = load_library()
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert set(EXPORTS) <= set(names)


def test_table_struct_matches_header():
    from alpacc_amd.runtime import AlpaccTables
    src = open(HEADER).read()
    body = src[src.index("typedef struct alpacc_tables {"):src.index("} alpacc_tables;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl or decl.startswith("typedef"):
            decl = decl.split("{")[-1].strip()
        if not decl:
            continue
        names = decl.split()[-1] if "," not in decl else decl
        for part in re.split(r",", decl):
            part = part.strip().lstrip("*")
            m = re.search(r"\*?\s*([a-z_0-9]+)$", part)
            if m:
                fields.append(m.group(1))
    assert fields == [f[0] for f in AlpaccTables._fields_]


def test_fill_tables_without_gpu():
    from alpacc_amd.gen.tables import generate
    from alpacc_amd.runtime import fill_tables
    b = generate(os.path.join(ROOT, "alpacc_amd", "grammars", "json.alp"))
    t, keep = fill_tables(b)
    assert t.num_endos == b.lexer.endo.num_endos and t.pair_index_len == 15 ** 4
    assert t.q == 1 and t.k == 3 and t.start_terminal == 13 and t.end_terminal == 14
