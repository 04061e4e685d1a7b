"""The C header written by alpacc_amd.gen.emit_c holds exactly the bundle's tables: a C program that
includes it dumps every array of the alpacc_tables value, and the bytes must equal the numpy arrays
the Python host passes through ctypes (no GPU)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GRAMMARS, ROOT

DUMP = r"""
#include <stdio.h>
#include TABLES_HEADER
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
static void put(const void *p, size_t n) { fwrite(&n, sizeof n, 1, stdout); fwrite(p, 1, n, stdout); }
int main(void) {
  const alpacc_tables *t = &CAT(PREFIX, _tables);
  unsigned sc[] = {t->num_endos, t->num_classes, t->num_states, t->init_state,
                   (unsigned)t->ignore_terminal, t->dead_terminal, t->q, t->k, t->num_terminals,
                   t->start_terminal, t->end_terminal, t->empty_terminal, t->num_productions};
  put(sc, sizeof sc);
  put(t->byte_class, 256);
  put(t->endo_step, 2ull * t->num_endos * t->num_classes);
  put(t->endo_comp, 2ull * t->num_endos * t->num_endos);
  put(t->endo_apply, 1ull * t->num_endos * t->num_states);
  put(t->dfa_step, 2ull * t->num_states * t->num_classes);
  put(t->state_accept, t->num_states);
  put(t->arities, 4ull * t->num_productions);
  put(t->prod_to_terminal, 4ull * t->num_productions);
  put(t->pair_index, 4ull * t->pair_index_len);
  put(t->pair_entry, 16ull * t->num_pair_entries);
  put(t->stacks, 2ull * t->stacks_len);
  put(t->productions, 2ull * t->productions_len);
  printf("%s|%s", CAT(PREFIX, _terminal_names)[0], CAT(PREFIX, _production_names)[0]);
  return 0;
}
"""


def _chunks(raw: bytes, count: int):
    out, pos = [], 0
    for _ in range(count):
        n = int.from_bytes(raw[pos:pos + 8], "little")
        out.append(raw[pos + 8:pos + 8 + n])
        pos += 8 + n
    return out, raw[pos:]


@pytest.mark.parametrize("g", ["arithmetic", "sexp", "json"])
def test_emitted_header_matches_bundle(tmp_path, g):
    from alpacc_amd.gen.emit_c import emit_header
    from alpacc_amd.gen.tables import generate
    b = generate(os.path.join(GRAMMARS, f"{g}.alp"))
    (tmp_path / "t.h").write_text(emit_header(b, g))
    (tmp_path / "dump.c").write_text(DUMP)
    exe = tmp_path / "dump"
    subprocess.run(["gcc", "-O0", "-I", os.path.join(ROOT, "include"), "-I", str(tmp_path),
                    '-DTABLES_HEADER="t.h"', f"-DPREFIX={g}", str(tmp_path / "dump.c"), "-o",
                    str(exe)], check=True)
    raw = subprocess.run([str(exe)], check=True, capture_output=True).stdout
    parts, tail = _chunks(raw, 13)
    lx, px = b.lexer, b.parser
    sc = np.frombuffer(parts[0], dtype=np.uint32)
    assert sc.tolist() == [lx.endo.num_endos, lx.num_classes, lx.num_states, lx.init_state,
                           lx.ignore & 0xFFFFFFFF, lx.dead_terminal, px.q, px.k, px.num_terminals,
                           px.start_terminal, px.end_terminal, px.empty_terminal,
                           len(px.production_names)]
    want = [(lx.byte_class, np.uint8), (lx.endo_step, np.uint16), (lx.endo_comp, np.uint16),
            (lx.endo_apply, np.uint8), (lx.dfa_step, np.uint16), (lx.state_info, np.uint8),
            (px.arities, np.int32), (px.production_to_terminal, np.int32),
            (px.pair_index, np.uint32), (px.pair_entry, np.uint32), (px.stacks, np.uint16),
            (px.productions, np.uint16)]
    assert len(parts) == 1 + len(want)
    for got, (arr, dt) in zip(parts[1:], want):
        assert got == np.ascontiguousarray(arr, dtype=dt).tobytes()
    assert tail.decode() == f"{b.terminal_names[0]}|{px.production_names[0]}"


def test_c_string_escapes():
    from alpacc_amd.gen.emit_c import _c_string
    assert _c_string('a"b\\c') == '"a\\"b\\\\c"'
    assert _c_string("\n") == '"\\012"'
