"""Build the gfx950 HIP library in-tree (``alpacc_amd/libalpacc_hip.so``) and, for each bundled
grammar, the stand-alone executable ``alpacc_amd/bin/alpacc_<grammar>`` (alpacc_main.c over the
C ABI, tables from the header alpacc_amd.gen.emit_c writes into ``alpacc_amd/bin/``)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "alpacc_hip.hip")
OUT = os.path.join(HERE, "libalpacc_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         # keep the compiler from emitting packed-FP32 instructions
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


GRAMMARS = ("json", "arithmetic", "sexp")
BIN = os.path.join(HERE, "bin")
CC = os.environ.get("CC", "gcc")


def build_cli(grammars=GRAMMARS, verbose: bool = False) -> list:
    """Emit each grammar's table header and link its executable against the HIP library."""
    import sys
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    from alpacc_amd.gen.emit_c import emit_header
    from alpacc_amd.gen.tables import generate
    os.makedirs(BIN, exist_ok=True)
    outs = []
    for g in grammars:
        header = os.path.join(BIN, f"{g}_tables.h")
        text = emit_header(generate(os.path.join(HERE, "grammars", f"{g}.alp")), g)
        with open(header, "w") as f:
            f.write(text)
        exe = os.path.join(BIN, f"alpacc_{g}")
        cmd = [CC, "-O2", "-std=c11", "-Wall", "-I", os.path.join(ROOT, "include"), "-I", BIN,
               f'-DALPACC_TABLES_HEADER="{g}_tables.h"', f"-DALPACC_PREFIX={g}",
               os.path.join(HERE, "csrc", "alpacc_main.c"), "-o", exe + ".tmp",
               "-L", HERE, "-lalpacc_hip", "-Wl,-rpath,$ORIGIN/.."]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(exe + ".tmp", exe)
        outs.append(exe)
    return outs


def build(verbose: bool = False) -> str:
    tmp = OUT + ".tmp"
    cmd = [HIPCC, *FLAGS, "-I", os.path.join(ROOT, "include"), SRC, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
    print("\n".join(build_cli(verbose=True)))
