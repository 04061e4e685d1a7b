"""Oracle wire formats of futhark/test.fut and the Data.Binary instances of src/Alpacc/Test/*.hs.

All integers are big-endian u64 (``encode_u64``, test.fut:9-18); booleans are one byte.
  * inputs:            u64 count, then per test u64 len + len bytes (Test/Lexer.hs:95-108)
  * lexer outputs:     u64 count, then u8 ok [+ u64 T + T x (kind, start, end)] (test.fut:40-52)
  * lexer+parser:      u64 count, then u8 ok [+ u64 N + N x (u8 tag, parent, id, start, end)]
                       with tag 0 = production (spans 0) and 1 = terminal (test.fut:134-157)
  * parser only:       inputs u64 count, per test u64 m + m x u64 terminal (test.fut:74-93);
                       outputs u64 count, then u8 ok [+ u64 N + N x u64 production] (:68-72)
"""
from __future__ import annotations

import struct


def encode_inputs(inputs) -> bytes:
    out = [struct.pack(">Q", len(inputs))]
    for s in inputs:
        out.append(struct.pack(">Q", len(s)))
        out.append(bytes(s))
    return b"".join(out)


def decode_inputs(blob: bytes) -> list:
    (n,) = struct.unpack_from(">Q", blob, 0)
    pos = 8
    out = []
    for _ in range(n):
        (ln,) = struct.unpack_from(">Q", blob, pos)
        pos += 8
        out.append(blob[pos:pos + ln])
        pos += ln
    return out


def encode_lexer_outputs(results) -> bytes:
    out = [struct.pack(">Q", len(results))]
    for r in results:
        if r is None:
            out.append(b"\x00")
            continue
        out.append(b"\x01" + struct.pack(">Q", len(r)))
        for kind, s, e in r:
            out.append(struct.pack(">QQQ", kind, s, e))
    return b"".join(out)


def encode_tree_outputs(results) -> bytes:
    out = [struct.pack(">Q", len(results))]
    for r in results:
        if r is None:
            out.append(b"\x00")
            continue
        out.append(b"\x01" + struct.pack(">Q", len(r)))
        for parent, is_t, ident, s, e in r:
            out.append(struct.pack(">BQQQQ", 1 if is_t else 0, parent, ident,
                                   s if is_t else 0, e if is_t else 0))
    return b"".join(out)


def decode_tree_outputs(blob: bytes) -> list:
    (n,) = struct.unpack_from(">Q", blob, 0)
    pos = 8
    out = []
    for _ in range(n):
        ok = blob[pos]
        pos += 1
        if not ok:
            out.append(None)
            continue
        (m,) = struct.unpack_from(">Q", blob, pos)
        pos += 8
        nodes = []
        for _ in range(m):
            tag, parent, ident, s, e = struct.unpack_from(">BQQQQ", blob, pos)
            pos += 33
            nodes.append((parent, bool(tag), ident, s, e))
        out.append(nodes)
    return out


def encode_terminal_inputs(inputs) -> bytes:
    out = [struct.pack(">Q", len(inputs))]
    for ts in inputs:
        out.append(struct.pack(">Q", len(ts)))
        out.extend(struct.pack(">Q", int(t)) for t in ts)
    return b"".join(out)


def encode_production_outputs(results) -> bytes:
    out = [struct.pack(">Q", len(results))]
    for r in results:
        if r is None:
            out.append(b"\x00")
            continue
        out.append(b"\x01" + struct.pack(">Q", len(r)))
        out.extend(struct.pack(">Q", int(p)) for p in r)
    return b"".join(out)
