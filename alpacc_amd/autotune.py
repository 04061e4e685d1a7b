"""Per-grammar autotuner of the backend's launch choices (alpacc_set_tuning).

The reference fixes its launch shape when it generates code: cuda/lexer.cu:1-48 derives the items
per thread from the shared memory per block and the state width, and
src/Alpacc/Generator/Cuda/Generator.hs:37-48 writes the constants into the generated source. Here
the kernels take their launch choices at run time, so they are timed instead:

* lexer_tables_in_lds: the lexer's tables in LDS where they fit, or all of them read from L2;
* lex_blocks_per_cu / keys_blocks_per_cu: persistent workgroups per CU of lex_reduce (its LDS holds
  the grammar's (endomorphism x byte class) table, whose size sets how many fit) and of parse_keys2
  (its LDS holds the pair index);
* search_block_threads / search_max_blocks: workgroup size and grid of the two searches;
* brackets_per_kib: the first size of the batched bracket scratch, from the brackets a sample batch
  needed (a batch that needs more than its estimate runs twice);
* graph_mode: whether batched calls that repeat their plan replay a HIP graph.

A tuning record is keyed by the grammar's table sizes (endomorphisms, byte classes, DFA states,
pair entries, key window q + k), so grammars of the same shape share it, and stored in one JSON file
(``ALPACC_TUNING_FILE``, else ``~/.cache/alpacc_hip/tuning.json``). The search is coordinate
descent over SPACE from the defaults: every candidate is the median of ``reps`` timed batched calls
on the caller's sample documents, after one untimed call.
"""
from __future__ import annotations

import json
import math
import os
import time

import numpy as np

DEFAULTS = {"graph_mode": 1, "lex_blocks_per_cu": 2, "keys_blocks_per_cu": 2,
            "search_max_blocks": 16384, "search_block_threads": 256, "brackets_per_kib": 256,
            "lexer_tables_in_lds": 1}

SPACE = (
    ("lexer_tables_in_lds", (1, 0)),
    ("lex_blocks_per_cu", (1, 2, 3, 4)),
    ("keys_blocks_per_cu", (1, 2, 4)),
    ("search_block_threads", (256, 512, 1024)),
    ("search_max_blocks", (4096, 8192, 16384, 65536)),
)


def table_key(bundle) -> str:
    """The table sizes a tuning record applies to."""
    lx, px = bundle.lexer, bundle.parser
    parts = []
    if lx is not None:
        parts += [f"e{lx.endo.num_endos}", f"C{lx.num_classes}", f"m{lx.num_states}"]
    if px is not None:
        parts += [f"E{int(px.pair_entry.shape[0])}", f"w{px.q + px.k}"]
    return "_".join(parts)


def default_path() -> str:
    return os.environ.get("ALPACC_TUNING_FILE") or os.path.join(
        os.path.expanduser("~"), ".cache", "alpacc_hip", "tuning.json")


def load(path: str | None = None) -> dict:
    path = path or default_path()
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save(records: dict, path: str | None = None) -> None:
    path = path or default_path()
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(records, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def apply(ctx, path: str | None = None) -> dict | None:
    """Set the stored launch choices of ctx's grammar, if any; returns them."""
    rec = load(path).get(table_key(ctx.bundle))
    if not rec:
        return None
    params = {k: int(v) for k, v in rec["params"].items() if k in DEFAULTS}
    ctx.set_tuning(**params)
    return params


class _Batch:
    """Sample documents resident on the device with output buffers sized for them."""

    def __init__(self, ctx, docs):
        from .runtime import DeviceBuffer
        self.ctx = ctx
        lengths = np.array([len(d) for d in docs], dtype=np.uint64)
        offsets = np.zeros(len(docs), dtype=np.uint64)
        if len(docs) > 1:
            offsets[1:] = np.cumsum(lengths[:-1])
        total = int(lengths.sum())
        self.d_in = DeviceBuffer(max(total, 1), ctx.lib)
        off = 0
        for d in docs:
            arr = np.frombuffer(bytes(d), dtype=np.uint8) if isinstance(d, (bytes, bytearray)) \
                else np.ascontiguousarray(d, dtype=np.uint8)
            if arr.shape[0]:
                self.d_in.upload(arr, offset=off)
            off += int(arr.shape[0])
        self.offsets, self.lengths, self.nbytes = offsets, lengths, total
        rc, need, _ = ctx.parse_batch_device(self.d_in.ptr, offsets, lengths, 0, 0, 0, 0)
        self.cap = int(need) + 16
        self.outs = [DeviceBuffer(4 * self.cap, ctx.lib), DeviceBuffer(4 * self.cap, ctx.lib),
                     DeviceBuffer(8 * self.cap, ctx.lib)]

    def run(self):
        rc, nodes, res = self.ctx.parse_batch_device(
            self.d_in.ptr, self.offsets, self.lengths, self.outs[0].ptr, self.outs[1].ptr,
            self.outs[2].ptr, self.cap)
        if rc != 1:
            raise RuntimeError(f"batched call failed during tuning (rc={rc})")
        return nodes, res

    def time(self, reps: int) -> float:
        self.run()
        self.run()  # a repeated plan: graph_mode 1 captures here
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            self.run()
            ts.append(time.perf_counter() - t0)
        return float(sorted(ts)[len(ts) // 2])

    def free(self):
        for b in self.outs + [self.d_in]:
            b.free()


def tune(bundle, docs, reps: int = 3, path: str | None = None, save_result: bool = True,
         log=None) -> dict:
    """Time the launch choices of ``bundle`` on a batch of sample documents (bytes or uint8
    arrays) and store the best under the grammar's table key. Returns the tuning record."""
    from .runtime import Context
    ctx = Context(bundle)
    batch = _Batch(ctx, docs)
    try:
        ctx.set_tuning(**DEFAULTS)
        nodes0, res0 = batch.run()
        # the bracket scratch estimate: what this batch needed, with a quarter more
        brackets = ctx.last_counts()["brackets"]
        kib = max(batch.nbytes / 1024.0, 1.0)
        params = dict(DEFAULTS)
        params["brackets_per_kib"] = int(min(1 << 20, max(1, math.ceil(1.25 * brackets / kib))))
        ctx.set_tuning(**params)
        base_ms = batch.time(reps) * 1e3
        best_ms = base_ms
        tried = []
        for name, values in SPACE:
            for v in values:
                if v == params[name]:
                    continue
                cand = dict(params, **{name: v})
                ctx.set_tuning(**cand)
                ms = batch.time(reps) * 1e3
                nodes, res = batch.run()
                if nodes != nodes0 or not np.array_equal(res, res0):
                    raise RuntimeError(f"results changed under {cand}")
                tried.append({name: v, "ms": round(ms, 4)})
                if log:
                    log(f"tune {name}={v}: {ms:.3f} ms (best {best_ms:.3f})")
                if ms < best_ms * 0.98:
                    best_ms, params = ms, cand
        # graph replay against queuing every kernel from the host
        ctx.set_tuning(**dict(params, graph_mode=0))
        direct_ms = batch.time(reps) * 1e3
        params["graph_mode"] = 1 if best_ms <= direct_ms else 0
        rec = {"key": table_key(bundle), "params": params, "ms": round(min(best_ms, direct_ms), 4),
               "default_ms": round(base_ms, 4), "direct_ms": round(direct_ms, 4),
               "docs": len(docs), "bytes": batch.nbytes, "tried": tried}
    finally:
        batch.free()
        ctx.close()
    if save_result:
        recs = load(path)
        recs[rec["key"]] = rec
        save(recs, path)
    return rec
