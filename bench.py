#!/usr/bin/env python3
"""JSON lex+parse throughput of the alpacc HIP backend on one MI355X.

Workload (default): grammars/json.alp, one seeded 256 MiB synthetic JSON document, full lex+parse
(bytes -> preorder tree with u32 parents, node words and spans), the document resident in HBM
before the timed region. A step is one ``alpacc_parse`` call over the document (it ends with the
verdict and node count back on the host).

Printed: one JSON line with the metric, the device-time roofline of the dominant kernel and a CPU
baseline (the C/OpenMP restatement of futhark/lexer.fut + futhark/parser.fut timed on this host
on a bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mib", type=int, default=256, help="document size in MiB")
    ap.add_argument("--workload", choices=["parse", "lex"], default="parse")
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--cpu-sample-mib", type=float, default=16.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--verify", action="store_true", help="check the tree against the C oracle")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(bundle, workload, mib, seed):
    """Time the C/OpenMP oracle on a bounded sample generated like the benchmark document."""
    from alpacc_amd import datagen
    from oracle.c import coracle
    n = int(mib * (1 << 20))
    doc = datagen.json_doc(n, seed=seed + 1)
    orc = coracle.Oracle(bundle)
    threads = os.cpu_count() or 1
    fn = orc.parse if workload == "parse" else orc.lex
    fn(doc)  # warm-up
    reps = 0
    t0 = time.perf_counter()
    while True:
        fn(doc)
        reps += 1
        if time.perf_counter() - t0 > 10.0 or reps >= 20:
            break
    dt = (time.perf_counter() - t0) / reps
    return {"value": round(n / dt / 1e9, 4), "unit": "GB/s", "cores": orc.threads or threads,
            "kind": "port",
            "sample": f"{mib:g} MiB seeded JSON document (seed {seed + 1}), {workload} with the "
                      f"C/OpenMP restatement of lexer.fut+parser.fut, mean of {reps} runs"}


def main():
    args = parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)
    import numpy as np
    from alpacc_amd import datagen
    from alpacc_amd.gen.tables import generate
    from alpacc_amd.runtime import Context, DeviceBuffer, load_library

    mode = "both" if args.workload == "parse" else "lexer"
    bundle = generate(os.path.join(ROOT, "alpacc_amd", "grammars", "json.alp"), mode=mode)
    n = args.mib << 20
    log(f"generating {args.mib} MiB JSON (seed {args.seed})")
    doc = datagen.json_doc(n, seed=args.seed)
    lib = load_library()
    ctx = Context(bundle)
    d_in = DeviceBuffer(n, lib)
    d_in.upload(doc)
    import ctypes
    if args.workload == "parse":
        cap = 2 * n
        outs = [DeviceBuffer(4 * cap, lib), DeviceBuffer(4 * cap, lib), DeviceBuffer(8 * cap, lib)]

        def step():
            rc, cnt = ctx.parse_device(d_in.ptr, n, outs[0].ptr, outs[1].ptr, outs[2].ptr, cap)
            return rc, cnt
    else:
        cap = n
        outs = [DeviceBuffer(4 * cap, lib), DeviceBuffer(4 * cap, lib), DeviceBuffer(4 * cap, lib)]

        def step():
            return ctx.lex_device(d_in.ptr, n, outs[0].ptr, outs[1].ptr, outs[2].ptr, cap)

    log("warmup")
    rc, count = step()
    if rc != 1:
        raise SystemExit(f"benchmark document rejected (rc={rc})")
    for _ in range(max(0, args.warmup - 1)):
        step()
    # token count of the document (for tokens/s and the lexer's algorithmic bytes)
    kb = DeviceBuffer(4 * n, lib)
    rc_t, T = ctx.lex_device(d_in.ptr, n, kb.ptr, kb.ptr, kb.ptr, n) if args.workload == "parse" \
        else (rc, count)
    kb.free()
    ctx.profile(True)
    log(f"timing {args.steps} steps")
    lib.alpacc_stream_sync(None)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rc, cnt = step()
        assert rc == 1 and cnt == count
    lib.alpacc_stream_sync(None)
    t1 = time.perf_counter()
    prof = ctx.read_profile()
    ms_step = (t1 - t0) * 1000.0 / args.steps
    value = n / (ms_step / 1000.0) / 1e9
    stage_ms = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in prof.items()}
    # per-stage algorithmic bytes (compulsory traffic only):
    #   lex_reduce: n (reads the document), lex_emit: n + 12 T (reads it again, writes the tokens),
    #   lex_scan: 0 (tile aggregates only), parse: 16 N (writes the tree; tokens are intermediates)
    alg = {"lex_reduce": n, "lex_scan": 0, "lex_emit": n + 12 * T}
    if args.workload == "parse":
        alg["parse"] = 16 * count
    dom = max((k for k in alg if stage_ms.get(k, 0) > 0), key=lambda k: stage_ms[k])
    dom_gbs = alg[dom] / (stage_ms[dom] / 1000.0) / 1e9 if alg[dom] else 0.0
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(dom_gbs, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(dom_gbs / HBM_PEAK_GBS, 4),
            "traffic": None, "avg_ms": round(stage_ms[dom], 4), "algorithmic_bytes": int(alg[dom])}
    lex_ms = sum(stage_ms.get(k, 0.0) for k in ("lex_reduce", "lex_scan", "lex_emit"))
    extra = {"tokens": int(T), "stage_ms": {k: round(v, 4) for k, v in stage_ms.items() if v},
             "lexer_pipeline_GBps": round((n + 12 * T) / (lex_ms / 1000.0) / 1e9, 1) if lex_ms else None,
             "tokens_per_s": round(T / (ms_step / 1000.0), 1)}
    if args.workload == "parse":
        N = count
        extra.update({"nodes": int(N),
                      "tree_bytes_frac_of_peak": round((n + 16 * N) / (ms_step / 1000.0) / 1e9
                                                       / HBM_PEAK_GBS, 4)})
    if args.verify:
        from oracle.c import coracle
        orc = coracle.Oracle(bundle)
        if args.workload == "parse":
            want = orc.parse(doc)
            got = (outs[0].download(np.uint32, count), outs[1].download(np.uint32, count),
                   outs[2].download(np.uint32, 2 * count))
            ok = want is not None and all(np.array_equal(a, b) for a, b in zip(got, want))
        else:
            want = orc.lex(doc)
            got = tuple(o.download(np.uint32, count) for o in outs)
            ok = want is not None and all(np.array_equal(a, b) for a, b in zip(got, want))
        extra["verified_against_oracle"] = bool(ok)
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(bundle, args.workload, args.cpu_sample_mib, args.seed)
        except Exception as exc:  # the baseline is reported, never required
            log(f"cpu baseline unavailable: {exc}")
    metric = ("json_lex_parse_input_GBps" if args.workload == "parse" else "json_lex_input_GBps")
    line = {
        "metric": metric, "value": round(value, 3), "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": f"json {'lex+parse' if args.workload == 'parse' else 'lex'} "
                               f"{args.mib} MiB document", "grammar": "grammars/json.alp",
                   "bytes": n, "seed": args.seed, "hbm_frac_of_input_rate": round(value / HBM_PEAK_GBS, 4)},
        "roofline": roof, "cpu_baseline": cpu, **extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
