#!/usr/bin/env python3
"""JSON lex+parse throughput of the alpacc HIP backend on one MI355X.

Workload (default): grammars/json.alp, one seeded 256 MiB synthetic JSON document, full lex+parse
(bytes -> preorder tree with u32 parents, node words and spans), the document resident in HBM
before the timed region. A step is one ``alpacc_parse`` call over the document (it ends with the
verdict and node count back on the host).

Printed: one JSON line with the metric, the device-time roofline of the dominant kernel (HIP
events around each kernel on the stream it runs on), the whole pipeline against its algorithmic
bytes, a CPU baseline (the C/OpenMP restatement of futhark/lexer.fut + futhark/parser.fut timed on
this host on a bounded sample of the same workload) and, unless --no-batch, the batched form:
grammars/sexp.alp over 4096 seeded documents totalling 2 GiB (log-normal lengths, median 64 KiB,
1 KiB..16 MiB) in one alpacc_parse_batch call against 4096 alpacc_parse calls.
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
    ap.add_argument("--cpu-sample-mib", type=float, default=16.0,
                    help="size of the one-thread CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--verify", action="store_true", help="check the tree against the C oracle")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched sexp comparison")
    ap.add_argument("--batch-docs", type=int, default=4096)
    ap.add_argument("--batch-total-mib", type=int, default=2048)
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving comparison")
    ap.add_argument("--no-lex", action="store_true", help="skip the 512 MiB lexer-only run")
    ap.add_argument("--lex-mib", type=int, default=512, help="document size of the lexer-only run")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run device-to-device copies of this size (the reference kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    ap.add_argument("--serve-requests", type=int, default=300)
    ap.add_argument("--serve-rate", type=float, default=1200.0,
                    help="offered load, requests/s (1200/s = 5.7 GB/s: past what one request per "
                         "call sustains)")
    ap.add_argument("--serve-workers", type=int, default=4)
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _median_time(fn, reps: int = 5) -> float:
    """Median wall time of `reps` calls after one warm-up call."""
    fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(sorted(ts)[len(ts) // 2])


def cpu_baseline(bundle, workload, doc, seed, one_thread_mib):
    """The C/OpenMP restatement of lexer.fut + parser.fut (the oracle) timed on this host's cores:
    the benchmark document itself in mode U (unchunked) and mode F (the 16 MiB chunk loop of
    lexer.fut:144-172, as the generated entries run it), on all threads and, on a bounded sample,
    on one thread; plus the CPU-only arithmetic workload (1 MiB, seed 42). Every figure is the
    median of 5 runs after one warm-up, input already in memory, output allocation included."""
    from alpacc_amd import datagen
    from alpacc_amd.gen.tables import generate
    from oracle.c import coracle
    n = int(doc.shape[0])
    orc = coracle.Oracle(bundle)
    threads = orc.threads
    fn = orc.parse if workload == "parse" else orc.lex
    chunk = 16 << 20
    t_u = _median_time(lambda: fn(doc))
    t_f = _median_time(lambda: fn(doc, chunk=chunk))
    f_accepts = fn(doc, chunk=chunk) is not None
    small = datagen.json_doc(int(one_thread_mib * (1 << 20)), seed=seed)
    arith_b = generate(os.path.join(ROOT, "alpacc_amd", "grammars", "arithmetic.alp"))
    arith = datagen.arith_doc(1 << 20, seed=seed)
    aorc = coracle.Oracle(arith_b)
    t_a = _median_time(lambda: aorc.parse(arith))
    orc.set_threads(1)
    try:
        t_1 = _median_time(lambda: fn(small))
        t_a1 = _median_time(lambda: aorc.parse(arith))
    finally:
        orc.set_threads(threads)
    gbs = lambda nb, t: round(nb / t / 1e9, 4)  # noqa: E731
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return {"value": gbs(n, t_u), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"the benchmark document itself ({n >> 20} MiB JSON, seed {seed}), "
                      f"{workload} with the C/OpenMP restatement of lexer.fut+parser.fut (mode U, "
                      f"unchunked), {threads} threads, median of 5 runs after one warm-up",
            "cpu_model": _cpu_model(), "nproc": cores,
            "mode_F_GBps": gbs(n, t_f),
            # the chunk loop lexes the overlap byte twice; a rejection there skips the parse
            "mode_F_accepts": f_accepts,
            "one_thread": {"GBps": gbs(small.shape[0], t_1),
                           "sample": f"{one_thread_mib:g} MiB JSON document, seed {seed}, mode U"},
            "arithmetic_1MiB": {"all_threads_GBps": gbs(arith.shape[0], t_a),
                                "one_thread_GBps": gbs(arith.shape[0], t_a1),
                                "ms_all_threads": round(t_a * 1e3, 2),
                                "workload": "grammars/arithmetic.alp, 1 MiB, seed 42, lex+parse"}}


def lexer_only(args, lib):
    """The GPU lexer alone on one 512 MiB JSON document (the lexer-only bundle: u32 kind, start and
    end per token), as the plain method's lexer configuration states it."""
    import numpy as np
    from alpacc_amd import datagen
    from alpacc_amd.gen.tables import generate
    from alpacc_amd.runtime import Context, DeviceBuffer
    bundle = generate(os.path.join(ROOT, "alpacc_amd", "grammars", "json.alp"), mode="lexer")
    n = args.lex_mib << 20
    log(f"generating {args.lex_mib} MiB JSON for the lexer (seed {args.seed})")
    doc = datagen.json_doc(n, seed=args.seed)
    ctx = Context(bundle)
    d_in = DeviceBuffer(n, lib)
    d_in.upload(doc)
    del doc
    outs = [DeviceBuffer(4 * n, lib) for _ in range(3)]

    def step():
        return ctx.lex_device(d_in.ptr, n, outs[0].ptr, outs[1].ptr, outs[2].ptr, n)

    for _ in range(2):
        rc, T = step()
    if rc != 1:
        raise SystemExit(f"lexer benchmark document rejected (rc={rc})")
    ctx.profile(True)
    lib.alpacc_stream_sync(None)
    t0 = time.perf_counter()
    steps = 5
    for _ in range(steps):
        rc, cnt = step()
        assert rc == 1 and cnt == T
    lib.alpacc_stream_sync(None)
    ms = (time.perf_counter() - t0) * 1e3 / steps
    prof = ctx.read_profile()
    ctx.profile(False)
    stage = {k: v[0] / v[1] for k, v in prof.items() if v[1]}
    emit_alg = n + 12 * T  # the document in, u32 kind, start and end per token out
    emit_tbs = emit_alg / (stage["lex_emit"] / 1e3) / 1e12
    for o in outs + [d_in]:
        o.free()
    ctx.close()
    return {"workload": f"json lex only, {args.lex_mib} MiB document (seed {args.seed}), "
                        "lexer-only bundle, u32 kind/start/end per token",
            "GBps": round(n / (ms / 1e3) / 1e9, 3), "ms_per_step": round(ms, 4), "steps": steps,
            "tokens": int(T), "tokens_per_s": round(T / (ms / 1e3), 1),
            "stage_ms": {k: round(v, 4) for k, v in stage.items()},
            "lex_emit_roofline": {"algorithmic_bytes": int(emit_alg), "achieved_TBps": round(emit_tbs, 3),
                                  "frac": round(emit_tbs * 1e3 / HBM_PEAK_GBS, 4)},
            "pipeline_roofline": {"algorithmic_bytes": int(emit_alg),
                                  "device_ms": round(sum(stage.values()), 4),
                                  "frac": round(emit_alg / (sum(stage.values()) / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}}


def batched(args, lib):
    """The batched form against one call per document, on the same resident documents."""
    import numpy as np
    from alpacc_amd import datagen
    from alpacc_amd.gen.tables import generate
    from alpacc_amd.runtime import Context, DeviceBuffer
    bundle = generate(os.path.join(ROOT, "alpacc_amd", "grammars", "sexp.alp"))
    total = args.batch_total_mib << 20
    sizes = datagen.sexp_batch_sizes(count=args.batch_docs, total=total)
    log(f"generating {args.batch_docs} sexp documents, {total >> 20} MiB")
    d_in = DeviceBuffer(total, lib)
    offsets = np.zeros(len(sizes), dtype=np.uint64)
    off = 0
    for i, sz in enumerate(sizes.tolist()):
        doc = datagen.sexp_doc(int(sz), seed=args.seed + 1000 + i)
        d_in.upload(doc, offset=off)
        offsets[i] = off
        off += int(sz)
    lengths = sizes.astype(np.uint64)
    # launch choices for this grammar, timed on the first 512 documents (about a quarter of the
    # bytes) by the autotuner
    from alpacc_amd import autotune
    log("autotuning the batched launch choices")
    sample = [datagen.sexp_doc(int(sz), seed=args.seed + 1000 + i)
              for i, sz in enumerate(sizes[:512].tolist())]
    tuned = autotune.tune(bundle, sample, reps=3, save_result=False)
    del sample
    ctx = Context(bundle)
    ctx.set_tuning(**tuned["params"])
    # size the outputs with one call (it reports the node count when the capacity is too small)
    rc, need, res = ctx.parse_batch_device(d_in.ptr, offsets, lengths, 0, 0, 0, 0)
    cap = int(need)
    outs = [DeviceBuffer(4 * cap, lib), DeviceBuffer(4 * cap, lib), DeviceBuffer(8 * cap, lib)]

    def batch_step():
        return ctx.parse_batch_device(d_in.ptr, offsets, lengths, outs[0].ptr, outs[1].ptr,
                                      outs[2].ptr, cap)

    def timed(steps):
        lib.alpacc_stream_sync(None)
        t0 = time.perf_counter()
        for _ in range(steps):
            rc_s, got_s, res_s = batch_step()
            if rc_s != 1 or got_s != got or not np.array_equal(res_s, res):
                raise SystemExit("batched results differ between calls")
        lib.alpacc_stream_sync(None)
        return (time.perf_counter() - t0) / steps

    # queued directly: every kernel launched from the host, one synchronisation at the end
    ctx.set_tuning(graph_mode=0)
    batch_step()
    rc, got, res = batch_step()
    if rc != 1:
        raise SystemExit(f"batched call failed (rc={rc})")
    accepted = int((res["status"] == 1).sum())
    t_direct = timed(args.batch_steps)
    # the same call replayed as one HIP graph launch (captured when the plan repeats)
    ctx.set_tuning(graph_mode=1)
    batch_step()
    batch_step()
    g0 = ctx.tuning()["graph_launches"]
    t_batch = timed(args.batch_steps)
    tune = ctx.tuning()
    graph_replays = tune["graph_launches"] - g0
    # stage times (HIP events around every stage: queued directly)
    ctx.profile(True)
    batch_step()
    prof = ctx.read_profile()
    ctx.profile(False)
    # plain: one alpacc_parse call per document, same documents and output buffers
    log("plain per-document calls")
    t0 = time.perf_counter()
    plain_nodes = 0
    plain_acc = 0
    for i in range(len(sizes)):
        rc_i, cnt = ctx.parse_device(d_in.ptr + int(offsets[i]), int(lengths[i]), outs[0].ptr,
                                     outs[1].ptr, outs[2].ptr, cap)
        plain_acc += rc_i == 1
        plain_nodes += cnt if rc_i == 1 else 0
    lib.alpacc_stream_sync(None)
    t_plain = time.perf_counter() - t0
    if plain_acc != accepted or plain_nodes != int(res["count"].sum()):
        raise SystemExit("batched and per-document results disagree")
    out = {"workload": f"sexp lex+parse, {len(sizes)} documents, {total >> 20} MiB, batched "
                       "(alpacc_parse_batch replayed as one HIP graph launch) vs one "
                       "alpacc_parse per document",
           "docs": len(sizes), "bytes": total, "accepted": accepted, "nodes": int(got),
           "batched_ms": round(t_batch * 1e3, 3), "plain_ms": round(t_plain * 1e3, 3),
           "batched_direct_ms": round(t_direct * 1e3, 3),
           "graph": {"replays_timed": int(graph_replays), "captures": int(tune["graph_captures"]),
                     "vs_direct": round(t_direct / t_batch, 3)},
           "autotuned": {"params": {k: tune[k] for k in autotune.DEFAULTS if k != "graph_mode"},
                         "sample": f"first 512 documents ({tuned['bytes'] >> 20} MiB)",
                         "sample_ms_default": tuned["default_ms"], "sample_ms_tuned": tuned["ms"]},
           "batched_docs_per_s": round(len(sizes) / t_batch, 1),
           "plain_docs_per_s": round(len(sizes) / t_plain, 1),
           "batched_GBps": round(total / t_batch / 1e9, 3),
           "vs_plain": round(t_plain / t_batch, 2),
           "batched_stage_ms": {k: round(v[0] / v[1], 4) for k, v in prof.items() if v[1]}}
    ctx.close()
    for o in outs:
        o.free()
    d_in.free()
    return out


def serving_bench(args):
    """Poisson stream of json lex+parse requests (log-uniform 256 B..64 MiB, 10 % invalid) served
    by continuous batching, and by one alpacc_parse call per request, at the same arrivals."""
    import numpy as np
    from alpacc_amd import datagen, serving
    from alpacc_amd.gen.tables import generate
    from alpacc_amd.runtime import Context
    bundle = generate(os.path.join(ROOT, "alpacc_amd", "grammars", "json.alp"))
    sizes = serving.request_sizes(args.serve_requests, seed=args.seed)
    rng = np.random.Generator(np.random.PCG64(args.seed + 7))
    log(f"generating {len(sizes)} requests, {int(sizes.sum()) >> 20} MiB")
    datas, invalid = [], 0
    for i, n in enumerate(sizes.tolist()):
        d = datagen.json_doc(max(int(n), 2), seed=args.seed + 20_000 + i)
        if rng.random() < 0.1:
            d = datagen.flip_bytes(d, seed=i, block=min(1 << 20, d.shape[0]))
            invalid += 1
        datas.append(d)
    arrivals = serving.poisson_arrivals(len(datas), args.serve_rate, seed=args.seed)
    # the plain loop first, its context warmed on the largest request
    ctx = Context(bundle)
    ctx.parse_int(max(datas, key=len))
    log("plain loop")
    plain_reqs = [serving.Request(i, np.ascontiguousarray(d)) for i, d in enumerate(datas)]
    serving.serve_plain(ctx, plain_reqs, arrivals)
    ctx.close()
    plain = serving.latency_stats(plain_reqs)
    # every worker's arenas, staging and output buffers take the size of the largest batch
    # (256 MiB) before the timed stream
    srv = serving.Server(bundle, workers=args.serve_workers,
                         warmup=datagen.json_doc(256 << 20, seed=args.seed + 9))
    serving.serve_poisson(srv, datas[:8], np.zeros(8))
    log(f"serving at {args.serve_rate:g} requests/s")
    reqs = serving.serve_poisson(srv, datas, arrivals)
    srv.close()
    batched = serving.latency_stats(reqs)
    batched["launches"] = len(srv.batch_sizes)
    batched["mean_batch"] = round(float(np.mean(srv.batch_sizes)), 2)
    same = all((a.status == 1) == (b.status == 1) and
               (a.tree is None or all(np.array_equal(x, y) for x, y in zip(a.tree, b.tree)))
               for a, b in zip(reqs, plain_reqs))
    return {"workload": f"json lex+parse requests, Poisson arrivals at {args.serve_rate:g}/s, "
                        f"sizes log-uniform 256 B..64 MiB, {invalid} of {len(datas)} invalid",
            "offered_GBps": round(float(sizes.sum()) / float(arrivals[-1]) / 1e9, 3),
            "continuous_batching": batched, "one_request_per_launch": plain,
            "results_identical": bool(same), "workers": args.serve_workers,
            "hbm_budget_GB": 200}


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

    if args.copy_probe_mib:
        lib = load_library()
        nb = args.copy_probe_mib << 20
        src, dst = DeviceBuffer(nb, lib), DeviceBuffer(nb, lib)
        for _ in range(4):
            lib.alpacc_copy_device(dst.ptr, src.ptr, nb, None)
        lib.alpacc_stream_sync(None)
        out = {"copy_probe_bytes": nb, "copies": 4}
        # practical rates for scale: device copy (read + write) and a fill (write), 20 calls each
        # between two host synchronisations
        import ctypes
        for name, fn, moved in (
                ("copy_TBps", lambda: lib.alpacc_copy_device(dst.ptr, src.ptr, nb, None), 2 * nb),
                ("fill_TBps", lambda: lib.alpacc_device_memset(ctypes.c_void_p(dst.ptr), 7, nb,
                                                                None), nb)):
            fn()
            lib.alpacc_stream_sync(None)
            t0 = time.perf_counter()
            for _ in range(20):
                fn()
            lib.alpacc_stream_sync(None)
            out[name] = round(moved * 20 / (time.perf_counter() - t0) / 1e12, 2)
        print(json.dumps(out), flush=True)
        return

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
    # Algorithmic bytes per launch of every single-kernel stage: what the kernel must read and
    # write at the least (per unit: byte n, token T, key position P = T + 2, bracket B, node N).
    #   lex_reduce    : n                       reads the document
    #   lex_emit      : n + 12 T                reads it again, writes (kind, start, end) u32; inside
    #                   n + 9 T                 alpacc_parse the kinds are bytes
    #   parse_expand  : 2 P + 8 T + 6 B + 16 N  pair entries in; token spans gathered; depth + symbol
    #                                           per bracket; s, node word and span per node out
    #   bracket_match : 6 B                     depth + symbol per bracket
    #   parse_parents : 8 N                     s in, parent out
    cnt = ctx.last_counts() if args.workload == "parse" else {}
    P, B, N_ = cnt.get("positions", 0), cnt.get("brackets", 0), cnt.get("productions", 0)
    alg = {"lex_reduce": n, "lex_emit": n + (9 if args.workload == "parse" else 12) * T}
    if args.workload == "parse":
        alg.update({"parse_expand": 2 * P + 8 * T + 6 * B + 16 * N_, "bracket_match": 6 * B,
                    "parse_parents": 8 * N_})
    dom = max((k for k in alg if stage_ms.get(k, 0) > 0), key=lambda k: stage_ms[k])
    dom_gbs = alg[dom] / (stage_ms[dom] / 1000.0) / 1e9 if alg[dom] else 0.0
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(dom_gbs, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(dom_gbs / HBM_PEAK_GBS, 4),
            "traffic": None, "avg_ms": round(stage_ms[dom], 4), "algorithmic_bytes": int(alg[dom])}
    # HBM bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes of this
    # workload (scaled by a plain copy of known size); only for the workload they were taken on
    tpath = os.path.join(ROOT, "profiles", "json_lex_parse_256MiB.traffic.json")
    kernel_of = {"lex_reduce": "lex_reduce", "lex_emit": "lex_emit", "parse_expand": "parse_expand2",
                 "bracket_match": "bracket_match", "parse_parents": "parse_parents"}
    if (args.workload == "parse" and args.mib == 256 and args.seed == 42
            and os.path.exists(tpath)):
        tk = json.load(open(tpath))["kernels"].get(kernel_of.get(dom, ""))
        if tk:
            roof["traffic"] = tk["traffic_bytes"]
            roof["traffic_source"] = ("profiles/json_lex_parse_256MiB.traffic.json: rocprofv3 "
                                      "FETCH_SIZE + WRITE_SIZE per dispatch, scaled by a 256 MiB copy")
    lex_ms = sum(stage_ms.get(k, 0.0) for k in ("lex_reduce", "lex_scan", "lex_emit"))
    dev_ms = lex_ms + stage_ms.get("parse", 0.0)
    pipe_bytes = n + 16 * count if args.workload == "parse" else n + 12 * T
    pipeline = {"algorithmic_bytes": int(pipe_bytes), "device_ms": round(dev_ms, 4),
                "achieved": round(pipe_bytes / (dev_ms / 1000.0) / 1e9, 1),
                "frac": round(pipe_bytes / (dev_ms / 1000.0) / 1e9 / HBM_PEAK_GBS, 4),
                "counts": cnt}
    extra = {"tokens": int(T), "pipeline_roofline": pipeline,
             "stage_ms": {k: round(v, 4) for k, v in stage_ms.items() if v},
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
            log("cpu baseline")
            cpu = cpu_baseline(bundle, args.workload, doc, args.seed, args.cpu_sample_mib)
        except Exception as exc:  # the baseline is reported, never required
            log(f"cpu baseline unavailable: {exc}")
    if not args.no_batch or not args.no_serve or not args.no_lex:
        for o in outs:
            o.free()
        d_in.free()
        ctx.close()
    if not args.no_lex and args.workload == "parse":
        extra["lexer_only"] = lexer_only(args, lib)
    if not args.no_batch:
        extra["batched"] = batched(args, lib)
    if not args.no_serve:
        extra["serving"] = serving_bench(args)
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
