"""Benchmark: tANS self-synchronising subsequence decoding on one MI355X.

Headline (`value`): the batched configuration -- 10,000 independent streams, uint8
sizes log-uniform in [4 KiB, 16 MiB], symbols s = min(255, floor(X)), X ~ Exp(lambda)
with lambda ~ LogUniform(0.005, 2.0) per stream, each with its own 4096-state table
normalised from its histogram -- decoded by one batched launch sequence
(mans_decode_batch).  Beside it: the same streams decoded by per-stream plain launches
(mans_decode once per stream, in order, one HIP stream), the single-stream at-scale
configuration (2^30 symbols for lambda in {0.01, 0.05, 0.2, 1.0}, 4096 states), the
stress configuration (2^32 symbols at lambda 0.002 and 3.0), the CPU-runnable
configuration (4 MiB, 1024 states, lambda 0.1: GPU beside the reference multicore
decoder at S = 4 and a tuned S), and the served configuration (continuous batching under
Poisson arrivals; p50/p99 latency).

Every decoded stream is checked bit-exact against its original symbols before timing.
Inputs are resident in HBM; timing is HIP events on the decode stream.  A step is one
decode of the whole batch.  `roofline` is for the dominant kernel of the batched
decode; `cpu_baseline` times the reference MulticoreDecoder (compiled from the
reference sources into oracle/_ref, else the oracle's restatement) on a bounded sample
on this host.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--streams N] [--no-single]
                    [--no-plain] [--no-cpu-baseline] [--calibrate-copy MIB]
                    [--serve-seconds S] [--serve-pool N] [--no-stress] [--no-c1]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "decoded GB/s, batched vs per-stream plain launch; p50/p99 latency; bit-exact"
HBM_PEAK_GBS = 8000.0
THREADS = min(os.cpu_count() or 1, 16)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_streams(specs, keep_symbols=True):
    """specs: list of (seed, lambda, n, L).  Host generation + encoding (product codec)."""
    from multians_amd import _lib
    from multians_amd.decoder import EncodedStream

    def one(sp):
        seed, lam, n, L = sp
        units, fb, fs, dec, sym = _lib.make_stream(seed, lam, n, L, keep_symbols)
        return EncodedStream(units, fb, fs, dec, L, n), sym

    with ThreadPoolExecutor(THREADS) as ex:
        return list(ex.map(one, specs))


def batch_specs(n_streams, seed=20261203):
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(np.log(4096), np.log(16 << 20), n_streams))).astype(np.int64)
    lams = np.exp(rng.uniform(np.log(0.005), np.log(2.0), n_streams))
    return [(seed * 100003 + i, float(lams[i]), int(sizes[i]), 4096) for i in range(n_streams)]


def verify_batch(b, items, torch):
    """D2H every stream's output and compare with its symbols (decode order = reversed)."""
    bad = []

    def chk(i):
        got = b.output(i).cpu().numpy()
        if not np.array_equal(got[::-1], items[i][1]):
            bad.append(i)

    with ThreadPoolExecutor(THREADS) as ex:
        list(ex.map(chk, range(len(items))))
    return bad


def time_events(torch, fn, reps):
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def bench_batched(args, torch):
    from multians_amd import _lib
    from multians_amd.decoder import DeviceBatch, workspace_bytes
    t0 = time.time()
    specs = batch_specs(args.streams)
    items = make_streams(specs)
    total_sym = sum(es.n_symbols for es, _ in items)
    total_units = sum(es.n_units for es, _ in items)
    log("batch: %d streams, %.2f GB symbols, %.2f GB compressed, generated in %.1fs" %
        (len(items), total_sym / 1e9, 4.0 * total_units / 1e9, time.time() - t0))
    b = DeviceBatch([es for es, _ in items], sync_points=not args.no_index)
    dec = (lambda: b.decode(one_pipeline=True)) if args.one_pipeline else b.decode
    dec()
    torch.cuda.synchronize()
    st = b.status()
    bad = verify_batch(b, items, torch)
    log("batched decode bit-exact: %s (status %d, mismatching streams %d)" % (not bad, st, len(bad)))
    if bad or st:
        raise SystemExit("batched decode mismatch: streams %s" % bad[:10])
    # kernel breakdown of one batched decode
    kt = _lib.KernelTimer()
    b.decode(kernel_events=kt.ptr)
    torch.cuda.synchronize()
    kms = kt.times_ms()
    kt.close()
    for _ in range(args.warmup):
        dec()
    ms_batch = time_events(torch, dec, args.steps)
    res = {"ms_batch": ms_batch, "ms_plain": None, "kernels_ms": kms,
           "total_symbols": total_sym, "total_units": total_units, "n_streams": len(items)}
    if not args.no_index:
        res["sync_points"] = bench_sync_points(args, torch, b, items)
    if args.no_plain:
        del b
        torch.cuda.empty_cache()
        return res, items
    # per-stream plain launches over the same device buffers (shared workspace)
    ws = torch.zeros(workspace_bytes(max(es.n_units for es, _ in items), 4096), dtype=torch.uint8,
                     device="cuda")
    plain = [b.plain_args(i, ws) for i in range(len(items))]
    L = _lib.lib()

    def run_plain():
        for a in plain:
            _lib.check(L.mans_decode(C.byref(a)))

    b.out.zero_()
    run_plain()
    torch.cuda.synchronize()
    bad = verify_batch(b, items, torch)
    log("per-stream plain launches bit-exact: %s" % (not bad))
    if bad:
        raise SystemExit("plain decode mismatch")
    res["ms_plain"] = time_events(torch, run_plain, 1)
    del b, ws
    torch.cuda.empty_cache()
    return res, items


def bench_sync_points(args, torch, b, items):
    """The batch from its recorded sync points at the tuned chunk lengths and at
    256-unit chunks, each also as windows decoded alone (mans_stream_range), and from
    trusted item records."""
    out = {"tuned_chunks": sync_point_leg(args, torch, b, items)}
    b.use_index(256)
    out["chunks_256"] = sync_point_leg(args, torch, b, items)
    b.use_index()
    out["trusted_items"] = trusted_leg(args, torch, b, items)
    return out


def trusted_leg(args, torch, b, items):
    """A decode at the tuned chunk lengths that also writes a record per 32 units
    (items_out), then the write pass alone from those records (MANS_BATCH_TRUST_
    SYNC_POINTS), and 64 Ki-symbol windows from them; all checked bit-exact."""
    from multians_amd import _lib
    from multians_amd.decoder import WindowDecoder
    total_sym = sum(es.n_symbols for es, _ in items)
    b.use_items()
    wdec = lambda: b.decode(index="items_out")
    wdec()
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        wdec()
    ms_w = time_events(torch, wdec, args.steps)
    b.out.zero_()
    b.decode(index="trusted", trust=True)
    torch.cuda.synchronize()
    bad = verify_batch(b, items, torch)
    log("trusted item records bit-exact: %s (status %d)" % (not bad, b.status()))
    if bad or b.status():
        raise SystemExit("trusted decode mismatch: streams %s" % bad[:10])
    kt = _lib.KernelTimer()
    b.decode(index="trusted", trust=True, kernel_events=kt.ptr)
    torch.cuda.synchronize()
    kms = kt.times_ms()
    kt.close()
    tdec = lambda: b.decode(index="trusted", trust=True)
    for _ in range(args.warmup):
        tdec()
    ms = time_events(torch, tdec, args.steps)
    n_rec = int(b.item_off[-1])
    out = {"writing_records_ms": round(ms_w, 4), "ms": round(ms, 4),
           "GBps": round(total_sym / ms / 1e6, 3),
           "kernels_ms": {k: round(v, 4) for k, v in kms.items() if v > 0.0},
           "records": n_rec, "record_bytes": 16 * n_rec}
    log("trusted item records", json.dumps(out))
    win, count = 65536, args.windows
    rng = np.random.default_rng(4)
    big = [i for i, (es, _) in enumerate(items) if es.n_symbols >= win]
    if big:
        wins = []
        for _ in range(count):
            i = int(rng.choice(big))
            a = int(rng.integers(0, items[i][0].n_symbols - win + 1))
            wins.append((i, a, a + win))
        planned = b.plan_windows(wins, items=True)
        wd = WindowDecoder()
        arena, spans = wd.decode(planned, wins, trust=True)
        torch.cuda.synchronize()
        host = arena.cpu().numpy()
        bad = [k for k, ((i, a, z), (o, ln)) in enumerate(zip(wins, spans))
               if not np.array_equal(host[o:o + ln], items[i][1][::-1][a:z])]
        log("trusted windows bit-exact: %s (status %d)" % (not bad, wd.status()))
        if bad or wd.status():
            raise SystemExit("trusted window mismatch: windows %s" % bad[:10])
        for _ in range(args.warmup):
            wd.decode(planned, wins, trust=True)
        wms = time_events(torch, lambda: wd.decode(planned, wins, trust=True), args.steps)
        out["windows"] = {"count": count, "window_symbols": win, "ms": round(wms, 4),
                          "windows_per_s": round(count / wms * 1e3, 1),
                          "window_GBps": round(count * win / wms / 1e6, 3)}
        log("trusted windows", json.dumps(out["windows"]))
    b.items = None
    return out


def sync_point_leg(args, torch, b, items):
    """The same batch again from its recorded sync points (index_in), and batches of
    windows decoded alone from them (mans_stream_range), both checked bit-exact."""
    from multians_amd import _lib
    from multians_amd.decoder import WindowDecoder
    total_sym = sum(es.n_symbols for es, _ in items)
    b.decode(index="out")
    torch.cuda.synchronize()
    n_rec = int(b.ioff[-1])
    b.out.zero_()
    b.decode(index="in")
    torch.cuda.synchronize()
    bad = verify_batch(b, items, torch)
    log("decode from sync points bit-exact: %s (status %d)" % (not bad, b.status()))
    if bad or b.status():
        raise SystemExit("decode from sync points mismatch: streams %s" % bad[:10])
    kt = _lib.KernelTimer()
    b.decode(index="in", kernel_events=kt.ptr)
    torch.cuda.synchronize()
    kms = kt.times_ms()
    kt.close()
    dec = (lambda: b.decode(index="in", one_pipeline=True)) if args.one_pipeline else \
        (lambda: b.decode(index="in"))
    for _ in range(args.warmup):
        dec()
    ms = time_events(torch, dec, args.steps)
    out = {"ms": round(ms, 4), "GBps": round(total_sym / ms / 1e6, 3),
           "kernels_ms": {k: round(v, 4) for k, v in kms.items()},
           "records": n_rec, "record_bytes": 16 * n_rec}
    log("from sync points", json.dumps(out))
    # windows of WIN symbols at random positions of the streams that hold one
    win, count = 65536, args.windows
    rng = np.random.default_rng(4)
    big = [i for i, (es, _) in enumerate(items) if es.n_symbols >= win]
    if not big:
        return out
    wins = []
    for _ in range(count):
        i = int(rng.choice(big))
        a = int(rng.integers(0, items[i][0].n_symbols - win + 1))
        wins.append((i, a, a + win))
    index_host = b.index.cpu().numpy().view(_lib.SYNC_POINT_DTYPE)
    t0 = time.time()
    planned = b.plan_windows(wins, index_host)
    plan_ms = (time.time() - t0) * 1e3
    wd = WindowDecoder()
    arena, spans = wd.decode(planned, wins)
    torch.cuda.synchronize()
    host = arena.cpu().numpy()
    bad = [k for k, ((i, a, z), (o, ln)) in enumerate(zip(wins, spans))
           if not np.array_equal(host[o:o + ln], items[i][1][::-1][a:z])]
    log("windows bit-exact: %s (status %d)" % (not bad, wd.status()))
    if bad or wd.status():
        raise SystemExit("window decode mismatch: windows %s" % bad[:10])
    decoded = int(planned[2][-1])
    for _ in range(args.warmup):
        wd.decode(planned, wins)
    wms = time_events(torch, lambda: wd.decode(planned, wins), args.steps)
    out["windows"] = {"count": count, "window_symbols": win, "ms": round(wms, 4),
                      "windows_per_s": round(count / wms * 1e3, 1),
                      "window_GBps": round(count * win / wms / 1e6, 3),
                      "decoded_symbols": decoded,
                      "decoded_GBps": round(decoded / wms / 1e6, 3),
                      "host_plan_ms": round(plan_ms, 2)}
    log("windows", json.dumps(out["windows"]))
    return out


def bench_encode(args, torch, items):
    """Encode every stream of the batch on the device (mans_encode_device, one call, each
    stream with its own table) from symbols resident in HBM; checked unit-exact against
    the host codec's streams."""
    from multians_amd import _lib
    n = len(items)
    soff = np.zeros(n + 1, np.int64)
    uoff = np.zeros(n + 1, np.int64)
    for i, (es, _) in enumerate(items):
        soff[i + 1] = soff[i] + ((es.n_symbols + 15) // 16) * 16
        uoff[i + 1] = uoff[i] + ((_lib.lib().mans_max_encoded_units(es.n_symbols) + 3) // 4) * 4
    sym = torch.empty(int(soff[-1]), dtype=torch.uint8, device="cuda")
    for i, (es, s) in enumerate(items):
        sym[int(soff[i]):int(soff[i]) + es.n_symbols].copy_(torch.from_numpy(s))
    toff = np.concatenate([[0], np.cumsum([es.L for es, _ in items])])
    tables = torch.from_numpy(np.concatenate([es.table for es, _ in items]).view(np.int32)).cuda()
    units = torch.empty(int(uoff[-1]), dtype=torch.int32, device="cuda")
    js = (_lib.EncodeJob * n)()
    sb, tb, ub = sym.data_ptr(), tables.data_ptr(), units.data_ptr()
    for i, (es, _) in enumerate(items):
        j = js[i]
        j.symbols, j.n = sb + int(soff[i]), es.n_symbols
        j.table, j.L = tb + 4 * int(toff[i]), es.L
        j.units, j.units_capacity = ub + 4 * int(uoff[i]), int(uoff[i + 1] - uoff[i])
    b = C.c_uint64()
    _lib.check(_lib.lib().mans_encode_workspace_bytes(js, n, C.byref(b)))
    ws = torch.empty(b.value, dtype=torch.uint8, device="cuda")
    L = _lib.lib()
    st = torch.cuda.current_stream().cuda_stream

    def enc():
        _lib.check(L.mans_encode_device(js, n, C.c_void_p(ws.data_ptr()), ws.numel(),
                                        C.c_void_p(st)))

    enc()
    res = (_lib.EncodeResult * n)()
    _lib.check(L.mans_encode_results(C.c_void_p(ws.data_ptr()), js, n, C.c_void_p(st), res))
    host = units.cpu().numpy().view(np.uint32)
    bad = [i for i, (es, _) in enumerate(items)
           if res[i].status or res[i].n_units != es.n_units or res[i].first_bit != es.first_bit
           or res[i].first_state != es.first_state
           or not np.array_equal(host[int(uoff[i]):int(uoff[i]) + es.n_units], es.units)]
    log("device encode unit-exact: %s" % (not bad))
    if bad:
        raise SystemExit("device encode mismatch: streams %s" % bad[:10])
    for _ in range(args.warmup):
        enc()
    ms = time_events(torch, enc, args.steps)
    total = sum(es.n_symbols for es, _ in items)
    out = {"streams": n, "symbols": total, "ms": round(ms, 3),
           "symbols_GBps": round(total / ms / 1e6, 3),
           "note": "one mans_encode_device call (plans copied in synchronously each call)"}
    log("encode", json.dumps(out))
    if args.no_encode_records:
        del sym, tables, units, ws
        torch.cuda.empty_cache()
        return out
    # the encoder also writes the decoder's 32-unit item records; the new streams then
    # decode by the write pass alone (MANS_BATCH_TRUST_SYNC_POINTS), checked bit-exact
    roff = np.zeros(n + 1, np.int64)
    for i in range(n):
        roff[i + 1] = roff[i] + _lib.sync_point_count(int(uoff[i + 1] - uoff[i]), 32)
    recs = torch.zeros(int(roff[-1]) * 16, dtype=torch.uint8, device="cuda")
    for i in range(n):
        js[i].items_out = recs.data_ptr() + 16 * int(roff[i])
    for _ in range(max(args.warmup, 1)):
        enc()
    ms_r = time_events(torch, enc, args.steps)
    _lib.check(L.mans_encode_results(C.c_void_p(ws.data_ptr()), js, n, C.c_void_p(st), res))
    del sym
    ooff = np.concatenate([[0], np.cumsum([es.n_symbols for es, _ in items])])
    outb = torch.empty(max(int(ooff[-1]), 1), dtype=torch.uint8, device="cuda")
    ds = (_lib.StreamDesc * n)()
    for i, (es, _) in enumerate(items):
        d = ds[i]
        d.units, d.n_units = ub + 4 * int(uoff[i]), res[i].n_units
        d.out, d.n_out = outb.data_ptr() + int(ooff[i]), es.n_symbols
        d.table, d.L = tb + 4 * int(toff[i]), es.L
        d.first_state, d.first_bit = res[i].first_state, res[i].first_bit
        d.chunk_units, d.index_in = 32, recs.data_ptr() + 16 * int(roff[i])
    b = C.c_uint64()
    _lib.check(L.mans_batch_workspace_bytes(ds, n, C.byref(b)))
    dws = torch.zeros(b.value, dtype=torch.uint8, device="cuda")
    o = _lib.BatchOpts()
    o.workspace, o.workspace_bytes, o.stream = dws.data_ptr(), dws.numel(), st
    o.flags = _lib.MANS_BATCH_TRUST_SYNC_POINTS

    def dec():
        _lib.check(L.mans_decode_batch(ds, n, C.byref(o), None))

    dec()
    torch.cuda.synchronize()

    def chk(i):
        got = outb[int(ooff[i]):int(ooff[i + 1])].cpu().numpy()
        return np.array_equal(got[::-1], items[i][1])

    with ThreadPoolExecutor(THREADS) as ex:
        ok = list(ex.map(chk, range(n)))
    bad = [i for i, v in enumerate(ok) if not v]
    log("decode from the encoder's item records bit-exact: %s" % (not bad))
    if bad:
        raise SystemExit("decode from encoder records mismatch: streams %s" % bad[:10])
    for _ in range(args.warmup):
        dec()
    ms_d = time_events(torch, dec, args.steps)
    out["with_item_records_ms"] = round(ms_r, 3)
    out["decode_from_encoder_records_ms"] = round(ms_d, 3)
    out["decode_from_encoder_records_GBps"] = round(total / ms_d / 1e6, 3)
    log("encode with item records, then decode from them", json.dumps(
        {k: out[k] for k in ("with_item_records_ms", "decode_from_encoder_records_ms",
                             "decode_from_encoder_records_GBps")}))
    del tables, units, ws, recs, outb, dws
    torch.cuda.empty_cache()
    return out


def _time_single(torch, gb, i, reps):
    """Decode stream i of a GeneratedBatch: bit-exact check on the device, then the median
    of `reps` HIP-event timings and a per-kernel breakdown."""
    from multians_amd import _lib
    from multians_amd.decoder import workspace_bytes
    n = int(gb.n_symbols[i])
    out = torch.empty(n, dtype=torch.uint8, device="cuda")
    ws = torch.zeros(workspace_bytes(int(gb.n_units[i]), int(gb.L[i])), dtype=torch.uint8,
                     device="cuda")
    gb.decode_one(i, out, ws)
    torch.cuda.synchronize()
    ok = gb.status(ws) == 0 and bool(torch.equal(out, gb.symbols(i).flip(0)))
    if not ok:
        raise SystemExit("single-stream decode mismatch (stream %s)" % (gb.specs[i],))
    for _ in range(2):
        gb.decode_one(i, out, ws)
    ms = sorted(time_events(torch, lambda: gb.decode_one(i, out, ws), 1) for _ in range(reps))
    kt = _lib.KernelTimer()
    gb.decode_one(i, out, ws, kernel_events=kt.ptr)
    torch.cuda.synchronize()
    kms = kt.times_ms()
    kt.close()
    del out, ws
    return ms[len(ms) // 2], kms


def bench_single(args, torch):
    """Single-stream at-scale configuration: 2^log2n symbols per lambda, L = 4096; streams
    made on the device (the host generator's symbols and the host codec's tables and units),
    median of 10 decodes each."""
    from multians_amd.decoder import GeneratedBatch
    lams = [0.01, 0.05, 0.2, 1.0]
    n = 1 << args.log2n
    per = []
    for i, lam in enumerate(lams):
        gb = GeneratedBatch([(1000 + i, lam, n, 4096)])
        ms, kms = _time_single(torch, gb, 0, 10)
        nu = int(gb.n_units[0])
        algo = 4.0 * nu + n
        per.append({"lambda": lam, "ms": round(ms, 4), "decoded_GBps": round(n / ms / 1e6, 2),
                    "ratio": round(n / (4.0 * nu), 3),
                    "hbm_frac": round(algo / ms / 1e6 / HBM_PEAK_GBS, 4), "bit_exact": True,
                    "kernels_ms": {k: round(v, 4) for k, v in kms.items()}})
        log("single", json.dumps(per[-1]))
        del gb
        torch.cuda.empty_cache()
    return per, None


def bench_plain_c1(args, torch):
    """The CPU-runnable configuration: 4 MiB of uint8 symbols, L = 1024, lambda = 0.1, one
    stream.  GPU decode (median of 20, HIP events) beside the reference MulticoreDecoder
    (oracle/_ref, else the oracle's restatement) on this host at S = 4 (the demo's) and at
    the best S of a short sweep, every output checked against the input and against each
    other (src/main.cc:46-164 prints the same pair per lambda)."""
    from multians_amd import _lib
    from multians_amd.decoder import DeviceStream, EncodedStream
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    n, L, lam = 4 << 20, 1024, 0.1
    units, fb, fs, dec, sym = _lib.make_stream(20261204, lam, n, L, True)
    es = EncodedStream(units, fb, fs, dec, L, n)
    ds = DeviceStream(es)
    out = ds.decode()
    torch.cuda.synchronize()
    gpu = out[:n].cpu().numpy()
    if ds.status() != 0 or not np.array_equal(gpu[::-1], sym):
        raise SystemExit("C1 GPU decode mismatch")
    for _ in range(3):
        ds.decode()
    gms = sorted(time_events(torch, ds.decode, 1) for _ in range(20))[10]
    padded = np.concatenate([units, np.zeros(1040, np.uint32)])
    use_ref = ol.ref_available()

    def multicore(S):
        o = np.zeros(n, np.uint8)
        t0 = time.perf_counter()
        if use_ref:
            us = ol.ref().ref_multicore(S, THREADS, padded, len(units), dec, L, fs, fb, o, n)
            dt = us / 1e3
        else:
            ol.oracle().mo_decode_multicore(S, THREADS, padded, len(units), dec, L, fs, fb, o, n)
            dt = (time.perf_counter() - t0) * 1e3
        return dt, o

    cpu = {}
    for S in (4, 16, 64, 256, 1024):
        ts = []
        for _ in range(5):
            dt, o = multicore(S)
            ts.append(dt)
        diff = np.nonzero(o != gpu)[0]
        # the reference drops trailing zero-bit symbols when n_units % S == 0 (the last
        # ones in decode order); any other difference is an error
        if len(diff) and diff.min() < n - 64:
            raise SystemExit("C1 multicore (S=%d) differs from the GPU decode" % S)
        cpu[S] = {"ms": round(sorted(ts)[2], 3), "bit_exact_vs_gpu": len(diff) == 0}
    best = min(cpu, key=lambda k: cpu[k]["ms"])
    out = {"symbols": n, "L": L, "lambda": lam, "units": len(units),
           "ratio": round(n / (4.0 * len(units)), 3), "gpu_ms": round(gms, 4),
           "gpu_decoded_GBps": round(n / gms / 1e6, 3), "bit_exact_vs_input": True,
           "cpu_kind": "reference" if use_ref else "port", "cpu_threads": THREADS,
           "cpu_S4_ms": cpu[4]["ms"], "cpu_best_S": best, "cpu_best_ms": cpu[best]["ms"],
           "cpu_sweep": {str(k): v for k, v in cpu.items()},
           "gpu_vs_cpu_S4": round(cpu[4]["ms"] / gms, 2),
           "gpu_vs_cpu_best": round(cpu[best]["ms"] / gms, 2)}
    log("C1", json.dumps(out))
    del ds
    return out


def bench_tables(args, torch):
    """Decoder tables built on the device for the first --tables streams of the batched
    configuration (one mans_build_tables_device call, symbols resident), against the host
    codec (threaded histogram, normalisation, spread) on the same streams; every device
    table is checked equal to the host one."""
    from multians_amd import _lib
    specs = batch_specs(args.streams)[:args.tables]
    syms = [_lib.gen_symbols(seed, lam, n, threads=THREADS) for (seed, lam, n, L) in specs]
    jobs = [(torch.from_numpy(s).cuda(), L, torch.zeros(L, dtype=torch.int32, device="cuda"), None)
            for s, (_, _, _, L) in zip(syms, specs)]
    status = _lib.build_tables_device(jobs)                   # warm-up + check below
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    ts = []
    for _ in range(3):
        a.record()
        _lib.build_tables_device(jobs)
        b.record()
        torch.cuda.synchronize()
        ts.append(a.elapsed_time(b))
    t0 = time.time()
    host = []
    for s, (_, _, _, L) in zip(syms, specs):
        c = _lib.normalize_counts(_lib.histogram(s, threads=THREADS), L)
        host.append(_lib.build_tables(c, L, want_enc=False)[1])
    host_ms = (time.time() - t0) * 1e3
    same = all(st == 0 and np.array_equal(j[2].cpu().numpy().view(np.uint32), h)
               for st, j, h in zip(status, jobs, host))
    if not same:
        raise SystemExit("device tables differ from the host codec's")
    out = {"tables": len(jobs), "symbols": int(sum(len(s) for s in syms)), "L": 4096,
           "device_ms": round(min(ts), 3), "host_ms": round(host_ms, 1),
           "host_threads": THREADS, "identical": True,
           "timing": "device: HIP events around one mans_build_tables_device call (job "
                     "upload included); host: wall clock of histogram + normalise + spread"}
    log("tables", json.dumps(out))
    return out


def bench_stress(args, torch):
    """Stress configuration: one stream of 2^32 uint8 symbols (4 GiB, 64-bit offsets) for
    lambda = 0.002 (symbol 255 holds 60% of the mass) and lambda = 3.0 (ratio > 7), 4096
    states, made on the device; bit-exact on the device, then timed (median of 5)."""
    from multians_amd.decoder import GeneratedBatch
    n = 1 << 32
    out = []
    for i, lam in enumerate([0.002, 3.0]):
        t0 = time.time()
        gb = GeneratedBatch([(4000 + i, lam, n, 4096)])
        torch.cuda.synchronize()
        nu = int(gb.n_units[0])
        log("stress lambda=%g: %d units (ratio %.2f), made on the device in %.1fs" %
            (lam, nu, n / (4.0 * nu), time.time() - t0))
        ms, kms = _time_single(torch, gb, 0, 5)
        out.append({"lambda": lam, "symbols": n, "ratio": round(n / (4.0 * nu), 3),
                    "ms": round(ms, 3), "decoded_GBps": round(n / ms / 1e6, 2),
                    "hbm_frac": round((4.0 * nu + n) / ms / 1e6 / HBM_PEAK_GBS, 4),
                    "bit_exact": True, "kernels_ms": {k: round(v, 4) for k, v in kms.items()}})
        log("stress", json.dumps(out[-1]))
        del gb
        torch.cuda.empty_cache()
    return out


def bench_interleaved(args, torch):
    """Two-state interleaved streams (the library's own format; the reference names
    multiple states / tables as future work): 2^log2n symbols for lambda = 0.2 and 1.0,
    L = 4096, one shared table, records every 256 units, host-encoded; bit-exact on the
    device, then the median of 10 decodes (HIP events).  `ratio` counts the records."""
    from multians_amd import _lib
    from multians_amd.decoder import DeviceInterleaved, encode_interleaved
    n = 1 << args.log2n
    out = []
    for i, lam in enumerate([0.2, 1.0]):
        t0 = time.time()
        sym = _lib.gen_symbols(5000 + i, lam, n)
        ie = encode_interleaved(sym, 4096, record_units=256)
        d = DeviceInterleaved(ie)
        log("interleaved lambda=%g: %d units, %d records, encoded on the host in %.1fs" %
            (lam, ie.n_units, len(ie.records), time.time() - t0))
        d.decode()
        torch.cuda.synchronize()
        ref = torch.from_numpy(sym).cuda()
        ok = d.status() == 0 and bool(torch.equal(d.out, ref.flip(0)))
        del ref
        if not ok:
            raise SystemExit("interleaved decode mismatch (lambda %g)" % lam)
        for _ in range(2):
            d.decode()
        ms = sorted(time_events(torch, d.decode, 1) for _ in range(10))[5]
        comp = 4.0 * ie.n_units + 16.0 * len(ie.records)
        out.append({"lambda": lam, "symbols": n, "record_units": 256,
                    "ratio": round(n / comp, 3), "ms": round(ms, 4),
                    "decoded_GBps": round(n / ms / 1e6, 2),
                    "hbm_frac": round((comp + n) / ms / 1e6 / HBM_PEAK_GBS, 4), "bit_exact": True})
        log("interleaved", json.dumps(out[-1]))
        del d, ie, sym
        torch.cuda.empty_cache()
    return out


def bench_served(args, torch):
    """Served configuration: open-loop Poisson arrivals at 70% of the measured served
    capacity on the request mix (L 1024/4096 50/50, uint8 sizes log-uniform 1 KiB-64 MiB,
    lambda ~ LogU(0.005, 2.0)), admitted under a 64 GB HBM budget, continuous batching.
    Each request brings its compressed stream and table from pinned host memory (a pool of
    client streams it is drawn from): copied to the device on admission, decoded, and
    released; the budget counts inputs, outputs and workspace, and latency runs from
    arrival to completion, copy included."""
    from multians_amd import serve
    t0 = time.time()
    specs = serve.request_mix(args.serve_pool)
    items = make_streams(specs)
    pool = serve.StreamPool([es for es, _ in items], [sym for _, sym in items], host_inputs=True)
    log("serve pool (host inputs): %d streams, %.2f GB symbols, %.2f GB compressed, built in "
        "%.1fs" % (pool.n, pool.n_sym.sum() / 1e9, pool.in_bytes.sum() / 1e9, time.time() - t0))
    srv = serve.Server(pool, 64e9, max_inflight=3)
    # capacity: every pool stream at t = 0 (one backlog), every result checked
    cap = srv.run(np.zeros(pool.n), np.arange(pool.n), check_every=1)
    if cap.mismatches or cap.failed:
        raise SystemExit("served capacity run: %d mismatching, %d failed" % (cap.mismatches, cap.failed))
    cap = srv.run(np.zeros(pool.n), np.arange(pool.n))
    capacity = cap.symbols / cap.wall_s
    in_rate = pool.in_bytes.sum() / cap.wall_s
    rate = 0.7 * capacity / float(pool.n_sym.mean())
    arr = serve.poisson_arrivals(rate, args.serve_seconds)
    ids = np.random.default_rng(11).integers(0, pool.n, len(arr))
    log("serve: capacity %.1f GB/s of symbols (%.1f GB/s of inputs over the host link), "
        "offered %.1f GB/s = %.0f requests/s for %gs" %
        (capacity / 1e9, in_rate / 1e9, 0.7 * capacity / 1e9, rate, args.serve_seconds))
    st = srv.run(arr, ids, check_every=16)
    if st.mismatches or st.failed:
        raise SystemExit("served run: %d mismatching, %d failed" % (st.mismatches, st.failed))
    out = st.summary()
    out.update({"capacity_GBps": round(capacity / 1e9, 3),
                "capacity_input_GBps": round(in_rate / 1e9, 3),
                "offered_GBps": round(0.7 * capacity / 1e9, 3),
                "offered_requests_per_s": round(rate, 1), "duration_s": args.serve_seconds,
                "client_pool_streams": pool.n, "budget_GB": 64, "max_inflight": 3,
                "latency": "arrival to completion observed on the host; each request's "
                           "compressed stream and table copied host-to-device on admission "
                           "(included)"})
    del srv
    if args.serve_records:
        # resident streams (not the served configuration): repeat requests decoded from the
        # item records each stream's first decode wrote
        rpool = serve.StreamPool([es for es, _ in items], [sym for _, sym in items], records=True)
        rsrv = serve.Server(rpool, 64e9, max_inflight=3, use_records=True)
        warm = rsrv.run(np.zeros(rpool.n), np.arange(rpool.n), check_every=1)
        if warm.mismatches:
            raise SystemExit("served records warm-up: %d mismatching streams" % warm.mismatches)
        rcap = rsrv.run(np.zeros(rpool.n), np.arange(rpool.n))
        keep = arr < min(args.serve_seconds, 30.0)
        rst = rsrv.run(arr[keep], ids[keep], check_every=16)
        if rst.mismatches:
            raise SystemExit("served from records: %d mismatching requests" % rst.mismatches)
        out["resident_from_item_records"] = dict(
            rst.summary(), capacity_GBps=round(rcap.symbols / rcap.wall_s / 1e9, 3),
            duration_s=min(args.serve_seconds, 30.0),
            note="streams resident on the device, repeat requests from item records")
        log("served from item records (resident)", json.dumps(out["resident_from_item_records"]))
        del rsrv, rpool
    del pool, items
    torch.cuda.empty_cache()
    return out


def cpu_baseline(items, budget_s=20.0, S=4):
    """Reference MulticoreDecoder on a bounded sample (streams of the benchmark), with
    subsequence size S (4 = the reference demo's)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    use_ref = ol.ref_available()
    total_bytes, total_s, nstreams, dropped_tail = 0, 0.0, 0, 0
    for es, sym in items:
        # the reference's overflow rounds never re-check a thread marked synchronised
        # when its predecessor's exit changes later (src/multicore_decoder.cc:57-75,
        # 109-111), so it is only correct when every thread's range is longer than the
        # self-synchronisation distance: sample streams with >= 8192 units per thread
        if es.n_units < 8192 * THREADS:
            continue
        units = np.concatenate([es.units, np.zeros(S + 16, np.uint32)])
        out = np.zeros(es.n_symbols, np.uint8)
        t0 = time.perf_counter()
        if use_ref:
            us = ol.ref().ref_multicore(S, THREADS, units, es.n_units, es.table, es.L,
                                        es.first_state, es.first_bit, out, es.n_symbols)
            dt = us / 1e6
        else:
            ol.oracle().mo_decode_multicore(S, THREADS, units, es.n_units, es.table, es.L,
                                            es.first_state, es.first_bit, out, es.n_symbols)
            dt = time.perf_counter() - t0
        if not np.array_equal(out[::-1], sym):
            # the reference never decodes zero-bit symbols that sit after the last unit
            # when n_units % S == 0 (the first-encoded symbols, decoded last); any other
            # difference is an error
            diff = np.nonzero(out[::-1] != sym)[0]
            if diff.max() >= 64:
                raise SystemExit("cpu baseline decode mismatch")
            dropped_tail += 1
        total_bytes += es.n_symbols
        total_s += dt
        nstreams += 1
        if total_s > budget_s:
            break
    return {"value": round(total_bytes / total_s / 1e9, 4), "unit": "GB/s", "cores": THREADS,
            "kind": "reference" if use_ref else "port",
            "sample": "MulticoreDecoder with S=%d%s and %d threads on the "
                      "first %d streams with >= 8192 units per thread (%.2f GB of "
                      "symbols) of the benchmark batch, one after another" %
                      (S, " (the reference demo's)" if S == 4 else "", THREADS, nstreams,
                       total_bytes / 1e9),
            "host_logical_cpus": os.cpu_count(),
            "streams_with_dropped_zero_bit_tail": dropped_tail}


def profiled_traffic(kernel, n_streams):
    """HBM bytes per launch of `kernel` on the 10,000-stream batch, measured with
    rocprofv3 FETCH_SIZE/WRITE_SIZE counter runs (scaled by a calibration kernel of
    known traffic) and recorded in profiles/batch10k_kernels.md; None if absent."""
    path = os.path.join(ROOT, "profiles", "batch10k_kernels.md")
    if n_streams != 10000 or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes "):
            vals = dict(kv.split("=") for kv in line.split()[1:])
            v = vals.get(kernel)
            return float(v) if v and float(v) > 0 else None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--streams", type=int, default=10000)
    ap.add_argument("--log2n", type=int, default=30)
    ap.add_argument("--no-single", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="duration of the served run (0 = skip)")
    ap.add_argument("--serve-pool", type=int, default=1000,
                    help="client streams (host memory) the served requests draw from")
    ap.add_argument("--serve-records", action="store_true",
                    help="also serve resident streams from item records (not the served "
                         "configuration)")
    ap.add_argument("--one-pipeline", action="store_true",
                    help="decode the batch as one pipeline (no concurrent parts), as the "
                         "per-kernel timing decode does; for per-dispatch profiles")
    ap.add_argument("--tables", type=int, default=0, metavar="N",
                    help="also build decoder tables on the device for the first N streams")
    ap.add_argument("--no-stress", action="store_true",
                    help="skip the two 2^32-symbol streams (lambda 0.002 and 3.0)")
    ap.add_argument("--no-interleaved", action="store_true",
                    help="skip the two-state interleaved leg")
    ap.add_argument("--no-c1", action="store_true",
                    help="skip the 4 MiB CPU-runnable configuration (GPU vs multicore)")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the per-stream plain-launch leg")
    ap.add_argument("--no-index", action="store_true",
                    help="skip the sync-point legs (decode from recorded sync points, windows)")
    ap.add_argument("--no-encode", action="store_true",
                    help="skip the device-encode leg")
    ap.add_argument("--no-encode-records", action="store_true",
                    help="in the device-encode leg, skip the encode with item records and "
                         "the decode from them")
    ap.add_argument("--windows", type=int, default=10000,
                    help="windows of 64 Ki symbols per call in the window leg")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run one elementwise kernel that reads and writes MIB MiB (known "
                         "HBM traffic, for scaling FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU; this build has no multi-GPU path (streams would "
              "shard across GPUs without exchange)", file=sys.stderr)
        raise SystemExit(2)
    import torch
    torch.cuda.set_device(0)
    if args.calibrate_copy:
        src = torch.ones(args.calibrate_copy << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.add(src, 1, out=dst)       # elementwise kernel: reads and writes MIB MiB
        torch.cuda.synchronize()
        log("calibration copy: %d MiB read + %d MiB written" % (args.calibrate_copy, args.calibrate_copy))
        del src, dst
    bres, bitems = bench_batched(args, torch)
    ms = bres["ms_batch"]
    value = bres["total_symbols"] / ms / 1e6
    plain_gbps = bres["total_symbols"] / bres["ms_plain"] / 1e6 if bres["ms_plain"] else None
    kms = bres["kernels_ms"]
    dom = max(kms, key=kms.get)
    # algorithmic bytes per launch: k_sync reads the compressed units once (4 B/unit);
    # k_write reads them and writes every decoded byte; other kernels are small.
    units4 = 4.0 * bres["total_units"]
    algo = {"sync": units4, "write": units4 + bres["total_symbols"]}.get(dom, units4)
    ach = algo / (kms[dom] / 1e3) / 1e9
    res = {"metric": METRIC, "value": round(value, 3), "unit": "GB/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
           "data": "synthetic (seeded Exp(lambda) symbols, tables and encoding on the host)",
           "config": {"workload": "batched: %d tANS streams, uint8 sizes log-uniform 4 KiB-16 MiB, "
                                  "lambda ~ LogU(0.005, 2.0), own 4096-state table each, one "
                                  "mans_decode_batch call per step (%s; kernels_ms and the "
                                  "roofline are one pipeline's kernels)" %
                                  (bres["n_streams"], "one pipeline" if args.one_pipeline else
                                   "up to three concurrent pipelines"),
                      "streams": bres["n_streams"], "total_symbols": bres["total_symbols"],
                      "compressed_bytes": int(units4), "L": 4096},
           "batched_GBps": round(value, 3),
           "per_stream_plain_GBps": round(plain_gbps, 3) if plain_gbps else None,
           "per_stream_plain_ms": round(bres["ms_plain"], 3) if plain_gbps else None,
           "batched_vs_plain": round(value / plain_gbps, 2) if plain_gbps else None,
           "kernels_ms": {k: round(v, 4) for k, v in kms.items()},
           "roofline": {"bound": "hbm", "kernel": "k_" + dom, "achieved": round(ach, 2),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                        "traffic": profiled_traffic("k_" + dom, bres["n_streams"]),
                        "traffic_source": "profiles/batch10k_kernels.md (rocprofv3 --pmc)",
                        "whole_decode_frac": round((units4 + bres["total_symbols"]) / ms / 1e6
                                                   / HBM_PEAK_GBS, 4)}}
    if "sync_points" in bres:
        res["from_sync_points"] = bres["sync_points"]
    log("batched", json.dumps({k: res[k] for k in ("batched_GBps", "per_stream_plain_GBps",
                                                   "batched_vs_plain", "kernels_ms")}))
    if not args.no_encode:
        res["device_encode"] = bench_encode(args, torch, bitems)
    if not args.no_single:
        per, _ = bench_single(args, torch)
        res["single_stream"] = per
    if args.tables:
        res["device_tables"] = bench_tables(args, torch)
    if not args.no_stress:
        res["stress"] = bench_stress(args, torch)
    if not args.no_interleaved:
        res["interleaved"] = bench_interleaved(args, torch)
    if not args.no_c1:
        res["c1_plain"] = bench_plain_c1(args, torch)
    if args.serve_seconds > 0:
        sv = bench_served(args, torch)
        res["served"] = sv
        res["latency_p50_ms"] = sv["p50_ms"]
        res["latency_p99_ms"] = sv["p99_ms"]
        log("served", json.dumps(sv))
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(bitems)
        log("cpu baseline", json.dumps(res["cpu_baseline"]))
        best = res.get("c1_plain", {}).get("cpu_best_S", 64)
        if best != 4:
            tuned = cpu_baseline(bitems, budget_s=10.0, S=best)
            res["cpu_baseline"]["tuned_S"] = {"S": best, "value": tuned["value"],
                                              "sample": tuned["sample"]}
            log("cpu baseline, tuned S", json.dumps(res["cpu_baseline"]["tuned_S"]))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
