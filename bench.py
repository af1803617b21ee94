"""Benchmark: tANS subsequence decoding on one MI355X.

Default workload ("plain", the at-scale single-stream configuration): one tANS stream
per lambda in {0.01, 0.05, 0.2, 1.0}, each 2^30 uint8 symbols s = min(255, floor(X)),
X ~ Exp(lambda), with a 4096-state table normalised from the stream's histogram.
A step decodes all four streams (inputs resident in HBM) through the C ABI; every
stream is checked bit-exact against its original symbols before timing.

Prints one JSON line.  `value` = decoded bytes of all streams / time of the timed
steps (decoded GB/s).  `roofline` is for the dominant kernel, timed with HIP events
on the decode stream; `cpu_baseline` times the reference MulticoreDecoder (compiled
from the reference sources into oracle/_ref, else the oracle's restatement) on a
bounded sample on this host.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--workload plain|batched]
"""
import argparse
import ctypes as C
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "decoded GB/s, batched vs per-stream plain launch; p50/p99 latency; bit-exact"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_streams(lams, n, L, seed0=1000):
    """Generate + encode one stream per lambda (host, product codec, threads)."""
    from multians_amd import _lib
    from multians_amd.decoder import EncodedStream
    res = [None] * len(lams)

    def one(i, lam):
        sym = _lib.gen_symbols(seed0 + i, lam, n, threads=4)
        hist = np.bincount(sym, minlength=256).astype(np.uint64)
        counts = _lib.normalize_counts(hist, L)
        enc, dec = _lib.build_tables(counts, L)
        units, fb, fs = _lib.encode(sym, enc, L)
        res[i] = (EncodedStream(units, fb, fs, dec, L, n, counts), sym)

    th = [threading.Thread(target=one, args=(i, lam)) for i, lam in enumerate(lams)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    return res


def _lib_names():
    from multians_amd import _lib
    return _lib.KERNEL_NAMES


def kernel_times(ds, torch, reps=3):
    """Per-kernel durations (ms) of one decode, from HIP events on the decode stream."""
    from multians_amd import _lib
    evs = [C.c_void_p() for _ in range(_lib.N_KERNEL_EVENTS)]
    L = _lib.lib()
    for e in evs:
        _lib.check(L.mans_event_create(C.byref(e)))
    arr = (C.c_void_p * _lib.N_KERNEL_EVENTS)(*[e.value for e in evs])
    tot = np.zeros(_lib.N_KERNEL_EVENTS - 1)
    for _ in range(reps):
        ds.decode(kernel_events=C.cast(arr, C.c_void_p))
        torch.cuda.synchronize()
        for k in range(_lib.N_KERNEL_EVENTS - 1):
            ms = C.c_float()
            _lib.check(L.mans_event_elapsed_ms(evs[k], evs[k + 1], C.byref(ms)))
            tot[k] += ms.value
    for e in evs:
        L.mans_event_destroy(e)
    return tot / reps


def cpu_baseline(streams, budget_s=20.0):
    """Reference MulticoreDecoder on a bounded prefix-sized sample of each stream."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    threads = min(os.cpu_count() or 1, 16)
    use_ref = ol.ref_available()
    total_bytes, total_s, samples = 0, 0.0, []
    for es, sym in streams:
        # sample: the same distribution, 64 Mi symbols, re-encoded (decode of a prefix
        # of units is not a valid stream)
        n = 1 << 26
        st, _ = ol.encode(sym[:n], es.counts, es.L)
        out = np.zeros(n, np.uint8)
        units = np.concatenate([st.units, np.zeros(16, np.uint32)])
        t0 = time.perf_counter()
        if use_ref:
            us = ol.ref().ref_multicore(4, threads, units, len(st.units), st.table, es.L,
                                        st.first_state, st.first_bit, out, n)
            dt = us / 1e6
        else:
            ol.oracle().mo_decode_multicore(4, threads, units, len(st.units), st.table, es.L,
                                            st.first_state, st.first_bit, out, n)
            dt = time.perf_counter() - t0
        assert np.array_equal(out[::-1], sym[:n]), "cpu baseline decode mismatch"
        total_bytes += n
        total_s += dt
        samples.append("%d Mi symbols" % (n >> 20))
        if total_s > budget_s:
            break
    return {"value": round(total_bytes / total_s / 1e9, 4), "unit": "GB/s", "cores": threads,
            "kind": "reference" if use_ref else "port",
            "sample": "MulticoreDecoder S=4 (the reference demo's), %d threads, on %d re-encoded "
                      "streams of 2^26 symbols from the same lambdas/tables" % (threads, len(samples)),
            "host_logical_cpus": os.cpu_count()}


def bench_plain(args, torch):
    from multians_amd.decoder import DeviceStream
    lams = [0.01, 0.05, 0.2, 1.0]
    n = 1 << args.log2n
    L = 4096
    t0 = time.time()
    streams = make_streams(lams, n, L)
    log("generated+encoded %d streams of %d symbols in %.1fs" % (len(lams), n, time.time() - t0))
    devs = []
    for (es, sym), lam in zip(streams, lams):
        ds = DeviceStream(es)
        out = ds.decode()
        ref = torch.from_numpy(sym).cuda()
        ok = bool(torch.equal(out[:n].flip(0), ref))
        st = ds.status()
        del ref
        log("lambda=%.3g units=%d ratio=%.2f bit-exact=%s status=%d" %
            (lam, es.n_units, n / (4.0 * es.n_units), ok, st))
        if not ok or st != 0:
            raise SystemExit("decode mismatch for lambda=%g" % lam)
        devs.append(ds)
    torch.cuda.synchronize()
    # per-lambda times (median of 5) and kernel breakdown
    per = []
    for ds, lam in zip(devs, lams):
        ts = []
        for _ in range(5):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            ds.decode()
            b.record()
            torch.cuda.synchronize()
            ts.append(a.elapsed_time(b))
        kt = kernel_times(ds, torch)
        ms = float(np.median(ts))
        algo = 4.0 * ds.es.n_units + n
        per.append({"lambda": lam, "ms": round(ms, 4), "decoded_GBps": round(n / ms / 1e6, 2),
                    "ratio": round(n / (4.0 * ds.es.n_units), 3),
                    "hbm_frac": round(algo / ms / 1e6 / HBM_PEAK_GBS, 4),
                    "kernels_ms": {k: round(float(v), 4) for k, v in
                                   zip(_lib_names(), kt)}})
        log(json.dumps(per[-1]))
    for _ in range(args.warmup):
        for ds in devs:
            ds.decode()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(args.steps):
        for ds in devs:
            ds.decode()
    b.record()
    torch.cuda.synchronize()
    ms_step = a.elapsed_time(b) / args.steps
    total = n * len(devs)
    value = total / ms_step / 1e6
    # dominant kernel over the four streams: summed kernel time
    ktot = {}
    for p in per:
        for k, v in p["kernels_ms"].items():
            ktot[k] = ktot.get(k, 0.0) + v
    dom = max(ktot, key=ktot.get)
    # algorithmic bytes per launch: k_sync reads the compressed units (4 B/unit); the
    # write kernel reads them and writes the decoded bytes; the others are small.
    units_all = sum(ds.es.n_units for ds in devs)
    algo = {"sync": 4.0 * units_all, "write": 4.0 * units_all + total}.get(dom, 4.0 * units_all)
    ach = algo / (ktot[dom] / 1e3) / 1e9
    roof = {"bound": "hbm", "kernel": "k_" + dom, "achieved": round(ach, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
            "whole_decode_frac": round((4.0 * units_all + total) / (ms_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
    res = {"metric": METRIC, "value": round(value, 3), "unit": "GB/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
           "data": "synthetic (seeded Exp(lambda) symbols, encoded on the host)",
           "config": {"workload": "plain single stream at scale: 4 streams x 2^%d uint8 symbols, "
                                  "lambda in {0.01,0.05,0.2,1.0}, L=4096" % args.log2n,
                      "streams": len(devs), "symbols_per_stream": n, "L": L},
           "per_lambda": per, "roofline": roof}
    return res, streams


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="plain", choices=["plain"])
    ap.add_argument("--log2n", type=int, default=30)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(streams shard across GPUs without exchange; see DESIGN.md)", file=sys.stderr)
        raise SystemExit(2)
    import torch
    torch.cuda.set_device(0)
    res, streams = bench_plain(args, torch)
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(streams)
        log("cpu baseline", json.dumps(res["cpu_baseline"]))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
