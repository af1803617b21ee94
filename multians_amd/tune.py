"""Autotuner for the sync-chunk length S (the decoder's one free geometry knob).

    python -m multians_amd.tune [--streams 10000] [--out DIR]

S trades lanes against failed walks: a short chunk gives a stream more lanes (lower
latency) but more walks that cross their chunk without merging (more chained and
re-walked work).  The built-in choice is

    S = clamp(pow2ceil(n_units / 256), floor_b, safe_b)

per symbols-per-unit bucket b (symbols per unit = n_symbols / n_units, the entropy
proxy the decoder knows before decoding), with the chunk count of one full 256-lane
workgroup as the fill target.  The tuner fits (floor_b, safe_b) by coordinate descent
on measured decode time of a workload:

  * mixed: the batched benchmark mix (uint8 sizes log-uniform 4 KiB-16 MiB,
    lambda ~ LogU(0.005, 2.0), 4096 states; other seeds than the benchmark's) decoded
    as one batch; the objective is the whole batch's decode time, so the slowest
    streams' critical paths count, not per-stream throughput;
  * single: one 2^30-symbol stream per lambda in {0.01, 0.05, 0.2, 1.0}, each tuned
    alone over S for streams of more than 2^24 units in its bucket.

Every table's output is checked against the symbols when it is adopted.  The result is
a rule table (mans_geometry_rule) written as JSON and as multians_amd/csrc/tuned_rules.inc,
which the library compiles in as its default.
"""
import argparse
import json
import os
import time

import numpy as np

from . import _lib

PKG = os.path.dirname(os.path.abspath(__file__))

SPU_EDGES = [4.6, 6.5, 10.0, 20.0, 40.0, 1e30]
# the built-in heuristic's constants (L = 4096; L <= 1024 halves safe)
FLOOR0 = [128, 256, 512, 512, 512, 512]
SAFE0 = [2048, 4096, 8192, 8192, 8192, 2048]
FLOORS = [128, 256, 512, 1024, 2048]
SAFES = [1024, 2048, 4096, 8192, 16384]
BIG = 1 << 24            # streams above this many units use the single-stream rules


def heuristic_rules(floor, safe, big=None):
    """Rule table reproducing clamp(pow2ceil(n/256), floor_b, safe_b) per spu bucket, for
    L = 1024 (safe halved) and L = 4096; `big` = {bucket: S} overrides streams > BIG."""
    rules = []
    for L, div in ((1024, 2), (4096, 1)):
        for b, spu_max in enumerate(SPU_EDGES):
            prev = None
            for k in range(8, 41):           # units in (2^(k-1), 2^k]
                units_max = 1 << k
                if big is not None and b in big and units_max > BIG:
                    S = big[b]
                else:
                    fill = max(1, units_max >> 8)
                    S = max(floor[b], min(max(safe[b] // div, 128), fill), 128)
                if prev is not None and prev[3] == S:
                    rules[-1] = (L, spu_max, units_max, S)
                else:
                    rules.append((L, spu_max, units_max, S))
                prev = rules[-1]
    return rules


def _make(specs, threads=16):
    from concurrent.futures import ThreadPoolExecutor
    from .decoder import EncodedStream

    def one(sp):
        seed, lam, n, L = sp
        units, fb, fs, dec, sym = _lib.make_stream(seed, lam, n, L, True)
        return EncodedStream(units, fb, fs, dec, L, n), sym
    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(one, specs))


def _evaluate(batch, items, rules, torch, reps=3, check=True):
    _lib.set_geometry_rules(rules)
    batch.decode()
    torch.cuda.synchronize()
    if batch.status() != 0:
        raise RuntimeError("decode status %d" % batch.status())
    if check:
        step = max(1, len(items) // 16)
        for i in range(0, len(items), step):
            if not np.array_equal(batch.output(i).cpu().numpy()[::-1], items[i][1]):
                raise RuntimeError("tuning decode mismatch, stream %d" % i)
    a = torch.cuda.Event(enable_timing=True)
    z = torch.cuda.Event(enable_timing=True)
    ts = []
    for _ in range(reps):
        a.record()
        batch.decode()
        z.record()
        torch.cuda.synchronize()
        ts.append(a.elapsed_time(z))
    return float(np.median(ts))


def tune_mixed(n_streams, log, sweeps=2):
    import torch
    from .decoder import DeviceBatch
    rng = np.random.default_rng(20261207)
    sizes = np.floor(np.exp(rng.uniform(np.log(4096), np.log(16 << 20), n_streams))).astype(np.int64)
    lams = np.exp(rng.uniform(np.log(0.005), np.log(2.0), n_streams))
    specs = [(777_000_000 + i, float(lams[i]), int(sizes[i]), 4096) for i in range(n_streams)]
    items = _make(specs)
    _lib.set_geometry_rules([(1 << 16, 1e30, 1 << 40, 128)])     # largest workspace
    batch = DeviceBatch([es for es, _ in items])
    floor, safe = list(FLOOR0), list(SAFE0)
    best = _evaluate(batch, items, heuristic_rules(floor, safe), torch)
    log("mixed: %d streams, start %.2f ms" % (n_streams, best))
    for sweep in range(sweeps):
        changed = False
        for b in range(len(SPU_EDGES)):
            for which, cands in (("floor", FLOORS), ("safe", SAFES)):
                for v in cands:
                    f2, s2 = list(floor), list(safe)
                    (f2 if which == "floor" else s2)[b] = v
                    if f2[b] > s2[b] or (f2 == floor and s2 == safe):
                        continue
                    t = _evaluate(batch, items, heuristic_rules(f2, s2), torch, check=False)
                    if t < best * 0.99:
                        t = _evaluate(batch, items, heuristic_rules(f2, s2), torch)   # checked
                        if t < best * 0.99:
                            log("  sweep %d bucket %d %s=%d: %.2f ms (was %.2f)" %
                                (sweep, b, which, v, t, best))
                            best, floor, safe, changed = t, f2, s2, True
        if not changed:
            break
    final = _evaluate(batch, items, heuristic_rules(floor, safe), torch)
    _lib.set_geometry_rules([])
    total = sum(es.n_symbols for es, _ in items)
    log("mixed: floor %s safe %s -> %.2f ms (%.1f GB/s)" % (floor, safe, final, total / final / 1e6))
    del batch, items
    torch.cuda.empty_cache()
    return floor, safe, final


def tune_single(log):
    import torch
    from .decoder import DeviceBatch
    big = {}
    out = []
    for i, lam in enumerate([0.01, 0.05, 0.2, 1.0]):
        items = _make([(5000 + i, lam, 1 << 30, 4096)])
        es = items[0][0]
        spu = es.n_symbols / es.n_units
        b = next(j for j, e in enumerate(SPU_EDGES) if spu <= e)
        _lib.set_geometry_rules([(1 << 16, 1e30, 1 << 40, 128)])
        batch = DeviceBatch([es])
        times = {}
        for S in [512, 1024, 2048, 4096, 8192]:
            times[S] = round(_evaluate(batch, items, [(1 << 16, 1e30, 1 << 40, S)], torch), 3)
        S_best = min(times, key=times.get)
        big[b] = S_best
        log("single lambda=%g (%.1f symbols/unit, bucket %d): %s -> S=%d" % (lam, spu, b, times, S_best))
        out.append({"lambda": lam, "spu": spu, "bucket": b, "ms": times, "best": S_best})
        _lib.set_geometry_rules([])
        del batch, items
        torch.cuda.empty_cache()
    return big, out


def write_rules(rules, meta, out_json, out_inc):
    os.makedirs(os.path.dirname(out_json), exist_ok=True)
    with open(out_json, "w") as f:
        json.dump(dict(meta, rules=[{"L_max": r[0], "spu_max": r[1], "units_max": r[2],
                                     "chunk_units": r[3]} for r in rules]), f, indent=1)
    with open(out_inc, "w") as f:
        f.write("// Generated by multians_amd/tune.py on MI355X (gfx950): built-in chunk-length\n"
                "// rules {L_max, chunk_units, spu_max, units_max}, matched first-hit.\n")
        for r in rules:
            spu = "1e30" if r[1] >= 1e29 else repr(float(r[1]))
            f.write("{%du, %du, %s, %dull},\n" % (r[0], r[3], spu, r[2]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--streams", type=int, default=10000)
    ap.add_argument("--out", default=os.path.join(PKG, "tuning"))
    ap.add_argument("--inc", default=os.path.join(PKG, "csrc", "tuned_rules.inc"))
    args = ap.parse_args()
    t0 = time.time()

    def log(*a):
        print(*a, flush=True)
    big, single = tune_single(log)
    floor, safe, ms = tune_mixed(args.streams, log)
    rules = heuristic_rules(floor, safe, big)
    meta = {"device": "MI355X (gfx950)", "floor": floor, "safe": safe, "spu_edges": SPU_EDGES,
            "mixed_ms": ms, "mixed_streams": args.streams, "single": single,
            "single_big_units": BIG}
    write_rules(rules, meta, os.path.join(args.out, "gfx950.json"), args.inc)
    log("wrote %d rules in %.0fs" % (len(rules), time.time() - t0))


if __name__ == "__main__":
    main()
