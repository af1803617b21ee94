"""Autotuner for the decoder's geometry: the sync-chunk length S and the speculative
warm-up W (units each chunk's guessed-state decode starts early), per stream shape.

    python -m multians_amd.tune [--streams 10000] [--out DIR]

S trades lanes against failed walks: a short chunk gives a stream more lanes (lower
latency) but more walks that cross their chunk without merging.  W trades work against
walk length: a warm-up longer than the merge distance puts most chunks' guessed paths on
the true path at the chunk start, so their walks merge at once.  Three rule sets are fitted
on the device, all keyed by symbols-per-unit bucket (n_symbols / n_units, the entropy
proxy the decoder knows before decoding) and stream size:

  * single (mans_decode, one stream on the GPU: latency): for one stream per (bucket,
    size) -- symbols 2^20 .. 2^30, lambda chosen to land in each bucket -- every (S, W) of
    a grid is timed (median of 3 decodes) and the fastest kept;
  * lineage (mans_decode, one stream): for one stream per (bucket, size) on finer buckets
    of their own, the chunked decode under the single rules against lineage decodes at
    spacings 32-256 units; each cell gets the fastest spacing when it wins by 3%, else
    spacing 0 (the chunked pipeline);
  * batch (mans_decode_batch: throughput of the whole batch): the built-in shape
    S = clamp(pow2ceil(n_units / 256), floor_b, safe_b), with warm-ups W1_b for streams
    of one 256-chunk range and W2_b for longer ones, fitted by coordinate descent on the
    decode time of the benchmark's mix (uint8 sizes log-uniform 4 KiB-16 MiB, lambda ~
    LogU(0.005, 2.0), 4096 states; other seeds than the benchmark's) as one batch.

Streams are made on the device (mans_gen_symbols_device, mans_build_tables_device,
mans_encode_device: the host generator's and host codec's bytes); every decode the tuner
times is compared with its symbols.  The result is a rule table (mans_geometry_rule)
written as JSON and as multians_amd/csrc/tuned_rules.inc, which the library compiles in as
its default.
"""
import argparse
import ctypes as C
import json
import os
import time

import numpy as np

from . import _lib

PKG = os.path.dirname(os.path.abspath(__file__))

SPU_EDGES = [4.6, 6.5, 10.0, 20.0, 40.0, 1e30]
# single-stream buckets (finer: a stream alone on the GPU is latency-bound, and its
# merge distance, which sets that latency, moves quickly with the entropy) and the lambda
# of each bucket's tuning stream (its symbols per unit at L = 4096 in brackets):
# 0.01 (4.3), 0.03 (5.0), 0.06 (5.8), 0.1 (6.6), 0.15 (7.6), 0.25 (9.4), 0.5 (13),
# 1.0 (21), 1.6 (32), 3.0 (106)
SINGLE_EDGES = [4.6, 5.2, 6.0, 7.0, 8.5, 11.0, 16.0, 25.0, 50.0, 1e30]
BUCKET_LAMBDA = [0.01, 0.03, 0.06, 0.1, 0.15, 0.25, 0.5, 1.0, 1.6, 3.0]
SINGLE_LOG2N = [20, 22, 24, 26, 28, 30]
SINGLE_LOG2N_LARGE_L = SINGLE_LOG2N + [32]      # 4096 states: also the 4 GiB streams
SINGLE_L = [1024, 4096]
SINGLE_S = [32, 64, 128, 256, 512, 1024, 2048, 4096]
SINGLE_W = [0, 256, 512, 1024, 2048, 4096]
# the built-in heuristic's constants for batches (L = 4096; L <= 1024 halves safe)
FLOOR0 = [128, 256, 512, 512, 512, 512]
SAFE0 = [2048, 4096, 8192, 8192, 8192, 2048]
# batch streams above BIG units take these chunk lengths (the single-stream sweep of
# 2^30-symbol streams per bucket: fewer, longer chunks than the fill rule would give;
# batches of the benchmark mix never hold such streams)
BIG = 1 << 24
BIG_S = [1024, 1024, 4096, 4096, 4096, 2048]
FLOORS = [128, 256, 512, 1024, 2048]
SAFES = [1024, 2048, 4096, 8192, 16384]


def batch_rules(floor, safe, warm1=None, warm2=None):
    """Batch rule table: clamp(pow2ceil(n/256), floor_b, safe_b) per spu bucket, for
    L = 1024 (safe halved) and L = 4096, warm-up warm1_b for the size classes whose streams
    all fit one 256-chunk range and warm2_b for the others (None: the built-in warm-up,
    512 units for a stream of one range and none for longer ones)."""
    rules = []
    for L, div in ((1024, 2), (4096, 1)):
        for b, spu_max in enumerate(SPU_EDGES):
            prev = None
            for k in range(8, 41):           # units in (2^(k-1), 2^k]
                units_max = 1 << k
                fill = max(1, units_max >> 8)
                S = max(floor[b], min(max(safe[b] // div, 128), fill), 128)
                if units_max > BIG:
                    S = BIG_S[b]
                if warm1 is None:
                    W = _lib.WARM_AUTO
                else:
                    W = warm1[b] if units_max <= 256 * S else warm2[b]
                r = (L, spu_max, units_max, S, W, _lib.RULE_BATCH)
                if prev is not None and prev[3:5] == r[3:5]:
                    rules[-1] = r
                else:
                    rules.append(r)
                prev = rules[-1]
    return rules


def _bucket(spu):
    return next(j for j, e in enumerate(SPU_EDGES) if spu <= e)


def _time(fn, torch, reps=3):
    ts = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        z = torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        z.record()
        torch.cuda.synchronize()
        ts.append(a.elapsed_time(z))
    return float(np.median(ts))


def tune_single(log):
    """Single-stream rules: per state count, bucket and size, the fastest (S, W) of the
    grid (L <= 1024 from the 1024-state streams, larger L from the 4096-state ones)."""
    import torch
    from .decoder import GeneratedBatch, workspace_bytes
    rules, rec = [], []
    for L in SINGLE_L:
        for b, lam in enumerate(BUCKET_LAMBDA):
            found = []
            for log2n in (SINGLE_LOG2N if L < SINGLE_L[-1] else SINGLE_LOG2N_LARGE_L):
                gb = GeneratedBatch([(9000 + 97 * b + log2n + L, lam, 1 << log2n, L)])
                nu, n = int(gb.n_units[0]), int(gb.n_symbols[0])
                out = torch.empty(n, dtype=torch.uint8, device="cuda")
                ws = torch.zeros(workspace_bytes(nu, L), dtype=torch.uint8, device="cuda")
                ref = gb.symbols(0).flip(0)
                times = {}
                for S in SINGLE_S:
                    if S > max(32, nu // 4):
                        continue
                    for W in SINGLE_W:
                        # the warm-up multiplies the spec work by (S + W) / S: bounded
                        # where the stream alone fills the GPU
                        if nu * (S + W) / S > max(1 << 26, 4 * nu):
                            continue
                        _lib.set_geometry_rules([(1 << 16, 1e30, 1 << 40, S, W,
                                                  _lib.RULE_SINGLE)])
                        gb.decode_one(0, out, ws)
                        torch.cuda.synchronize()
                        if gb.status(ws) != 0 or not torch.equal(out, ref):
                            raise RuntimeError("tuning decode mismatch (S=%d, W=%d)" % (S, W))
                        times[(S, W)] = _time(lambda: gb.decode_one(0, out, ws), torch)
                best = min(times, key=times.get)
                spu = n / nu
                log("single L=%d bucket %d (lambda %g, %.1f symbols/unit), 2^%d symbols, %d "
                    "units: S=%d W=%d %.3f ms" % (L, b, lam, spu, log2n, nu, best[0], best[1],
                                                  times[best]))
                found.append((nu, best))
                rec.append({"L": L, "bucket": b, "lambda": lam, "log2n": log2n, "units": nu,
                            "spu": spu, "S": best[0], "W": best[1], "ms": round(times[best], 4),
                            "grid_ms": {"%d/%d" % k: round(v, 4) for k, v in times.items()}})
                del gb, out, ws, ref
                torch.cuda.empty_cache()
            # size boundaries: geometric midpoints between the tuned sizes
            for i, (nu, (S, W)) in enumerate(found):
                hi = int(np.sqrt(nu * found[i + 1][0])) if i + 1 < len(found) else 1 << 40
                rules.append((L if L < SINGLE_L[-1] else 1 << 14, SINGLE_EDGES[b], hi, S, W,
                              _lib.RULE_SINGLE))
    _lib.set_geometry_rules([])
    return rules, rec


LIN_D = [32, 64, 96, 128, 192, 256]
# lineage buckets (finer than the single-stream ones around 8 symbols per unit, where the
# choice between lineages and chunks turns at 2^30 symbols) and their tuning streams'
# lambda (symbols per unit at L = 4096 in brackets): 0.01 (4.3), 0.03 (5.0), 0.06 (5.8),
# 0.1 (6.6), 0.15 (7.6), 0.2 (8.5), 0.25 (9.4), 0.5 (13), 1.0 (21), 1.6 (32), 3.0 (106)
LIN_EDGES = [4.6, 5.2, 6.0, 7.0, 8.0, 9.0, 11.0, 16.0, 25.0, 50.0, 1e30]
LIN_LAMBDA = [0.01, 0.03, 0.06, 0.1, 0.15, 0.2, 0.25, 0.5, 1.0, 1.6, 3.0]


def tune_lineage(single_rules, log, Ls=None):
    """Lineage rules: for every (state count, bucket, size) of the single-stream sweep, the
    chunked decode under the single-stream rules against lineage decodes at each spacing of
    LIN_D (median of 5 each, every decode checked against its symbols and required not to
    fall back).  A cell takes the fastest spacing when it beats the chunked decode by 3%,
    else a rule with spacing 0 (the chunked pipeline), so the first-hit rules never give a
    cell a spacing that lost there."""
    import torch
    from .decoder import GeneratedBatch, lineage_stats, workspace_bytes
    rules, rec = [], []
    for L in (Ls or SINGLE_L):
        for b, lam in enumerate(LIN_LAMBDA):
            found = []
            for log2n in (SINGLE_LOG2N if L < SINGLE_L[-1] else SINGLE_LOG2N_LARGE_L):
                gb = GeneratedBatch([(9000 + 97 * b + log2n + L, lam, 1 << log2n, L)])
                nu, n = int(gb.n_units[0]), int(gb.n_symbols[0])
                out = torch.empty(n, dtype=torch.uint8, device="cuda")
                ws = torch.zeros(workspace_bytes(nu, L), dtype=torch.uint8, device="cuda")
                ref = gb.symbols(0).flip(0)
                times = {}
                for D in [0] + LIN_D:
                    lin = [(1 << 16, 1e30, 1 << 40, D, 0, _lib.RULE_LINEAGE)] if D else []
                    _lib.set_geometry_rules(single_rules + lin)
                    out.zero_()
                    gb.decode_one(0, out, ws)
                    torch.cuda.synchronize()
                    if gb.status(ws) != 0 or not torch.equal(out, ref):
                        raise RuntimeError("tuning decode mismatch (D=%d)" % D)
                    if D and lineage_stats(ws)[0]:
                        continue                    # fell back: not a lineage time
                    times[D] = _time(lambda: gb.decode_one(0, out, ws), torch, reps=5)
                lin_t = {D: t for D, t in times.items() if D}
                best = min(lin_t, key=lin_t.get) if lin_t else 0
                pick = best if best and lin_t[best] < 0.97 * times[0] else 0
                log("lineage L=%d bucket %d (lambda %g), 2^%d symbols, %d units: chunked %.3f "
                    "ms, best lineage D=%d %.3f ms -> %s" % (
                        L, b, lam, log2n, nu, times[0], best, lin_t.get(best, float("nan")),
                        ("D=%d" % pick) if pick else "chunked"))
                found.append((nu, pick))
                rec.append({"L": L, "bucket": b, "lambda": lam, "log2n": log2n, "units": nu,
                            "chunked_ms": round(times[0], 4), "D": pick,
                            "lineage_ms": {str(D): round(t, 4) for D, t in lin_t.items()}})
                del gb, out, ws, ref
                torch.cuda.empty_cache()
            for i, (nu, D) in enumerate(found):
                hi = int(np.sqrt(nu * found[i + 1][0])) if i + 1 < len(found) else 1 << 40
                r = (L if L < SINGLE_L[-1] else 1 << 14, LIN_EDGES[b], hi, D, 0,
                     _lib.RULE_LINEAGE)
                if rules and rules[-1][0] == r[0] and rules[-1][1] == r[1] and rules[-1][3] == D:
                    rules[-1] = r                   # same spacing as the smaller sizes
                else:
                    rules.append(r)
    _lib.set_geometry_rules([])
    return rules, rec


def tune_mixed(n_streams, single_rules, log, sweeps=2):
    """Batch rules by coordinate descent on the mixed batch's decode time."""
    import torch
    from .decoder import GeneratedBatch
    rng = np.random.default_rng(20261207)
    sizes = np.floor(np.exp(rng.uniform(np.log(4096), np.log(16 << 20), n_streams))).astype(np.int64)
    lams = np.exp(rng.uniform(np.log(0.005), np.log(2.0), n_streams))
    specs = [(777_000_000 + i, float(lams[i]), int(sizes[i]), 4096) for i in range(n_streams)]
    gb = GeneratedBatch(specs)
    ooff = np.concatenate([[0], np.cumsum(gb.n_symbols)])
    out = torch.empty(int(ooff[-1]), dtype=torch.uint8, device="cuda")
    descs = gb.descs(out, ooff)
    b = C.c_uint64()
    _lib.set_geometry_rules([(1 << 16, 1e30, 1 << 40, 128, 4096)])     # largest workspace
    _lib.check(_lib.lib().mans_batch_workspace_bytes(descs, gb.n, C.byref(b)))
    ws = torch.zeros(b.value, dtype=torch.uint8, device="cuda")
    o = _lib.BatchOpts()
    o.workspace, o.workspace_bytes = ws.data_ptr(), ws.numel()
    o.stream = torch.cuda.current_stream().cuda_stream

    def run():
        _lib.check(_lib.lib().mans_decode_batch(descs, gb.n, C.byref(o), None))

    def evaluate(rules, check):
        _lib.set_geometry_rules(single_rules + rules)
        run()
        torch.cuda.synchronize()
        if check:
            for i in range(gb.n):
                if not torch.equal(out[int(ooff[i]):int(ooff[i + 1])], gb.symbols(i).flip(0)):
                    raise RuntimeError("tuning decode mismatch, stream %d" % i)
        return _time(run, torch)

    # the warm-up of batched streams stays the built-in one: a search over warm1_b / warm2_b
    # (0-2048 units) on this mix found no change worth 1% of the batch's time
    floor, safe, w1, w2 = list(FLOOR0), list(SAFE0), None, None
    best = evaluate(batch_rules(floor, safe, w1, w2), True)
    log("mixed: %d streams, start %.2f ms" % (n_streams, best))
    for sweep in range(sweeps):
        changed = False
        for bk in range(len(SPU_EDGES)):
            for which, cands in (("floor", FLOORS), ("safe", SAFES)):
                for v in cands:
                    cur = {"floor": floor, "safe": safe, "warm1": w1, "warm2": w2}
                    new = {k: (list(x) if x is not None else None) for k, x in cur.items()}
                    new[which][bk] = v
                    if new["floor"][bk] > new["safe"][bk] or new == cur:
                        continue
                    r = batch_rules(new["floor"], new["safe"], new["warm1"], new["warm2"])
                    t = evaluate(r, False)
                    if t < best * 0.99:
                        t = evaluate(r, True)                              # checked
                        if t < best * 0.99:
                            log("  sweep %d bucket %d %s=%d: %.2f ms (was %.2f)" %
                                (sweep, bk, which, v, t, best))
                            best, changed = t, True
                            floor, safe, w1, w2 = (new["floor"], new["safe"], new["warm1"],
                                                   new["warm2"])
        if not changed:
            break
    final = evaluate(batch_rules(floor, safe, w1, w2), True)
    _lib.set_geometry_rules([])
    total = int(ooff[-1])
    log("mixed: floor %s safe %s warm1 %s warm2 %s -> %.2f ms (%.1f GB/s)" %
        (floor, safe, w1, w2, final, total / final / 1e6))
    del gb, out, ws
    torch.cuda.empty_cache()
    return (floor, safe, w1, w2), final


def write_rules(rules, meta, out_json, out_inc):
    os.makedirs(os.path.dirname(out_json), exist_ok=True)
    with open(out_json, "w") as f:
        json.dump(dict(meta, rules=[{"L_max": r[0], "spu_max": r[1], "units_max": r[2],
                                     "chunk_units": r[3], "warm_units": r[4], "scope": r[5]}
                                    for r in rules]), f, indent=1)
    with open(out_inc, "w") as f:
        f.write("// Generated by multians_amd/tune.py on MI355X (gfx950): built-in geometry\n"
                "// rules {L_max, chunk_units, spu_max, units_max, warm_units, scope},\n"
                "// matched first-hit within their scope (1 single-stream decode, 2 batch).\n")
        for r in rules:
            spu = "1e30" if r[1] >= 1e29 else repr(float(r[1]))
            f.write("{%du, %du, %s, %dull, %du, %du},\n" % (r[0], r[3], spu, r[2], r[4], r[5]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--streams", type=int, default=10000)
    ap.add_argument("--out", default=os.path.join(PKG, "tuning"))
    ap.add_argument("--inc", default=os.path.join(PKG, "csrc", "tuned_rules.inc"))
    ap.add_argument("--no-single", action="store_true",
                    help="keep the single-stream rules of the existing JSON")
    ap.add_argument("--no-mixed", action="store_true",
                    help="keep the batch parameters of the existing JSON")
    ap.add_argument("--no-lineage", action="store_true",
                    help="keep the lineage rules of the existing JSON")
    args = ap.parse_args()
    t0 = time.time()

    def log(*a):
        print(*a, flush=True)
    import torch
    torch.cuda.set_device(0)
    out_json = os.path.join(args.out, "gfx950.json")
    if args.no_single:
        meta0 = json.load(open(out_json))
        single = [tuple(r[k] for k in ("L_max", "spu_max", "units_max", "chunk_units",
                                       "warm_units", "scope"))
                  for r in meta0["rules"] if r["scope"] == _lib.RULE_SINGLE]
        single_rec = meta0.get("single", [])
    else:
        single, single_rec = tune_single(log)
    if args.no_mixed:
        meta0 = json.load(open(out_json))
        floor, safe = meta0["floor"], meta0["safe"]
        w1, w2 = meta0.get("warm1"), meta0.get("warm2")
        ms = meta0["mixed_ms"]
    else:
        (floor, safe, w1, w2), ms = tune_mixed(args.streams, single, log)
    if args.no_lineage:
        meta0 = json.load(open(out_json))
        lineage = [tuple(r[k] for k in ("L_max", "spu_max", "units_max", "chunk_units",
                                        "warm_units", "scope"))
                   for r in meta0["rules"] if r["scope"] == _lib.RULE_LINEAGE]
        lineage_rec = meta0.get("lineage", [])
    else:
        lineage, lineage_rec = tune_lineage(single, log)
    rules = single + lineage + batch_rules(floor, safe, w1, w2)
    meta = {"device": "MI355X (gfx950)", "floor": floor, "safe": safe, "warm1": w1,
            "warm2": w2, "spu_edges": SPU_EDGES, "single_edges": SINGLE_EDGES,
            "mixed_ms": ms, "mixed_streams": args.streams, "single": single_rec,
            "lineage": lineage_rec}
    write_rules(rules, meta, out_json, args.inc)
    log("wrote %d rules in %.0fs" % (len(rules), time.time() - t0))


if __name__ == "__main__":
    main()
