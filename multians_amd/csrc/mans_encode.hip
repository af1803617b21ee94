// mans_encode.hip -- tANS encoder on the device, bit-exact with the host codec's
// mans_encode (which restates ANSEncoder::encode, src/ans_encoder.cc:13-97, and the unit
// reversal of CUHDInputBuffer, src/multians_input_buffer.cc:29): state x starts at L,
// each symbol appends its code MSB-first and moves x on, the last unit keeps first_bit
// bits left-aligned (a zero unit if none), and the units are stored in decoder order.
//
// Encoding from a wrong state also converges on the true state path (the state is
// shifted right before each lookup, which maps many states to one), so a stream is cut
// into chunks that encode in parallel:
//   k_enc_sync   per 256-chunk range: each lane encodes its chunk from the start state
//                L (a guess for chunks >= 1), storing the state and the bits emitted at
//                every checkpoint (kEK symbols).
//   k_enc_walk   per range: each lane encodes its chunk again from its predecessor's
//                spec exit until the state meets the stored one.
//   k_enc_fix    per stream: rounds that re-walk every chunk whose recorded start is not
//                its predecessor's exit (range seams, rare failed walks), then the
//                exclusive scan of bits per chunk, the stream's results, and zeroed
//                units where chunks share one.
//   k_enc_write  per range: each lane encodes its chunk from its true start and writes
//                its bits at their offset (whole units by plain stores, units shared with
//                a neighbour chunk by atomicOr into the zeroed unit).
// The encoder's table comes from the decoder table (the decoder's own input): symbol s
// owns f_s states y whose `next` fields are f_s .. 2f_s - 1; encoding x emits
// nb = m_s - (x < f_s << m_s) bits, m_s = log2(L) - floor(log2 f_s), and moves to the
// state y of s with next = x >> nb.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "mans_c_api.h"
#include "mans_internal.h"

namespace {

constexpr uint32_t kEWG = 256;          // lanes (chunks) per workgroup
constexpr uint32_t kEC = 4096;          // symbols per chunk
constexpr uint32_t kEK = 32;            // symbols per checkpoint
constexpr uint32_t kECK = kEC / kEK;    // checkpoints per chunk
constexpr uint32_t kNoneE = 0xFFFFFFFFu;

// status bits (mans_encode_result.status)
constexpr uint32_t kEncBadTable = 1u;   // a `next` field outside its symbol's range
constexpr uint32_t kEncNoCode = 2u;     // a symbol with no state in the table
constexpr uint32_t kEncCapacity = 4u;   // more units than units_capacity
constexpr uint32_t kEncNoConverge = 8u; // fixup rounds stopped at their bound, inconsistent

struct EncPlan {
    const uint8_t* sym;
    const uint32_t* table;
    uint32_t* units;
    mans_sync_point* items;     // optional: the decoder's 32-unit item records
    uint64_t n;
    uint64_t cap;
    uint32_t L, R;
    uint32_t n_chunks, chunk_base;
    uint64_t ck_base;
    uint32_t task_base, pad;
};

struct EncWork {
    const EncPlan* plans;
    uint32_t n_jobs, n_tasks;
    // per chunk
    uint32_t *entry, *mq, *wbits, *wexit, *wtot, *sexit, *stot, *fl, *fl2, *ff;
    uint64_t* off;
    // per checkpoint
    uint32_t *ckx, *ckb;
    // per job
    mans_encode_result* res;
    uint32_t round_limit;       // k_enc_fix rounds per stream (0 = n_chunks + 1, the bound)
};

__device__ uint32_t job_of_task(const EncWork& w, uint32_t t) {
    uint32_t lo = 0, hi = w.n_jobs - 1u;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1u) >> 1;
        if (w.plans[mid].task_base <= t) lo = mid; else hi = mid - 1u;
    }
    return lo;
}

// Encoder tables in LDS: per symbol one 64-bit word {f | base << 16, thr | m << 16}
// (f: the symbol's state count, base: its first slot in ys, m = log2 L - floor(log2 f),
// thr = f << m), read with one ds_read_b64 per step; ys[base + i - f] = the state index y
// of the symbol with next(y) = i.
struct EncTab {
    const uint2* sym;
    uint16_t* ys;
};

__host__ __device__ constexpr size_t enc_lds_bytes(uint32_t L) { return 4u * 256u * 4u + 2u * L; }

// Builds the tables for plan p (all threads); returns status bits (uniform).
__device__ uint32_t stage_enc(const EncPlan& p, uint32_t* lds, EncTab& t) {
    __shared__ uint32_t s_bad;
    uint2* sym = (uint2*)lds;                  // 256 x 8 bytes
    uint32_t* cnt = lds + 512;
    t.sym = sym;
    t.ys = (uint16_t*)(lds + 1024);
    if (threadIdx.x == 0) s_bad = 0u;
    cnt[threadIdx.x] = 0u;
    for (uint32_t k = threadIdx.x; k < p.L; k += kEWG) t.ys[k] = 0xFFFFu;
    __syncthreads();
    for (uint32_t y = threadIdx.x; y < p.L; y += kEWG) atomicAdd(&cnt[(p.table[y] >> 16) & 0xFFu], 1u);
    __syncthreads();
    if (threadIdx.x < 64u) {               // exclusive prefix of the counts (one wave)
        uint32_t v[4], tot = 0;
        for (uint32_t k = 0; k < 4; ++k) { v[k] = cnt[threadIdx.x * 4u + k]; tot += v[k]; }
        uint32_t incl = tot;
        for (uint32_t d = 1; d < 64u; d <<= 1) {
            const uint32_t o = __shfl_up(incl, d, 64);
            if (threadIdx.x >= d) incl += o;
        }
        uint32_t run = incl - tot;
        for (uint32_t k = 0; k < 4; ++k) {
            const uint32_t sy = threadIdx.x * 4u + k, f = v[k];
            const uint32_t m = f ? p.R - (31u - (uint32_t)__builtin_clz(f)) : 0u;
            sym[sy] = make_uint2(f | (run << 16), (f << m) | (m << 16));
            run += f;
        }
    }
    __syncthreads();
    for (uint32_t y = threadIdx.x; y < p.L; y += kEWG) {
        const uint32_t e = p.table[y];
        const uint32_t s = (e >> 16) & 0xFFu, i = e & 0xFFFFu;
        const uint32_t f = sym[s].x & 0xFFFFu;
        if (i < f || i >= 2u * f) {
            s_bad = 1u;
        } else {
            t.ys[(sym[s].x >> 16) + i - f] = (uint16_t)y;
        }
    }
    __syncthreads();
    // every (symbol, next) pair exactly once: no slot left unwritten
    for (uint32_t k = threadIdx.x; k < p.L; k += kEWG)
        if (t.ys[k] == 0xFFFFu) s_bad = 1u;
    __syncthreads();
    return s_bad ? kEncBadTable : 0u;
}

// One symbol with table word q: x (in [L, 2L)) -> new state; *code / *nb the emitted
// bits; *bad set when the symbol has no state.
__device__ __forceinline__ uint32_t enc_step(const EncTab& t, uint32_t L, uint2 q, uint32_t x,
                                             uint32_t* code, uint32_t* nb, uint32_t* bad) {
    const uint32_t f = q.x & 0xFFFFu;
    if (f == 0u) {
        *bad = 1u;
        *code = 0u;
        *nb = 0u;
        return x;
    }
    const uint32_t n = (q.y >> 16) - (x < (q.y & 0xFFFFu) ? 1u : 0u);
    *nb = n;
    *code = x & ((1u << n) - 1u);
    return L + t.ys[(q.x >> 16) + (x >> n) - f];
}

// Symbols of one chunk, read 16 at a time (two u64 in registers, bytes taken by shifts).
struct SymReader {
    const uint8_t* p;
    uint64_t i, e, lo, hi;
    uint32_t k;
    __device__ void init(const uint8_t* base, uint64_t b, uint64_t end) {
        p = base;
        i = b;
        e = end;
        k = 16u;
    }
    __device__ __forceinline__ uint32_t next() {
        if (k == 16u) {
            if (i + 16u <= e && (((uintptr_t)(p + i)) & 15u) == 0u) {
                const uint4 v = *(const uint4*)(p + i);
                lo = (uint64_t)v.x | ((uint64_t)v.y << 32);
                hi = (uint64_t)v.z | ((uint64_t)v.w << 32);
            } else {                       // edge or unaligned: byte loads
                lo = 0u;
                hi = 0u;
                for (uint32_t q = 0; q < 16u && i + q < e; ++q) {
                    const uint64_t byte = p[i + q];
                    if (q < 8u) lo |= byte << (8u * q); else hi |= byte << (8u * (q - 8u));
                }
            }
            k = 0u;
            i += 16u;
        }
        const uint64_t word = k < 8u ? lo : hi;
        const uint32_t s = (uint32_t)(word >> (8u * (k & 7u))) & 0xFFu;
        ++k;
        return s;
    }
};

__device__ __forceinline__ uint64_t chunk_begin(uint32_t c) { return (uint64_t)c * kEC; }

// Symbols staged through LDS for loops in which every lane of a workgroup takes all of
// its chunk's symbols in order (spec and write passes): rounds of kER symbols per lane.
// The 16-byte pieces of all lanes' segments are loaded by consecutive threads (kEP
// threads per 64-byte segment, instead of each lane loading its own 16 bytes at a 4 KiB
// stride), into registers one round ahead, then through LDS to their lane.
constexpr uint32_t kER = 64;            // symbols per lane per round
constexpr uint32_t kEP = kER / 16;      // 16-byte pieces per lane segment
constexpr size_t kStageBytes = (size_t)kEWG * kER;

__device__ __forceinline__ size_t stage_offset(uint32_t L) { return (enc_lds_bytes(L) + 15u) & ~15ull; }

__device__ __forceinline__ uint4 load_piece(const EncPlan& p, uint32_t c, uint64_t o) {
    const uint64_t end = min(p.n, chunk_begin(c) + kEC);
    if (o + 16u <= end && (((uintptr_t)(p.sym + o)) & 15u) == 0u) return *(const uint4*)(p.sym + o);
    uint64_t lo = 0u, hi = 0u;                 // edge or unaligned: byte loads
    for (uint32_t q = 0; q < 16u && o + q < end; ++q) {
        const uint64_t byte = p.sym[o + q];
        if (q < 8u) lo |= byte << (8u * q); else hi |= byte << (8u * (q - 8u));
    }
    return make_uint4((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32));
}

__device__ __forceinline__ void stage_load(const EncPlan& p, uint32_t c0, uint32_t r, uint4* pf) {
#pragma unroll
    for (uint32_t j = 0; j < kEP; ++j) {
        const uint32_t idx = threadIdx.x + kEWG * j;
        const uint32_t c = c0 + idx / kEP;
        pf[j] = c < p.n_chunks ? load_piece(p, c, chunk_begin(c) + (uint64_t)r * kER + 16u * (idx % kEP))
                               : make_uint4(0u, 0u, 0u, 0u);
    }
}

// Runs f(i, symbol) for i = 0 .. n-1 of this lane's chunk (lanes with run == false only
// take part in the loads and barriers).  All threads of the workgroup must call it.
struct NoRoundHook {
    __device__ void operator()(uint32_t, const uint4*) const {}
};

// With kPrev, each lane takes its predecessor lane's chunk instead (lane 0 must not run).
template <bool kPrev = false, typename F, typename G = NoRoundHook>
__device__ void staged_symbols(const EncPlan& p, uint32_t c0, uint4* sbuf, bool run, uint32_t n,
                               F&& f, G&& after_round = G()) {
    const uint64_t left = p.n - chunk_begin(c0);
    const uint32_t rounds = (uint32_t)((min((uint64_t)kEC, left) + kER - 1u) / kER);
    uint4 pf[kEP];
    stage_load(p, c0, 0u, pf);
    for (uint32_t r = 0; r < rounds; ++r) {
        __syncthreads();                       // previous round's reads done
#pragma unroll
        for (uint32_t j = 0; j < kEP; ++j) sbuf[threadIdx.x + kEWG * j] = pf[j];
        __syncthreads();
        if (r + 1u < rounds) stage_load(p, c0, r + 1u, pf);
        if (!run) continue;
#pragma unroll
        for (uint32_t k = 0; k < kEP; ++k) {
            const uint4 v = sbuf[(threadIdx.x - (kPrev ? 1u : 0u)) * kEP + k];
            const uint32_t wd[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (uint32_t b = 0; b < 16u; ++b) {
                const uint32_t i = r * kER + k * 16u + b;
                if (i < n) f(i, (wd[b >> 2] >> (8u * (b & 3u))) & 0xFFu);
            }
        }
        after_round(r, sbuf + threadIdx.x * kEP);
    }
}

// Exit state and bits of chunk c along its current record.
__device__ __forceinline__ uint32_t enc_exit(const EncWork& w, const EncPlan& p, uint32_t c) {
    const uint32_t g = p.chunk_base + c;
    return w.mq[g] == kNoneE ? w.wexit[g] : w.sexit[g];
}
__device__ __forceinline__ uint32_t enc_bits(const EncWork& w, const EncPlan& p, uint32_t c) {
    const uint32_t g = p.chunk_base + c;
    const uint32_t m = w.mq[g];
    if (m == kNoneE) return w.wtot[g];
    return w.wbits[g] + (w.stot[g] - w.ckb[p.ck_base + (uint64_t)c * kECK + m]);
}

// Walk chunk c (c >= 1) from state e until it meets the stored spec state at a
// checkpoint; records entry, and mq / wbits or (no meeting) wexit / wtot.  Every lane
// stores all five fields once, at the end: with the stores in the two exits, the device
// compiler merged them into stores whose base pointers it selected per lane from a
// scratch copy of the work struct, and records came out mixed up.  Aligned symbols are
// taken 64 at a time with the loads kept in flight: a walk is a serial chain, and the
// last fixup rounds run few walks on a nearly idle device.
__device__ void enc_walk(const EncWork& w, const EncPlan& p, const EncTab& t, uint32_t c,
                         uint32_t e, uint32_t* bad) {
    const uint32_t g = p.chunk_base + c;
    const uint64_t b = chunk_begin(c), end = min(p.n, b + kEC);
    const uint64_t q0 = p.ck_base + (uint64_t)c * kECK;
    uint32_t mq = kNoneE, wb = 0u, x = e, bits = 0u;
    if (w.ckx[q0] == e) {
        mq = 0u;
    } else if ((((uintptr_t)p.sym) & 15u) == 0u) {
        // Symbols 16-byte aligned (chunk starts are multiples of 16 too): 64 symbols per
        // iteration from four block slots, each reloaded in place with the block four
        // ahead once taken, and the states of the next even and odd checkpoints read two
        // checkpoints ahead.  Every load is unconditional straight-line code, so the
        // compiler keeps them in flight instead of waiting for each.  Block addresses are
        // clamped to the last block holding a symbol of the chunk; bytes past its end are
        // loaded but never taken as symbols (at most n steps).
        typedef uint32_t v4 __attribute__((ext_vector_type(4)));
        typedef const __attribute__((address_space(1))) v4* gv4p;
        const uint32_t n = (uint32_t)(end - b);
        gv4p qb = (gv4p)(p.sym + b);
        const gv4p last = (gv4p)(((uintptr_t)(p.sym + end) - 1u) & ~(uintptr_t)15);
        v4 s0 = *(qb < last ? qb : last);
        v4 s1 = *(qb + 1 < last ? qb + 1 : last);
        v4 s2 = *(qb + 2 < last ? qb + 2 : last);
        v4 s3 = *(qb + 3 < last ? qb + 3 : last);
        qb += 4;
        uint32_t ckE = 0u, ckO = w.ckx[q0 + 1u];
        uint2 qn = t.sym[s0.x & 0xFFu];
        uint32_t i = 0u;
        // 16 steps from block s (nx: the next block, for the look-ahead at its end);
        // true when the chunk's symbols are all taken
        auto blk = [&](v4& s, const v4& nx) __attribute__((always_inline)) -> bool {
#pragma unroll
            for (uint32_t kk = 0; kk < 16u; ++kk) {
                if (i >= n) return true;
                const uint32_t k1 = kk + 1u;
                const uint32_t w1 = k1 < 4u ? s.x : k1 < 8u ? s.y : k1 < 12u ? s.z : s.w;
                const uint32_t nbyte = kk < 15u ? (w1 >> (8u * (k1 & 3u))) & 0xFFu : nx.x & 0xFFu;
                const uint2 cur = qn;
                qn = t.sym[nbyte];
                uint32_t code, nbits;
                x = enc_step(t, p.L, cur, x, &code, &nbits, bad);
                bits += nbits;
                ++i;
            }
            s = *(qb < last ? qb : last);           // the block four ahead
            ++qb;
            return false;
        };
        for (uint32_t m = 0;; m += 2u) {            // i = 32 m here
            if (m && i < n && ckE == x) {
                mq = m;
                wb = bits;
                break;
            }
            ckE = w.ckx[q0 + min(m + 2u, kECK - 1u)];
            if (blk(s0, s1) || blk(s1, s2)) break;
            if (i < n && ckO == x) {                // i = 32 (m + 1)
                mq = m + 1u;
                wb = bits;
                break;
            }
            ckO = w.ckx[q0 + min(m + 3u, kECK - 1u)];
            if (blk(s2, s3) || blk(s3, s0)) break;
        }
    } else {
        SymReader rd;
        rd.init(p.sym, b, end);
        const uint32_t n = (uint32_t)(end - b);
        uint2 q = t.sym[rd.next()];
        for (uint32_t i = 0; i < n; ++i) {
            if (i && (i % kEK) == 0u && w.ckx[q0 + i / kEK] == x) {
                mq = i / kEK;
                wb = bits;
                break;
            }
            const uint2 cur = q;
            if (i + 1u < n) q = t.sym[rd.next()];      // next symbol's word, ahead
            uint32_t code, nb;
            x = enc_step(t, p.L, cur, x, &code, &nb, bad);
            bits += nb;
        }
    }
    w.entry[g] = e;
    w.mq[g] = mq;
    w.wbits[g] = wb;
    w.wexit[g] = mq == kNoneE ? x : 0u;
    w.wtot[g] = mq == kNoneE ? bits : 0u;
}

__global__ __launch_bounds__(kEWG) void k_enc_sync(EncWork w) {
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t s_job, s_warm;
    if (threadIdx.x == 0) s_job = job_of_task(w, blockIdx.x);
    __syncthreads();
    const EncPlan p = w.plans[__builtin_amdgcn_readfirstlane(s_job)];
    EncTab t;
    const uint32_t st = stage_enc(p, lds, t);
    const uint32_t c0 = (blockIdx.x - p.task_base) * kEWG;
    const uint32_t c = c0 + threadIdx.x;
    const bool valid = c < p.n_chunks;
    uint32_t bad = 0u;
    uint32_t x = p.L;
    // Low-entropy tables (sum of f_s * m_s below 2L: under about two bits per symbol)
    // shed few bits per symbol, so a state guessed at a chunk's start can take more than
    // a chunk to meet the true path.  For them every lane but the range's first encodes
    // its predecessor's chunk first (a warm-up, from state L, nothing stored), so that its
    // spec starts near the true path and fewer walks fail into k_enc_fix.
    if (threadIdx.x < 64u) {
        uint32_t a = 0u;
        for (uint32_t k = 0; k < 4u; ++k) {
            const uint2 q = t.sym[threadIdx.x * 4u + k];
            a += (q.x & 0xFFFFu) * (q.y >> 16);
        }
        for (uint32_t d = 32u; d; d >>= 1) a += __shfl_xor(a, d, 64);
        if (threadIdx.x == 0) s_warm = a < 2u * p.L ? 1u : 0u;
    }
    __syncthreads();
    if (!st) {
        // spec from the start state, every checkpoint stored
        const uint64_t q0 = p.ck_base + (uint64_t)c * kECK;
        const uint32_t n = valid ? (uint32_t)(min(p.n, chunk_begin(c) + kEC) - chunk_begin(c)) : 0u;
        uint32_t bits = 0u;
        uint4* sbuf = (uint4*)((char*)lds + stage_offset(p.L));
        if (s_warm) {                          // uniform; chunk c - 1 is whole for c >= 1
            staged_symbols<true>(p, c0, sbuf, valid && threadIdx.x > 0u, kEC,
                                 [&](uint32_t, uint32_t sy) {
                uint32_t code, nb;
                x = enc_step(t, p.L, t.sym[sy], x, &code, &nb, &bad);
            });
        }
        staged_symbols(p, c0, sbuf, valid, n, [&](uint32_t i, uint32_t sy) {
            if ((i % kEK) == 0u) {
                w.ckx[q0 + i / kEK] = x;
                w.ckb[q0 + i / kEK] = bits;
            }
            uint32_t code, nb;
            x = enc_step(t, p.L, t.sym[sy], x, &code, &nb, &bad);
            bits += nb;
        });
        if (valid) {
            const uint32_t g = p.chunk_base + c;
            w.sexit[g] = x;
            w.stot[g] = bits;
            w.ff[g] = 0u;
        }
    }
    if (valid && !st && c == c0) {
        // chunk 0 starts on the true path; a range's first chunk (a seam) gets an entry
        // that matches no exit, so k_enc_fix walks it (k_enc_walk walks the others)
        const uint32_t g = p.chunk_base + c;
        w.entry[g] = c == 0u ? p.L : kNoneE;
        w.mq[g] = 0u;
        w.wbits[g] = 0u;
        w.wexit[g] = 0u;
        w.wtot[g] = 0u;
    }
    if (bad || (st && threadIdx.x == 0)) atomicOr(&w.res[s_job].status, bad ? kEncNoCode : st);
}

// Per range (grid as k_enc_sync): each chunk but the range's first encodes again from its
// predecessor's spec exit until it meets its own spec at a checkpoint.  A kernel of its
// own, so that the walks get the registers that keep their loads in flight (enc_walk).
__global__ __launch_bounds__(kEWG) void k_enc_walk(EncWork w) {
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t s_job;
    if (threadIdx.x == 0) s_job = job_of_task(w, blockIdx.x);
    __syncthreads();
    const EncPlan p = w.plans[__builtin_amdgcn_readfirstlane(s_job)];
    EncTab t;
    if (stage_enc(p, lds, t)) return;          // uniform; k_enc_sync reported it
    const uint32_t c0 = (blockIdx.x - p.task_base) * kEWG;
    const uint32_t c = c0 + threadIdx.x;
    uint32_t bad = 0u;
    if (c > c0 && c < p.n_chunks) enc_walk(w, p, t, c, w.sexit[p.chunk_base + c - 1u], &bad);
    if (bad) atomicOr(&w.res[s_job].status, kEncNoCode);
}

// Per stream: consistency rounds, scan, results, zeroed shared units.
__global__ __launch_bounds__(kEWG) void k_enc_fix(EncWork w) {
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t s_n, s_n2;
    __shared__ uint64_t s_w[kEWG / 64];
    const uint32_t j = blockIdx.x;
    const EncPlan p = w.plans[j];
    mans_encode_result* r = &w.res[j];
    if (p.n == 0u) {                       // no symbols: one zero unit, state L
        if (threadIdx.x == 0) {
            r->n_units = 1u;
            r->first_bit = 0u;
            r->first_state = p.L;
            if (p.cap < 1u) r->status |= kEncCapacity; else p.units[0] = 0u;
            if (p.items) {                      // item 0: state L, no bits
                mans_sync_point q;
                q.state = 0u;
                q.bit = 32u;
                q.symbols_before = 0u;
                p.items[0] = q;
            }
        }
        return;
    }
    EncTab t;
    const uint32_t st = stage_enc(p, lds, t);
    if (st || r->status) return;           // status set before this kernel (uniform read)
    uint32_t* list = w.fl + p.chunk_base;
    uint32_t* next = w.fl2 + p.chunk_base;
    uint32_t bad = 0u;
    if (threadIdx.x == 0) s_n = 0u;
    __syncthreads();
    for (uint32_t c = 1u + threadIdx.x; c < p.n_chunks; c += kEWG)
        if (w.entry[p.chunk_base + c] != enc_exit(w, p, c - 1u)) list[atomicAdd(&s_n, 1u)] = c;
    __syncthreads();
    // The leftmost listed chunk starts from a proven exit and is consistent after its
    // round, so at most n_chunks rounds run; stop after n_chunks + 1 (or w.round_limit)
    // and report a stream still inconsistent then.
    const uint32_t cap = w.round_limit && w.round_limit <= p.n_chunks ? w.round_limit
                                                                       : p.n_chunks + 1u;
    for (uint32_t round = 0; s_n; ++round) {
        if (round >= cap) {                 // uniform (s_n read after a barrier)
            if (threadIdx.x == 0) r->status |= kEncNoConverge;
            return;
        }
        const uint32_t n = s_n;
        for (uint32_t i = threadIdx.x; i < n; i += kEWG) {
            next[i] = enc_exit(w, p, list[i] - 1u);
            w.ff[p.chunk_base + list[i]] = 1u;      // listed this round
        }
        __syncthreads();                    // entries captured before any walk writes
        for (uint32_t i = threadIdx.x; i < n; i += kEWG) enc_walk(w, p, t, list[i], next[i], &bad);
        __threadfence_block();
        __syncthreads();
        if (threadIdx.x == 0) s_n2 = 0u;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < n; i += kEWG) {
            const uint32_t c = list[i];
            if (w.entry[p.chunk_base + c] != enc_exit(w, p, c - 1u)) next[atomicAdd(&s_n2, 1u)] = c;
            if (c + 1u < p.n_chunks && !w.ff[p.chunk_base + c + 1u] &&
                w.entry[p.chunk_base + c + 1u] != enc_exit(w, p, c))
                next[atomicAdd(&s_n2, 1u)] = c + 1u;       // each chunk at most once
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < n; i += kEWG) w.ff[p.chunk_base + list[i]] = 0u;
        if (threadIdx.x == 0) s_n = s_n2;
        uint32_t* tmp = list; list = next; next = tmp;
        __syncthreads();
    }
    // exclusive scan of bits per chunk
    uint64_t carry = 0u;
    for (uint32_t b0 = 0; b0 < p.n_chunks; b0 += kEWG) {
        const uint32_t c = b0 + threadIdx.x;
        const uint64_t v = c < p.n_chunks ? enc_bits(w, p, c) : 0u;
        uint64_t incl = v;
        const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
        for (uint32_t d = 1; d < 64u; d <<= 1) {
            const uint64_t o = __shfl_up(incl, d, 64);
            if (lane >= d) incl += o;
        }
        if (lane == 63u) s_w[wv] = incl;
        __syncthreads();
        uint64_t add = carry;
        for (uint32_t k = 0; k < wv; ++k) add += s_w[k];
        if (c < p.n_chunks) w.off[p.chunk_base + c] = add + incl - v;
        carry += s_w[0] + s_w[1] + s_w[2] + s_w[3];
        __syncthreads();
    }
    const uint64_t T = carry;
    const uint64_t nu = T / 32u + 1u;
    if (threadIdx.x == 0) {
        r->n_units = nu;
        r->first_bit = (uint32_t)(T % 32u);
        r->first_state = enc_exit(w, p, p.n_chunks - 1u);
        if (nu > p.cap) r->status |= kEncCapacity;
        if (bad) r->status |= kEncNoCode;
    }
    if (nu > p.cap) return;
    // units shared by chunks (and the last one) start at zero: the write pass ORs into them
    for (uint32_t c = threadIdx.x; c < p.n_chunks; c += kEWG) {
        const uint64_t u = w.off[p.chunk_base + c] / 32u;
        p.units[nu - 1u - u] = 0u;
    }
    if (threadIdx.x == 0) p.units[0] = 0u;      // forward unit T / 32
}

__global__ __launch_bounds__(kEWG) void k_enc_write(EncWork w) {
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t s_job;
    if (threadIdx.x == 0) s_job = job_of_task(w, blockIdx.x);
    __syncthreads();
    const EncPlan p = w.plans[__builtin_amdgcn_readfirstlane(s_job)];
    if (w.res[s_job].status) return;            // uniform: written by k_enc_fix
    EncTab t;
    if (stage_enc(p, lds, t)) return;
    const uint32_t c0 = (blockIdx.x - p.task_base) * kEWG;
    const uint32_t c = c0 + threadIdx.x;
    const bool valid = c < p.n_chunks;
    const uint64_t nu = w.res[s_job].n_units;
    const uint32_t n = valid ? (uint32_t)(min(p.n, chunk_begin(c) + kEC) - chunk_begin(c)) : 0u;
    uint32_t x = p.L;
    uint64_t pos = 0u;
    if (valid) {
        x = c == 0u ? p.L : w.entry[p.chunk_base + c];
        pos = w.off[p.chunk_base + c];
    }
    uint64_t u = pos / 32u;                      // forward index of the unit being filled
    uint32_t fill = (uint32_t)(pos % 32u);       // bits of that unit before this chunk
    bool shared = fill != 0u;                    // its leading bits belong to a predecessor
    uint64_t acc = 0u;
    uint32_t bad = 0u;
    // Decoder item records: the decoder reads the forward bits backwards (its position P
    // is forward bit 32 nu - 1 - P), so the step for symbol s starts at P = 32 nu - E_s,
    // E_s the forward bits through symbol s, with the state after symbol s.  Item i
    // (decoder units 32i ..) starts at the last symbol s with P_s >= 1024 i: it is
    // recorded when the next symbol's code takes P below that boundary.
    const uint64_t d0 = 32u * nu - pos;         // P of the chunk's first step
    const uint32_t doff = 1023u - (uint32_t)(d0 & 1023u);
    uint32_t le = 0u;                            // bits of this chunk so far
    const uint64_t sym0 = chunk_begin(c);
    uint4* sbuf = (uint4*)((char*)lds + stage_offset(p.L));
    // a round (kER symbols of at most 14 bits) crosses at most one boundary (1024 bits),
    // whose record is written after the round
    static_assert(kER * 14u < 1024u, "one item boundary per round");
    // P = d0 - le crosses a multiple of 1024 exactly when (le + doff) / 1024 changes.  The
    // symbol loop only flags a crossing; the round is then replayed from its start state
    // (rx, rle) out of LDS to find the crossing step (keeping that step's values in the
    // unrolled loop took some 70 more registers and scratch).
    uint32_t crossed = 0u, rx = x, rle = 0u;
    staged_symbols(p, c0, sbuf, valid, n, [&](uint32_t, uint32_t sy) {
        uint32_t code, nb;
        x = enc_step(t, p.L, t.sym[sy], x, &code, &nb, &bad);
        const uint32_t lnew = le + nb;
        crossed |= ((le + doff) ^ (lnew + doff)) >> 10;
        le = lnew;
        acc = (acc << nb) | code;
        fill += nb;
        if (fill >= 32u) {
            fill -= 32u;
            const uint32_t word = (uint32_t)(acc >> fill);
            acc &= (1ull << fill) - 1u;
            uint32_t* dst = p.units + (nu - 1u - u);
            if (u < nu) {                        // always, when the records are right
                if (shared) atomicOr(dst, word); else *dst = word;
            }
            shared = false;
            ++u;
        }
    }, [&](uint32_t r, const uint4* mine) {
        if (crossed && p.items) {
            // replay this round to the step that crosses: item k = P / 1024 starts at the
            // step before it (symbol s - 1, in state cx); the stream's first symbol
            // leaves a trailing item without steps; k = nu / 32 (the stream's end) is no
            // item
            uint32_t xx = rx, ll = rle, dummy = 0u;
            for (uint32_t j = 0; j < kER; ++j) {
                const uint32_t i = r * kER + j;
                if (i >= n) break;
                const uint4 v = mine[j >> 4];
                const uint32_t q = j & 15u;
                const uint32_t wdq = q < 8u ? (q < 4u ? v.x : v.y) : (q < 12u ? v.z : v.w);
                const uint32_t sy = (wdq >> (8u * (q & 3u))) & 0xFFu;
                uint32_t code, nb;
                const uint32_t xn = enc_step(t, p.L, t.sym[sy], xx, &code, &nb, &dummy);
                if ((((ll + doff) ^ (ll + nb + doff)) >> 10) != 0u) {
                    const uint64_t cd = d0 - ll;     // P before the crossing step
                    const uint64_t k = cd >> 10;
                    if (k < (nu + 31u) / 32u) {
                        const uint64_t sg = sym0 + i;
                        mans_sync_point rec;
                        rec.state = sg ? xx - p.L : 0u;
                        rec.bit = sg ? (uint32_t)(cd - (k << 10)) : 0u;
                        rec.symbols_before = p.n - sg;
                        p.items[k] = rec;
                    }
                    break;
                }
                xx = xn;
                ll += nb;
            }
        }
        crossed = 0u;
        rx = x;
        rle = le;
    });
    if (valid && fill) {
        // the unit this chunk ends in is shared with the next chunk (or is the last one):
        // acc holds this chunk's bits of it, placed below any predecessor's bits
        const uint32_t word = (uint32_t)(acc << (32u - fill));
        if (u < nu) atomicOr(p.units + (nu - 1u - u), word);
    }
    if (valid && p.items && c + 1u == p.n_chunks) {
        // item 0: the decoder's start (the state after the last symbol, first_bit bits)
        mans_sync_point r;
        r.state = x - p.L;
        r.bit = (uint32_t)(d0 - le);            // 32 - first_bit (32 when first_bit = 0)
        r.symbols_before = 0u;
        p.items[0] = r;
    }
}

uint64_t align256(uint64_t v) { return (v + 255u) & ~255ull; }

struct EncLayout {
    uint64_t plans, res, chunk[10], off, ckx, ckb, bytes;
};

EncLayout enc_layout(size_t n_jobs, uint64_t chunks, uint64_t cks) {
    EncLayout l;
    uint64_t o = 0;
    auto take = [&](uint64_t b) { uint64_t r = o; o = align256(o + b); return r; };
    l.plans = take(sizeof(EncPlan) * n_jobs);
    l.res = take(sizeof(mans_encode_result) * n_jobs);
    for (int i = 0; i < 10; ++i) l.chunk[i] = take(4 * chunks + 4);
    l.off = take(8 * chunks + 8);
    l.ckx = take(4 * cks + 4);
    l.ckb = take(4 * cks + 4);
    l.bytes = o;
    return l;
}

int enc_plans(const mans_encode_job* jobs, size_t n, std::vector<EncPlan>& plans,
              uint64_t* chunks, uint64_t* cks, uint32_t* tasks, uint32_t* maxL) {
    if (!jobs && n) return set_error(MANS_EINVAL, "jobs is NULL");
    if (n == 0 || n > 0xFFFFFFu) return set_error(MANS_EINVAL, "%zu jobs", n);
    plans.resize(n);
    uint64_t cb = 0, qb = 0, tb = 0;
    uint32_t ml = 2;
    for (size_t i = 0; i < n; ++i) {
        const mans_encode_job& jb = jobs[i];
        if (jb.L < 2u || jb.L > 16384u || (jb.L & (jb.L - 1u)))
            return set_error(MANS_EINVAL, "job %zu: L=%u is not a power of two in [2, 16384]", i, jb.L);
        if (!jb.table || !jb.units || (jb.n && !jb.symbols))
            return set_error(MANS_EINVAL, "job %zu: NULL pointer", i);
        if (jb.n >= (1ull << 42)) return set_error(MANS_EINVAL, "job %zu: n too large", i);
        EncPlan& p = plans[i];
        std::memset(&p, 0, sizeof(p));
        p.sym = (const uint8_t*)jb.symbols;
        p.table = jb.table;
        p.units = jb.units;
        p.items = jb.items_out;
        if (((uintptr_t)jb.items_out) & 7u)
            return set_error(MANS_EINVAL, "job %zu: items_out must be 8-byte aligned", i);
        p.n = jb.n;
        p.cap = jb.units_capacity;
        p.L = jb.L;
        uint32_t R = 0;
        while ((1u << R) < jb.L) ++R;
        p.R = R;
        const uint64_t nc = (jb.n + kEC - 1u) / kEC;
        p.n_chunks = (uint32_t)nc;
        p.chunk_base = (uint32_t)cb;
        p.ck_base = qb;
        p.task_base = (uint32_t)tb;
        cb += std::max<uint64_t>(nc, 1u);
        qb += std::max<uint64_t>(nc, 1u) * kECK;
        tb += std::max<uint64_t>((nc + kEWG - 1u) / kEWG, 1u);
        ml = std::max(ml, jb.L);
    }
    if (cb >= 0xFFFFFFF0ull || tb >= 0x7FFFFFFFull) return set_error(MANS_EINVAL, "too many chunks");
    *chunks = cb;
    *cks = qb;
    *tasks = (uint32_t)tb;
    *maxL = ml;
    return MANS_OK;
}

EncWork enc_work(void* ws, const EncLayout& l, size_t n_jobs, uint32_t tasks) {
    char* b = (char*)ws;
    EncWork w;
    w.plans = (const EncPlan*)(b + l.plans);
    w.n_jobs = (uint32_t)n_jobs;
    w.n_tasks = tasks;
    uint32_t** pc[10] = {&w.entry, &w.mq, &w.wbits, &w.wexit, &w.wtot, &w.sexit, &w.stot, &w.fl,
                         &w.fl2, &w.ff};
    for (int i = 0; i < 10; ++i) *pc[i] = (uint32_t*)(b + l.chunk[i]);
    w.off = (uint64_t*)(b + l.off);
    w.ckx = (uint32_t*)(b + l.ckx);
    w.ckb = (uint32_t*)(b + l.ckb);
    w.res = (mans_encode_result*)(b + l.res);
    w.round_limit = 0u;
    return w;
}

}  // namespace

extern "C" int mans_encode_workspace_bytes(const mans_encode_job* jobs, size_t n_jobs,
                                           uint64_t* bytes) {
    if (!bytes) return set_error(MANS_EINVAL, "bytes is NULL");
    std::vector<EncPlan> plans;
    uint64_t chunks, cks;
    uint32_t tasks, ml;
    int rc = enc_plans(jobs, n_jobs, plans, &chunks, &cks, &tasks, &ml);
    if (rc) return rc;
    *bytes = enc_layout(n_jobs, chunks, cks).bytes;
    return MANS_OK;
}

extern "C" int mans_encode_device(const mans_encode_job* jobs, size_t n_jobs, void* workspace,
                                  uint64_t workspace_bytes, void* stream) {
    std::vector<EncPlan> plans;
    uint64_t chunks, cks;
    uint32_t tasks, ml;
    int rc = enc_plans(jobs, n_jobs, plans, &chunks, &cks, &tasks, &ml);
    if (rc) return rc;
    const EncLayout l = enc_layout(n_jobs, chunks, cks);
    if (!workspace || workspace_bytes < l.bytes)
        return set_error(MANS_ENOMEM, "workspace of %llu bytes < %llu needed",
                         (unsigned long long)workspace_bytes, (unsigned long long)l.bytes);
    hipStream_t st = (hipStream_t)stream;
    EncWork w = enc_work(workspace, l, n_jobs, tasks);
    w.round_limit = sync_round_limit();
    // plans, and zeroed results (status words are ORed into)
    std::vector<char> img(l.chunk[0] - l.plans, 0);
    std::memcpy(img.data(), plans.data(), sizeof(EncPlan) * n_jobs);
    hipError_t e = hipMemcpyAsync((char*)workspace + l.plans, img.data(), img.size(),
                                  hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);    // the host image is freed on return
    if (e != hipSuccess) return set_error(MANS_EHIP, "plan upload: %s", hipGetErrorString(e));
    const size_t lds = enc_lds_bytes(ml);
    const size_t lds_staged = ((lds + 15u) & ~(size_t)15u) + kStageBytes;
    hipLaunchKernelGGL(k_enc_sync, dim3(tasks), dim3(kEWG), lds_staged, st, w);
    hipLaunchKernelGGL(k_enc_walk, dim3(tasks), dim3(kEWG), lds, st, w);
    hipLaunchKernelGGL(k_enc_fix, dim3((uint32_t)n_jobs), dim3(kEWG), lds, st, w);
    hipLaunchKernelGGL(k_enc_write, dim3(tasks), dim3(kEWG), lds_staged, st, w);
    e = hipGetLastError();
    if (e != hipSuccess) return set_error(MANS_EHIP, "encode launch: %s", hipGetErrorString(e));
    return MANS_OK;
}

extern "C" int mans_encode_results(const void* workspace, const mans_encode_job* jobs,
                                   size_t n_jobs, void* stream, mans_encode_result* out) {
    if (!workspace || !out) return set_error(MANS_EINVAL, "NULL argument");
    std::vector<EncPlan> plans;
    uint64_t chunks, cks;
    uint32_t tasks, ml;
    int rc = enc_plans(jobs, n_jobs, plans, &chunks, &cks, &tasks, &ml);
    if (rc) return rc;
    const EncLayout l = enc_layout(n_jobs, chunks, cks);
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = hipMemcpyAsync(out, (const char*)workspace + l.res,
                                  sizeof(mans_encode_result) * n_jobs, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return set_error(MANS_EHIP, "results: %s", hipGetErrorString(e));
    return MANS_OK;
}
