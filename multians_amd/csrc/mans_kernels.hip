// mans_kernels.hip -- gfx950 kernels of the tANS subsequence decoder and the
// decode entry points of the C ABI (mans_decode, mans_decode_batch).
//
// Replaces cuhd::CUHDGPUDecoder::decode and its kernels
// (src/multians_gpu_decoder.cu:195-687 of the reference):
//   phase1_decode_subseq           -> k_sync   (speculative decode + checkpoints, then
//                                     merge walks from each chunk's predecessor exit)
//   phase2_synchronise_blocks+loop -> k_check + k_fixup (consistency check, fixup
//                                     rounds; on device, no host round trip)
//   phase3 copies + thrust scan    -> k_scan (u64 exclusive scan, single pass with
//                                     decoupled look-back)
//   phase4_decode_write_output     -> k_write  (count-driven decode + store)
// The per-lane logic lives in mans_core.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "mans_core.h"
#include "mans_internal.h"
#include "mans_lineage.h"

using namespace mans;

namespace {

constexpr int kWG = 256;           // lanes per workgroup (4 waves)
constexpr uint32_t kRepairPasses = 4;   // in-range repair passes after the walks (k_sync)
// Write passes a k_write workgroup takes per grab of the pass counter (blocks of
// consecutive passes, usually of one stream; 2 measured 0.8% faster than 4 on the mixed
// batch, 1 and 8 slower)
constexpr uint32_t kWritePer = 2;

// LDS words by byte address.  The staged table's words carry LDS dword indices (table
// start + next-state base), so a decode step forms the address of the next table word
// with one add and one shift.
typedef const uint32_t __attribute__((address_space(3)))* lds_cptr;

__device__ __forceinline__ uint32_t lds_addr(const void* p) {
    return (uint32_t)(size_t)(const __attribute__((address_space(3))) void*)p;
}

__device__ __forceinline__ uint32_t lds_word(uint32_t byte_addr) {
    return *(lds_cptr)(size_t)byte_addr;
}

// The 32-bit window at LDS bit address pb: the two words from the one holding bit pb,
// shifted by pb mod 32 (alignbit uses the low five bits of its shift).
__device__ __forceinline__ void lds_window_load(uint32_t pb, uint32_t& lo, uint32_t& hi) {
    const uint32_t a = (pb >> 3) & ~3u;
    lo = lds_word(a);
    hi = lds_word(a + 4u);
}

// A staged table: word i (state i) sits at LDS dword tb + i and holds nb (bits 0-4),
// the symbol (bits 8-15) and tb + base (bits 16-31): state i steps to tb + base + bits.
struct LdsTab {
    uint32_t tb;
};

// First LDS word after a staged table of L words, 16-byte aligned (the unit buffers
// after it take 16- and 8-byte stores).
__host__ __device__ __forceinline__ uint32_t table_words(uint32_t L) {
    return (L + 3u) & ~3u;
}

struct KArgs {
    StreamPlan one;                 // used when plans == nullptr
    const StreamPlan* plans;        // batch
    const uint32_t* order;          // batch: chunk tasks, longest expected first
    const uint32_t* sorder;         // batch: streams, longest expected first
    uint32_t n_streams;
    uint32_t n_tasks;
    uint32_t wait_polls;             // k_sync: polls of a range's wait for its predecessor
    uint32_t round_limit;            // k_fixup: rounds per stream (0 = n_chunks + 1, the bound)
    uint32_t n_scan_blocks;          // k_scan blocks (descriptors k_init clears)
    uint32_t any_index_out;          // some stream asks for its sync points (k_index runs)
    uint32_t direct;                 // trusted item records: the write pass alone
    Work w;
    // optional: the reference decoder's sync state (mans_ref_sync), single streams only
    uint32_t* ref_sync;
    uint32_t* ref_sizes;
    uint8_t* ref_flags;
    uint32_t ref_S, ref_nsub, ref_nblocks;
    // optional (lineage decodes): k_scan takes its per-chunk counts from lin_count
    const uint64_t* lin_count;
    // optional: the kernel runs only when *gate == gate_val (a device word written by an
    // earlier kernel of the same sequence, so the choice needs no host round trip)
    const uint32_t* gate;
    uint32_t gate_val;
    // optional (k_refsync after a lineage decode): the decode's 32-unit item records, used
    // when its lineage failure word (lin_fail_word) is 0
    const SyncPoint* lin_rec;
    const uint32_t* lin_fail_word;
    uint32_t lin_nq;
    // chunks per sync task (the sync kernels' workgroup size): 0 = kWG
    uint32_t range;
};

__host__ __device__ __forceinline__ uint32_t range_of(const KArgs& a) {
    return a.range ? a.range : (uint32_t)kWG;
}

// True when this launch is switched off by its gate word (uniform over the grid).
__device__ __forceinline__ bool gated_off(const KArgs& a) {
    return a.gate &&
           __hip_atomic_load(a.gate, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != a.gate_val;
}

// Chunk task of workgroup b: batches run their tasks longest-expected first, so the
// long ones do not start last and hold the end of the kernel.
__device__ __forceinline__ uint32_t chunk_task(const KArgs& a, uint32_t b) {
    return a.order ? a.order[b] : b;
}

__device__ __forceinline__ const StreamPlan* plan_of(const KArgs& a, uint32_t s) {
    return a.plans ? &a.plans[s] : &a.one;
}

// Stream owning task t (tasks of each stream are contiguous, ascending in s).
template <bool kItems>
__device__ uint32_t stream_of_task(const KArgs& a, uint32_t t) {
    if (!a.plans) return 0u;
    uint32_t lo = 0, hi = a.n_streams - 1u;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1u) >> 1;
        uint32_t b = kItems ? a.plans[mid].task_item : a.plans[mid].task_chunk;
        if (b <= t) lo = mid; else hi = mid - 1u;
    }
    return lo;
}

__device__ uint32_t stream_of_chunk(const KArgs& a, uint32_t gc) {
    if (!a.plans) return 0u;
    uint32_t lo = 0, hi = a.n_streams - 1u;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1u) >> 1;
        if (a.plans[mid].chunk_base <= gc) lo = mid; else hi = mid - 1u;
    }
    return lo;
}

// Stage the derived table of stream sp into LDS.  Returns (uniform) flags: bit 0 when
// the table has zero-bit cycles (see derive_entry), bit 1 when it has rejected entries
// (they stage as zero, a zero-bit self-loop a round would never leave, so nothing of
// that stream is decoded); rejected tables set status[0] bit 0.
// With count_only, the symbol byte (bits 8-15) is staged as zero: the sync kernels keep
// the length of a folded zero-bit run there instead (fold_zero_bit_runs).
__device__ uint32_t stage_table(const StreamPlan& sp, uint32_t* lds, uint32_t* status,
                                bool count_only = false) {
    __shared__ uint32_t s_flags;
    if (threadIdx.x == 0) s_flags = 0u;
    __syncthreads();
    int bad = 0, cyc = 0;
    // tb + base < 2^16: tb <= 160 KiB / 4 and base < L <= 16384
    const uint32_t tb = lds_addr(lds) >> 2;
    // kG entries per thread per group, all loads of a group issued before any is used
    // (one memory latency per group instead of one per entry)
    constexpr uint32_t kG = 16;
    for (uint32_t i0 = threadIdx.x; i0 < sp.L; i0 += blockDim.x * kG) {
        uint32_t e[kG];
#pragma unroll
        for (uint32_t k = 0; k < kG; ++k) e[k] = load_unit(sp.table, sp.L, i0 + k * blockDim.x);
#pragma unroll
        for (uint32_t k = 0; k < kG; ++k) {
            const uint32_t i = i0 + k * blockDim.x;
            if (i < sp.L) {
                const uint32_t d = derive_entry(e[k], sp.L, i, &bad, &cyc);
                lds[i] = (count_only ? (d & ~0xFF00u) : d) + (tb << 16);
            }
        }
    }
    if (bad) {
        atomicOr(&status[0], 1u);
        if (sp.status_out) atomicOr(sp.status_out, 1u);
        atomicOr(&s_flags, 2u);
    }
    if (cyc) atomicOr(&s_flags, 1u);
    __syncthreads();
    const uint32_t f = s_flags;
    if ((f & 1u) && sp.n_ck > 1u && threadIdx.x == 0) {
        atomicOr(&status[0], 1u);
        if (sp.status_out) atomicOr(sp.status_out, 1u);
    }
    return __builtin_amdgcn_readfirstlane(f);
}

// Round geometry (see the round engine below).
#ifndef MANS_ROUND_UNITS
#define MANS_ROUND_UNITS 16
#endif
constexpr uint32_t kR = MANS_ROUND_UNITS;  // units decoded per round (= checkpoint stride)
constexpr uint32_t kRL = kR + 4;   // units staged per lane per round (kR + lookahead)
// LDS stride per lane in units: the round plus the two units its last window can reach
// (p stays below (kR + 1) * 32 after the last pair, see round_decode); the rest of the
// last 16-byte load is not stored.  18 is 2 mod 4: lanes at the same offset of their segments
// (lockstep rounds keep them close) spread over twice the banks a 16-byte multiple would
// give, the segments take 8-byte halves (ds_write2_b64), and four sync workgroups fit a
// CU's LDS.
constexpr uint32_t kLS = kR + 2u;
static_assert(kLS <= kRL && kLS % 4u == 2u, "round segment stride");
constexpr uint32_t kPieces = kRL / 4;   // 16-byte loads per lane segment
static_assert(kR % 4 == 0 && kR >= 8, "round length");

// Zero-bit runs folded into the staged table, for the kernels that count steps without
// emitting symbols (k_sync, k_fixup).  A zero-bit step reads no bits, so from a state
// whose step reads none the next states are fixed until one whose step reads bits: the
// word of state i becomes that state j's word, with the number z of zero-bit steps from i
// to j (at most kMaxRun) in bits 8-15, where the emitted symbol was.  One lookup then
// takes 1 + z steps, all starting at the same bit; a lane's state is still the state
// before its next step, so checkpoints (the entry of a unit boundary) are unchanged.
// The runs come from pointer jumping in LDS over `scratch` (L words: next | z << 16;
// each word is read and written whole, so jumps taken in any order stay consistent):
// about log2(run length) passes.  Tables with zero-bit cycles (flag bit 0) or rejected
// entries keep their plain words; so do tables of more than 4096 states.  Tables larger
// than `cap_words` (the scratch) fold in place instead (below).  The table must be staged
// count_only (symbol bytes zero).
constexpr uint32_t kMaxRun = 255;

// The same folding in place, for workgroups whose round buffer cannot hold L words: the
// zero-bit words become pointer words (bits 16-31: LDS index of the state reached, 8-15:
// zero-bit steps z to it, nb 0) that jump along their runs; words whose step reads bits
// are never written, and a zero-bit word is read whole in either form, so jumps taken in
// any order stay consistent.  Then a word whose run ends on a state that reads bits takes
// that state's word with z (1 + z steps per lookup, as above); a run stopped by kMaxRun
// keeps its pointer with z - 1 (z - 1 zero-bit steps, then one more to the state it
// points at).  Bit 7 marks the latter between two barriers.
template <int W>
__device__ void fold_zero_bit_runs_in_place(uint32_t* lds, uint32_t L) {
    const uint32_t tb = lds_addr(lds) >> 2;
    uint32_t any = 0u;
    for (uint32_t i = threadIdx.x; i < L; i += W) {
        const uint32_t w = lds[i];
        if ((w & 31u) == 0u) {
            lds[i] = (w & 0xFFFF0000u) | (1u << 8);
            any = 1u;
        }
    }
    if (!__syncthreads_or(any)) return;
    for (;;) {
        uint32_t changed = 0u;
        for (uint32_t i = threadIdx.x; i < L; i += W) {
            const uint32_t w = lds[i];
            if (w & 31u) continue;                  // a state whose step reads bits
            const uint32_t q = lds[(w >> 16) - tb];
            if (q & 31u) continue;                  // resolved
            const uint32_t z = ((w >> 8) & 0xFFu) + ((q >> 8) & 0xFFu);
            if (z > kMaxRun) continue;              // at the run cap
            lds[i] = (q & 0xFFFF0000u) | (z << 8);
            changed = 1u;
        }
        if (!__syncthreads_or(changed)) break;
    }
    for (uint32_t i = threadIdx.x; i < L; i += W) {
        const uint32_t w = lds[i];
        if ((w & 31u) == 0u && (lds[(w >> 16) - tb] & 31u) == 0u) lds[i] = w | 0x80u;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < L; i += W) {
        const uint32_t w = lds[i];
        if (w & 31u) continue;
        lds[i] = (w & 0x80u) ? (w & ~0x80u) - (1u << 8) : lds[(w >> 16) - tb] | (w & 0xFF00u);
    }
    __syncthreads();
}

template <int W>
__device__ void fold_zero_bit_runs(uint32_t* lds, uint32_t L, uint32_t* scratch,
                                   uint32_t cap_words) {
    if (L > 4096u) return;                          // uniform
    if (L > cap_words) {
        fold_zero_bit_runs_in_place<W>(lds, L);
        return;
    }
    const uint32_t tb = lds_addr(lds) >> 2;
    uint32_t any = 0u;
    for (uint32_t i = threadIdx.x; i < L; i += blockDim.x) {
        const uint32_t w = lds[i];
        const uint32_t zb = (w & 31u) == 0u ? 1u : 0u;
        scratch[i] = zb ? (((w >> 16) - tb) | (1u << 16)) : i;
        any |= zb;
    }
    if (!__syncthreads_or(any)) return;
    for (;;) {
        uint32_t changed = 0u;
        for (uint32_t i = threadIdx.x; i < L; i += blockDim.x) {
            const uint32_t p = scratch[i];
            const uint32_t z = p >> 16;
            if (!z) continue;                       // a state whose step reads bits
            const uint32_t q = scratch[p & 0xFFFFu];
            const uint32_t zq = q >> 16;
            if (!zq || z + zq > kMaxRun) continue;  // resolved, or at the run cap
            scratch[i] = (q & 0xFFFFu) | ((z + zq) << 16);
            changed = 1u;
        }
        if (!__syncthreads_or(changed)) break;
    }
    // the folded words (read all, barrier, write all: lds[] is read by other threads);
    // L <= min(cap_words, 4096) here, cap_words <= W * kLS, so each of the W threads holds
    // at most kPer of them
    constexpr uint32_t kCap = (uint32_t)W * kLS < 4096u ? (uint32_t)W * kLS : 4096u;
    constexpr uint32_t kPer = (kCap + W - 1u) / W;
    uint32_t nw[kPer];
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        const uint32_t i = threadIdx.x + k * W;
        nw[k] = 0u;
        if (i < L) {
            const uint32_t p = scratch[i];
            const uint32_t z = p >> 16;
            nw[k] = z ? (lds[p & 0xFFFFu] | (z << 8)) : lds[i];   // symbol bytes are 0
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        const uint32_t i = threadIdx.x + k * W;
        if (i < L) lds[i] = nw[k];
    }
    __syncthreads();
}

// Clears the status words, tickets and fixup counters, and marks the last exit of each
// range that another range waits for (k_sync) as not yet published.
__global__ void k_init(KArgs a, uint32_t n_status) {
    if (gated_off(a)) return;
    const Work& w = a.w;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_status) w.status[i] = 0u;
    if (i < 4u) w.ctick[i] = 0u;
    if (i < a.n_scan_blocks) w.bsum[i] = 0ull;          // k_scan's look-back descriptors
    if (i < a.n_streams) {
        w.fcnt[i] = 0u;
        uint32_t* so = plan_of(a, i)->status_out;
        if (so) *so = 0u;
    }
    if (i < a.n_tasks) {
        const StreamPlan& sp = *plan_of(a, stream_of_task<false>(a, i));
        const uint32_t c0 = (i - sp.task_chunk) * range_of(a);
        if (c0 > 0u && c0 < sp.n_chunks) w.exitv[sp.chunk_base + c0 - 1u] = kNone;
    }
}

// ------------------------------------------------------------ round engine
// Spec and walk decodes run in lockstep rounds per workgroup: in round r, lane l
// decodes the steps starting in units [rs_l, rs_l + kR) of its region, rs_l =
// rbase_l + r*kR, reading its units from an LDS buffer.  The units of round r+1
// (kRL per lane, one 16-byte load per 4 units) are loaded into registers before
// round r is decoded and written to LDS after it, so global-memory latency overlaps
// a whole round of decoding.  A round is exactly one checkpoint interval (kR = K),
// so checkpoints and merge tests happen between rounds, outside the step loop.

// The round engine is instantiated per workgroup size W (lanes = threads).
struct RoundIO {
    uint32_t* buf;        // LDS, W * kLS units
    uint32_t* rbase;      // LDS, W: first unit of round 0 per lane (kNone = no load)
    const uint32_t* units;
    uint32_t n;
    uint32_t lim;         // lanes [lim, W) never decode: their segments are not staged
};

struct Prefetch { uint4 v[kPieces]; };

// Piece j of thread t: the (lane, piece) pairs of all lane segments are numbered
// lane-major, kPieces per lane, and thread t takes numbers t, t + W, ...; consecutive
// threads read consecutive 16-byte pieces of a segment.
template <int W>
__device__ __forceinline__ void piece_of(uint32_t j, uint32_t* lane, uint32_t* piece) {
    const uint32_t idx = threadIdx.x + W * j;
    *lane = idx / kPieces;
    *piece = idx - *lane * kPieces;
}

// issue the loads of round r for all lanes
template <int W>
__device__ __forceinline__ void round_load(const RoundIO& io, uint32_t r, Prefetch& pf) {
#pragma unroll
    for (uint32_t j = 0; j < kPieces; ++j) {
        uint32_t lane, piece;
        piece_of<W>(j, &lane, &piece);
        if (lane >= io.lim) continue;
        // lanes without a load (rbase kNone) read an in-stream block they never use;
        // no branch and no mask, so the loads stay in flight while the round decodes
        const uint32_t rb = io.rbase[lane];
        const U4 u = load_block_raw(io.units, io.n, rb != kNone ? rb + r * kR + 4u * piece : 0u);
        pf.v[j] = make_uint4(u.x, u.y, u.z, u.w);
    }
}

template <int W>
__device__ __forceinline__ void round_store(const RoundIO& io, const Prefetch& pf) {
#pragma unroll
    for (uint32_t j = 0; j < kPieces; ++j) {
        uint32_t lane, piece;
        piece_of<W>(j, &lane, &piece);
        if (lane >= io.lim) continue;
        uint2* d = reinterpret_cast<uint2*>(io.buf + lane * kLS + 4u * piece);
        if (4u * piece < kLS) d[0] = make_uint2(pf.v[j].x, pf.v[j].y);
        if (4u * piece + 2u < kLS) d[1] = make_uint2(pf.v[j].z, pf.v[j].w);
    }
}

// Per-lane decode state.
struct RLane {
    uint32_t u, o, xi, cnt, u_stop, rbase;
    uint32_t active;
};

// Decode the lane's steps of round r (those starting in units [rs, min(rs+kR, u_stop))).
// The lane's position is a bit offset p into its staged segment; each iteration reads
// the 32-bit window at p (two LDS words, one ds_read2) and decodes two steps from it:
// a step takes at most 15 bits, so the second step's bits (offset nb1, width nb2) lie in
// the same window.  The second step runs whether or not the first one ended the round;
// if it did, the first step's state is kept (both table words and the window at the
// next position are in range: p stays below (kR + 1) * 32).
__device__ __forceinline__ void round_decode(const LdsTab& tab, const uint32_t* lbuf, RLane& L,
                                             uint32_t r) {
    if (!L.active) return;
    const uint32_t rs = L.rbase + r * kR;
    const uint32_t re = min(rs + kR, L.u_stop);
    if (L.u >= re) return;
    // p: LDS bit address (the segment start is a multiple of 32)
    const uint32_t pb0 = lds_addr(lbuf) * 8u;
    const uint32_t end = pb0 + (re - rs) * 32u;
    uint32_t p = pb0 + (L.u - rs) * 32u + L.o;
    uint32_t lo, hi;
    lds_window_load(p, lo, hi);
    uint32_t i = L.xi + tab.tb;
    uint32_t w = lds_word(i << 2);
    uint32_t cnt = L.cnt, i1, p1, z2;
    do {
        const uint32_t win = funnel_shr(hi, lo, p);
        i1 = (w >> 16) + low_bits(win, w);
        const uint32_t w1 = lds_word(i1 << 2);
        p1 = p + (w & 0xFFu);
        i = (w1 >> 16) + __builtin_amdgcn_ubfe(win, w, w1);   // offset nb1, width nb2
        // each lookup takes 1 + z steps (z: folded zero-bit run, bits 8-15)
        cnt += 2u + __builtin_amdgcn_ubfe(w, 8u, 8u);
        z2 = __builtin_amdgcn_ubfe(w1, 8u, 8u);
        cnt += z2;
        w = lds_word(i << 2);
        p = p1 + (w1 & 0xFFu);
        lds_window_load(p, lo, hi);
    } while (p < end);
    if (p1 >= end) {                        // the round ended after the first lookup
        i = i1;
        p = p1;
        cnt -= 1u + z2;
    }
    p -= pb0;
    L.u = rs + (p >> 5); L.o = p & 31u; L.xi = i - tab.tb; L.cnt = cnt;
}

// Run rounds (at most nr_max, uniform) until no lane is active.  After each round,
// at_end(L, re) is called for lanes that reached a round boundary re < u_stop (their
// state is the entry of unit re); it may deactivate the lane.  at_stop(L) is called in
// the round in which a lane reaches the end of its region.  A lane that stops has its
// unit loads switched off from the next round on.
struct NoStop {
    __device__ void operator()(RLane&) const {}
};

template <int W, typename F, typename G = NoStop>
__device__ void run_rounds(const LdsTab& tab, RoundIO& io, RLane& L, uint32_t nr_max, F&& at_end,
                           G&& at_stop = G()) {
    __syncthreads();                       // rbase[] visible
    Prefetch pf;
    round_load<W>(io, 0u, pf);
    for (uint32_t r = 0; r < nr_max; ++r) {
        __syncthreads();                   // previous round's reads of buf done
        round_store<W>(io, pf);
        __syncthreads();
        if (r + 1u < nr_max) round_load<W>(io, r + 1u, pf);
        round_decode(tab, io.buf + threadIdx.x * kLS, L, r);
        if (L.active) {
            const uint32_t re = L.rbase + (r + 1u) * kR;
            if (re < L.u_stop) {
                if (L.u == re) at_end(L, re);
            } else {
                L.active = 0u;             // region done
                at_stop(L);
            }
            if (!L.active) io.rbase[threadIdx.x] = kNone;
        }
        if (!__syncthreads_or(L.active)) break;
    }
}

__device__ __forceinline__ uint32_t ld_coherent(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// mans::chunk_exit for records other lanes of this kernel have just written.
__device__ __forceinline__ uint32_t chunk_exit_now(const StreamPlan& sp, const Work& w, uint32_t c) {
    const uint32_t gc = sp.chunk_base + c;
    return ld_coherent(&w.mq[gc]) == kNone ? ld_coherent(&w.wexit[gc]) : ld_coherent(&w.exitv[gc]);
}

// Walk into chunk c (c >= 1) of sp from `entry` for the lanes with want != 0, in rounds.
// Same semantics as mans::lane_walk with K == kR.  Returns 1 if merged.
__device__ uint32_t walk_rounds(const StreamPlan& sp, const Work& w, const LdsTab& tab,
                                RoundIO& io, uint32_t want, uint32_t c, uint32_t entry) {
    const uint32_t gc = sp.chunk_base + c;
    RLane L;
    L.active = 0u;
    uint32_t merged = 0u, late = 0u, mq = 0u, wcm = 0u;
    if (want) {
        const uint32_t u0 = c * sp.S;
        L.u_stop = min(u0 + sp.S, sp.n_units);
        L.rbase = u0;
        w.wentry[gc] = entry;
        const uint32_t q0 = u0 >> sp.log2K;
        if (ld_coherent(&w.ck[ck_slot(sp, q0)]) == entry) {
            w.mq[gc] = q0;
            w.wcm[gc] = 0u;
            merged = 1u;
        } else {
            w.wb[ck_slot(sp, q0)] = entry;
            w.wbc[ck_slot(sp, q0)] = 0u;
            L.active = 1u;
            L.xi = entry_state(entry);
            L.u = u0 + (entry_off(entry) >> 5);
            L.o = entry_off(entry) & 31u;
            L.cnt = 0u;
        }
    }
    io.rbase[threadIdx.x] = L.active ? L.rbase : kNone;
    run_rounds<kWG>(tab, io, L, (sp.S + kR - 1u) / kR, [&](RLane& l, uint32_t re) {
        const uint32_t q = re >> sp.log2K;
        const uint32_t v = pack_entry(l.xi, l.o);
        if (ld_coherent(&w.ck[ck_slot(sp, q)]) == v) {
            mq = q;
            wcm = l.cnt;
            merged = 1u;
            late = 1u;
            l.active = 0u;
            return;
        }
        w.wb[ck_slot(sp, q)] = v;
        w.wbc[ck_slot(sp, q)] = l.cnt;
    });
    if (want && late) {                 // merged during the rounds
        w.mq[gc] = mq;
        w.wcm[gc] = wcm;
    } else if (want && !merged) {
        w.mq[gc] = kNone;
        w.wexit[gc] = pack_entry(L.xi, L.o);
        w.wtotal[gc] = L.cnt;
    }
    return merged;
}

// Walk from `entry` at the start of chunk c (c >= 1) and, while it has not merged with
// the stored spec path, on through the following chunks up to chunk c_end (exclusive),
// for the lanes with want != 0.  Each chunk the walk enters gets its full record
// (wentry, wb/wbc up to the merge, then mq/wcm or wexit/wtotal), written in the round
// the walk reaches that point.  Walks of one workgroup run in lockstep rounds, so when
// several walks pass through one chunk the one that started further back writes last;
// that walk's start is the exit recorded for its predecessor chunk, so the records of
// a range stay consistent.  Every walk starts at a chunk start, so chunk boundaries fall
// in the same rounds for all lanes; in such a round the walks leaving a chunk write its
// closing record (mq = kNone, wexit, wtotal) before a barrier, and the walks entering a
// chunk write theirs (wentry, and mq/wcm if they merge at its first checkpoint) after
// it: a walk that started one chunk further back can enter chunk g in the same round in
// which another walk leaves g.
// nr_max (uniform over the workgroup) bounds the rounds of every lane's walk.  Every four
// rounds the walks still running are compacted into the lowest lanes when that frees a
// wave (their state moves through LDS), so a long tail of a few walks keeps one wave
// issuing instead of four.
// With mrg (LDS, one word per origin lane, zeroed by the caller) the walks publish when
// they merge in their first chunk, and a walk leaves its first chunk only if the walk of
// the chunk before it merged inside its own chunk: otherwise its start (the previous chunk's spec exit) is not on
// the true path, and continuing would overwrite good records with a wrong path (the
// previous lane's continuing walk covers those chunks instead).
template <int W>
__device__ void walk_chain(const StreamPlan& sp, const Work& w, const LdsTab& tab,
                           RoundIO& io, uint32_t want, uint32_t c, uint32_t c_end,
                           uint32_t entry, uint32_t nr_max, uint32_t* mrg = nullptr) {
    __shared__ uint32_t s_wany[W / 64], s_wcnt[W / 64];
    uint32_t gc = sp.chunk_base + c;       // chunk the walk is in
    uint32_t gc0 = gc;                     // its first chunk
    uint32_t org = threadIdx.x;            // its lane before any compaction (mrg index)
    RLane L;
    L.active = 0u;
    if (want) {
        const uint32_t u0 = c * sp.S;
        L.u_stop = min(c_end * sp.S, sp.n_units);
        L.rbase = u0;
        w.wentry[gc] = entry;
        const uint32_t q0 = u0 >> sp.log2K;
        if (ld_coherent(&w.ck[ck_slot(sp, q0)]) == entry) {
            w.mq[gc] = q0;
            w.wcm[gc] = 0u;
            if (mrg) mrg[org] = 1u;
        } else {
            w.wb[ck_slot(sp, q0)] = entry;
            w.wbc[ck_slot(sp, q0)] = 0u;
            L.active = 1u;
            L.xi = entry_state(entry);
            L.u = u0 + (entry_off(entry) >> 5);
            L.o = entry_off(entry) & 31u;
            L.cnt = 0u;
        }
    }
    io.rbase[threadIdx.x] = L.active ? L.rbase : kNone;
    const uint32_t r_first = sp.S / kR;    // rounds of a chunk (a power of two)
    const uint32_t wave = threadIdx.x >> 6;
    __syncthreads();                       // rbase[] visible
    Prefetch pf;
    round_load<W>(io, 0u, pf);
    for (uint32_t r = 0; r < nr_max; ++r) {
        __syncthreads();                   // previous round's reads of buf done
        round_store<W>(io, pf);
        __syncthreads();
        if (r + 1u < nr_max) round_load<W>(io, r + 1u, pf);
        // the spec checkpoint at this round's end, read before the round decodes (the
        // range's spec records are complete before its walks start)
        uint32_t ckv = kNone;
        if (L.active) {
            const uint32_t re = L.rbase + (r + 1u) * kR;
            if (re < L.u_stop) ckv = ld_coherent(&w.ck[ck_slot(sp, re >> sp.log2K)]);
        }
        round_decode(tab, io.buf + threadIdx.x * kLS, L, r);
        const bool bnd = ((r + 1u) & (r_first - 1u)) == 0u;   // chunk boundary round (uniform)
        uint32_t test = 0u, v = 0u, q = 0u;            // merge test at checkpoint q
        if (L.active) {
            const uint32_t re = L.rbase + (r + 1u) * kR;
            if (re < L.u_stop) {
                if (L.u == re) {
                    v = pack_entry(L.xi, L.o);
                    q = re >> sp.log2K;
                    test = 1u;
                    if (bnd) {                     // close this chunk
                        w.mq[gc] = kNone;
                        w.wexit[gc] = v;
                        w.wtotal[gc] = L.cnt;
                        if (mrg && gc == gc0 && org > 0u && !mrg[org - 1u]) {
                            L.active = 0u;         // start not proven: stop here
                            test = 0u;
                        } else {
                            ++gc;                  // enter the next (after the barrier)
                            L.cnt = 0u;
                        }
                    }
                }
            } else {                       // end of the range without a merge
                L.active = 0u;
                w.mq[gc] = kNone;
                w.wexit[gc] = pack_entry(L.xi, L.o);
                w.wtotal[gc] = L.cnt;
            }
        }
        if (bnd) __syncthreads();          // closing records before entering ones
        if (test) {
            if (bnd) w.wentry[gc] = v;
            if (ckv == v) {                        // ck[q], read before the round
                w.mq[gc] = q;
                w.wcm[gc] = L.cnt;
                if (mrg && gc == gc0) mrg[org] = 1u;
                L.active = 0u;
            } else {
                w.wb[ck_slot(sp, q)] = v;
                w.wbc[ck_slot(sp, q)] = L.cnt;
            }
        }
        if (!L.active) io.rbase[threadIdx.x] = kNone;
        const uint32_t n_act = (uint32_t)__syncthreads_count(L.active);
        if (n_act == 0u) break;
        // Compaction: every four rounds, when it frees a wave, the walks still running
        // move to the lowest lanes (carrying their first chunk and origin lane, which the
        // merge flags are indexed by), so that the walks left keep as few waves issuing
        // as their number needs -- also in the first chunk, where most walks end.
        if (((r + 1u) & 3u) != 0u || r + 1u >= nr_max) continue;
        const uint64_t bal = __ballot(L.active);
        if ((threadIdx.x & 63u) == 0u) {
            s_wany[wave] = bal ? 1u : 0u;
            s_wcnt[wave] = (uint32_t)__popcll(bal);
        }
        __syncthreads();
        uint32_t waves_busy = 0u, before = 0u;
        for (uint32_t k = 0; k < W / 64u; ++k) {
            waves_busy += s_wany[k];
            if (k < wave) before += s_wcnt[k];
        }
        if ((n_act + 63u) / 64u >= waves_busy) continue;       // uniform decision
        uint32_t* x = io.buf;                                    // round data is dead here
        if (L.active) {
            const uint32_t pos = before + __builtin_amdgcn_mbcnt_hi(
                (uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
            uint32_t* d = x + pos * 8u;
            d[0] = L.u; d[1] = L.o; d[2] = L.xi; d[3] = L.cnt;
            // a walk stays inside its 256-chunk range: gc - gc0 and org fit a byte each
            d[4] = L.u_stop; d[5] = L.rbase; d[6] = gc; d[7] = (gc - gc0) | (org << 8);
        }
        __syncthreads();
        L.active = threadIdx.x < n_act ? 1u : 0u;
        if (L.active) {
            const uint32_t* d = x + threadIdx.x * 8u;
            L.u = d[0]; L.o = d[1]; L.xi = d[2]; L.cnt = d[3];
            L.u_stop = d[4]; L.rbase = d[5]; gc = d[6];
            gc0 = gc - (d[7] & 0xFFu);
            org = d[7] >> 8;
        }
        io.rbase[threadIdx.x] = L.active ? L.rbase : kNone;
        __syncthreads();                   // new rbase[] visible, exchange reads done
        round_load<W>(io, r + 1u, pf);     // reload the next round for the new lanes
    }
}

// Spec pass, then merge walks, for a range of 256 consecutive chunks of one stream.
// Workgroups take their task from a ticket counter in task order, in which the ranges of
// a stream are ascending, so a range's predecessor has always started before it.
// The walk into chunk c0 + l (lane l) starts from the spec exit of chunk c0 + l - 1 and
// continues until it merges, at most to the end of the range; chunk 0 is the true start
// and is not walked.  Lane 0's start is the previous range's last exit, published by
// that range's workgroup as soon as its spec pass ends (k_init set it to kNone): the
// workgroup waits for it, a bounded number of polls (a wait that runs out skips that
// walk and marks the chunk's start kNone, so k_check lists it and k_fixup walks it).
// W: lanes per workgroup = chunks per range (the launch's range_of(a)).
template <int W>
__global__ __launch_bounds__(W) void k_sync(KArgs a) {
    if (gated_off(a)) return;
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t s_task, s_stream, s_entry;
    __shared__ uint32_t s_rbase[W];
    __shared__ uint32_t s_exit[W];
    __shared__ uint32_t s_mrg[W];
    if (threadIdx.x == 0) {
        const uint32_t t = chunk_task(a, atomicAdd(&a.w.ctick[1], 1u));
        s_task = t;
        s_stream = stream_of_task<false>(a, t);
    }
    __syncthreads();
    const uint32_t t = s_task;
    const StreamPlan sp = *plan_of(a, __builtin_amdgcn_readfirstlane(s_stream));
    const uint32_t sf = stage_table(sp, lds, a.w.status, true);
    const uint32_t cyc = sf & 1u;
    if (!sf) fold_zero_bit_runs<W>(lds, sp.L, lds + table_words(sp.L), W * kLS);
    LdsTab tab{lds_addr(lds) >> 2};
    RoundIO io{lds + table_words(sp.L), s_rbase, sp.units, sp.n_units, (uint32_t)W};
    const Work& w = a.w;
    const uint32_t c0 = (t - sp.task_chunk) * W;
    const uint32_t c = c0 + threadIdx.x;
    if (sf & 2u) {
        // rejected table (uniform): empty, consistent records, so that the later
        // kernels read in-range values (the stream's output is not written)
        if (c < sp.n_chunks) {
            const uint32_t gc = sp.chunk_base + c;
            w.mq[gc] = kNone;
            w.wentry[gc] = 0u;
            w.wexit[gc] = 0u;
            w.wtotal[gc] = 0u;
            w.exitv[gc] = 0u;
            w.own_total[gc] = 0u;
        }
        return;
    }
    if (cyc) {
        // zero-bit cycles: only a single checkpoint interval is decodable (count-driven)
        if (c == 0u && sp.n_chunks) {
            const uint32_t off0 = 32u - sp.first_bit;
            w.ck[ck_slot(sp, 0u)] = pack_entry(sp.first_state - sp.L, off0);
            w.ckc[ck_slot(sp, 0u)] = 0u;
            w.mq[sp.chunk_base] = 0u;
            w.wcm[sp.chunk_base] = 0u;
        }
        return;
    }
    const bool valid = c < sp.n_chunks;
    const uint32_t c_end = min(c0 + (uint32_t)W, sp.n_chunks);
    io.lim = c_end - c0;                       // lanes with a chunk of this range
    // ---- spec (a chunk c >= 1 starts its guessed decode sp.warm units early; its first
    // checkpoint is the state at the round boundary u0, and its counts start there)
    RLane L;
    L.active = valid ? 1u : 0u;
    const uint32_t u0 = c * sp.S;
    const uint32_t wu = c == 0u ? 0u : min(sp.warm, u0);
    if (valid) {
        uint32_t off0 = c == 0u ? 32u - sp.first_bit : 0u;
        L.xi = sp.first_state - sp.L;
        if (sp.index_in && c != 0u) {      // recorded sync point, masked into range
            const SyncPoint p = sp.index_in[c];
            off0 = p.bit & 31u;
            L.xi = p.state & (sp.L - 1u);
        }
        L.rbase = u0 - wu;
        L.u_stop = min(u0 + sp.S, sp.n_units);
        L.u = L.rbase + (off0 >> 5);
        L.o = off0 & 31u;
        L.cnt = 0u;
        if (wu == 0u) {
            w.ck[ck_slot(sp, u0 >> sp.log2K)] = pack_entry(L.xi, off0);
            w.ckc[ck_slot(sp, u0 >> sp.log2K)] = 0u;
        }
        if (c == 0u) {
            w.mq[sp.chunk_base] = 0u;
            w.wcm[sp.chunk_base] = 0u;
            w.wentry[sp.chunk_base] = pack_entry(L.xi, off0);
        }
    }
    s_rbase[threadIdx.x] = valid ? L.rbase : kNone;
    run_rounds<W>(tab, io, L, (sp.S + sp.warm + kR - 1u) / kR, [&](RLane& l, uint32_t re) {
        if (re < u0) return;                       // still warming up
        if (re == u0) l.cnt = 0u;                  // chunk start: counts begin here
        w.ck[ck_slot(sp, re >> sp.log2K)] = pack_entry(l.xi, l.o);
        w.ckc[ck_slot(sp, re >> sp.log2K)] = l.cnt;
    });
    const uint32_t my_exit = pack_entry(L.xi, L.o);
    if (valid) {
        if (c + 1u == c_end && c_end < sp.n_chunks)    // the next range waits for this one
            __hip_atomic_store(&w.exitv[sp.chunk_base + c], my_exit, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        else
            w.exitv[sp.chunk_base + c] = my_exit;
        w.own_total[sp.chunk_base + c] = L.cnt;
    }
    s_exit[threadIdx.x] = my_exit;
    s_mrg[threadIdx.x] = c == 0u ? 1u : 0u;   // the stream start counts as merged
    if (threadIdx.x == 0) {
        uint32_t e = 0u;
        if (c0 > 0u) {
            const uint32_t* p = &w.exitv[sp.chunk_base + c0 - 1u];
            for (uint32_t i = 0;; ++i) {
                e = ld_coherent(p);
                if (e != kNone || i >= a.wait_polls) break;
                __builtin_amdgcn_s_sleep(4);
            }
            if (e == kNone) atomicAdd(&w.status[7], 1u);    // waits run out (statistics)
        }
        s_entry = e;
    }
    __threadfence();
    __syncthreads();
    // ---- walks
    uint32_t want = (c >= 1u && c < c_end) ? 1u : 0u;
    uint32_t entry = 0u;
    if (want) entry = threadIdx.x ? s_exit[threadIdx.x - 1u] : s_entry;
    if (want && entry == kNone) {          // predecessor not published in time
        w.wentry[sp.chunk_base + c] = kNone;
        want = 0u;
    }
    const uint32_t nr = ((c_end - c0) * sp.S + kR - 1u) / kR;
    walk_chain<W>(sp, w, tab, io, want, c, c_end, entry, nr, s_mrg);
    // ---- repair passes: a chunk whose recorded start is not its predecessor's exit (a
    // walk that stopped at its chunk's end although its start was right, found right by a
    // longer walk) is walked again from that exit.  The walks of a pass start together and
    // continue until they merge (entries are captured before any of them writes); a pass
    // can change an exit that a later chunk started from, so passes repeat, at most
    // kRepairPasses.  What remains (and chunk c0's entry) is left to k_check / k_fixup.
    for (uint32_t pass = 0; pass < kRepairPasses; ++pass) {
        __threadfence();
        __syncthreads();
        uint32_t want2 = 0u, entry2 = 0u;
        if (c > c0 && c < c_end) {
            entry2 = chunk_exit_now(sp, w, c - 1u);
            want2 = ld_coherent(&w.wentry[sp.chunk_base + c]) != entry2 ? 1u : 0u;
        }
        if (!__syncthreads_or(want2)) break;
        walk_chain<W>(sp, w, tab, io, want2, c, c_end, entry2, nr);
    }
}

// Consistency check over all chunks (grid as k_sync): a chunk whose recorded walk start
// differs from its predecessor's exit goes on its stream's fixup list.  Clears fflag.
__global__ __launch_bounds__(kWG) void k_check(KArgs a) {
    if (gated_off(a)) return;
    __shared__ uint32_t s_stream;
    const uint32_t t = chunk_task(a, blockIdx.x);
    if (threadIdx.x == 0) s_stream = stream_of_task<false>(a, t);
    __syncthreads();
    const uint32_t sid = __builtin_amdgcn_readfirstlane(s_stream);
    const StreamPlan sp = *plan_of(a, sid);
    const Work& w = a.w;
    const uint32_t c = (t - sp.task_chunk) * blockDim.x + threadIdx.x;   // range = blockDim.x
    if (c >= sp.n_chunks) return;
    const uint32_t g = sp.chunk_base + c;
    w.fflag[g] = 0u;
    if (c >= 1u && w.wentry[g] != chunk_exit(sp, w, c - 1u)) {
        const uint32_t slot = atomicAdd(&w.fcnt[sid], 1u);
        w.flist[sp.chunk_base + slot] = g;
    }
}

// Fixup rounds.  Workgroup g handles streams g, g+G, ...; a stream's list (in its
// own chunk range of flist/flist2, count in fcnt) holds chunks whose recorded walk
// start (wentry) may differ from the current exit of their predecessor.  Each round
// walks them again from that exit and lists any chunk (itself, or its successor)
// still inconsistent afterwards.  The leftmost listed chunk always starts from a
// proven exit and is consistent after its round, so at most n_chunks rounds run; the
// loop stops after n_chunks + 1 (or a.round_limit) rounds and reports a stream that is
// still inconsistent then in status[0] bit 1 instead of looping on.  Streams are
// independent, so no grid-wide sync.
__global__ __launch_bounds__(kWG) void k_fixup(KArgs a) {
    if (gated_off(a)) return;
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t s_n, s_n2;
    __shared__ uint32_t s_rbase[kWG];
    const Work& w = a.w;
    for (uint32_t si = blockIdx.x; si < a.n_streams; si += gridDim.x) {
        const uint32_t sid = a.sorder ? a.sorder[si] : si;
        __syncthreads();
        if (threadIdx.x == 0) s_n = ld_coherent(&w.fcnt[sid]);
        __syncthreads();
        if (s_n == 0u) continue;
        const StreamPlan sp = *plan_of(a, sid);
        if (stage_table(sp, lds, w.status, true)) continue;
        fold_zero_bit_runs<kWG>(lds, sp.L, lds + table_words(sp.L), kWG * kLS);
        LdsTab tab{lds_addr(lds) >> 2};
        RoundIO io{lds + table_words(sp.L), s_rbase, sp.units, sp.n_units, (uint32_t)kWG};
        uint32_t* list = w.flist + sp.chunk_base;
        uint32_t* next = w.flist2 + sp.chunk_base;
        if (threadIdx.x == 0) {                                    // statistics
            atomicAdd(&w.status[5], 1u);
            atomicMax(&w.status[6], s_n);
        }
        const uint32_t cap = a.round_limit && a.round_limit <= sp.n_chunks ? a.round_limit
                                                                             : sp.n_chunks + 1u;
        for (uint32_t round = 0;; ++round) {
            __syncthreads();
            const uint32_t n = s_n;
            if (n == 0u) {
                if (threadIdx.x == 0) atomicMax(&w.status[4], round);
                break;
            }
            if (round >= cap) {                    // not converged (uniform): report, stop
                if (threadIdx.x == 0) {
                    atomicOr(&w.status[0], 2u);
                    if (sp.status_out) atomicOr(sp.status_out, 2u);
                }
                break;
            }
            for (uint32_t i = threadIdx.x; i < n; i += kWG) {        // capture + mark
                const uint32_t g = list[i];
                next[i] = chunk_exit(sp, w, g - sp.chunk_base - 1u);
                w.fflag[g] = 1u;
            }
            __threadfence();
            __syncthreads();
            for (uint32_t b0 = 0; b0 < n; b0 += kWG) {               // walk again
                const uint32_t i = b0 + threadIdx.x;
                const uint32_t mine = i < n ? 1u : 0u;
                const uint32_t g = mine ? list[i] : sp.chunk_base + 1u;
                walk_rounds(sp, w, tab, io, mine, g - sp.chunk_base, mine ? next[i] : 0u);
            }
            __threadfence();
            if (threadIdx.x == 0) s_n2 = 0u;
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < n; i += kWG) {        // consistency
                const uint32_t g = list[i];
                const uint32_t c = g - sp.chunk_base;
                if (w.wentry[g] != chunk_exit(sp, w, c - 1u)) next[atomicAdd(&s_n2, 1u)] = g;
                if (c + 1u < sp.n_chunks && !w.fflag[g + 1u] &&
                    w.wentry[g + 1u] != chunk_exit(sp, w, c))
                    next[atomicAdd(&s_n2, 1u)] = g + 1u;
            }
            __threadfence();
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < n; i += kWG) w.fflag[list[i]] = 0u;
            if (threadIdx.x == 0) {
                s_n = s_n2;
                atomicAdd(&w.status[3], 1u);                      // rounds (statistics)
            }
            __threadfence();
            uint32_t* tmp = list; list = next; next = tmp;
        }
    }
}

// Exclusive scan of chunk step counts over all chunks of all streams (three passes:
// block reduce, scan of block sums, block scan + carry).  A stream's offsets are the
// global scan minus the value at its first chunk (see checkpoint_offset).
constexpr uint32_t kScanBlock = 2048;   // chunks per scan block (256 lanes x 8)

__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t v) {
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t d = 1; d < 64u; d <<= 1) {
        uint64_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// inclusive scan over the workgroup (kWG lanes); returns the total in *tot
__device__ uint64_t wg_incl_scan(uint64_t v, uint64_t* s_w, uint64_t* tot) {
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    v = wave_incl_scan(v);
    if (lane == 63u) s_w[wv] = v;
    __syncthreads();
    uint64_t add = 0;
    for (uint32_t i = 0; i < wv; ++i) add += s_w[i];
    *tot = s_w[0] + s_w[1] + s_w[2] + s_w[3];
    __syncthreads();
    return v + add;
}


// Single-pass exclusive scan with decoupled look-back (replaces the reference's phase-3
// copies + thrust::exclusive_scan).  Workgroups take blocks of kScanBlock chunks in
// ticket order (ctick[2]), so every block a workgroup looks back at belongs to a
// workgroup that started earlier.  Each block scans its counts in registers and LDS,
// publishes its total (flag 1), then one wave walks back over the preceding blocks'
// descriptors 64 at a time, adding totals down to the nearest inclusive prefix (flag 2),
// publishes its own inclusive prefix and writes its chunk offsets.  A descriptor is one u64: flag in bits 62-63,
// value below (offsets are far below 2^62).  k_init zeroes the descriptors.  A
// predecessor publishes its total without waiting on anything, so every wait ends; the
// wait is bounded anyway (kScanPolls polls per window) and a run-out sets status bit 2
// (reported as MANS_ECORRUPT).
constexpr uint64_t kScanAgg = 1ull << 62, kScanIncl = 2ull << 62, kScanVal = kScanAgg - 1ull;
constexpr uint32_t kScanPolls = 1u << 22;   // bound of one look-back wait (never reached)

__global__ __launch_bounds__(kWG) void k_scan(KArgs a, uint32_t total, uint64_t* desc) {
    if (gated_off(a)) return;
    __shared__ uint64_t s_w[4];
    __shared__ uint32_t s_blk;
    __shared__ uint64_t s_prefix;
    if (threadIdx.x == 0) s_blk = atomicAdd(&a.w.ctick[2], 1u);
    __syncthreads();
    const uint32_t blk = s_blk;
    const uint32_t per = kScanBlock / kWG;
    const uint32_t base = blk * kScanBlock + threadIdx.x * per;
    uint64_t c[kScanBlock / kWG];
    uint64_t v = 0;
    // the thread's chunks are consecutive: find the first one's stream, then step on
    uint32_t sid = base < total ? stream_of_chunk(a, base) : 0u;
    for (uint32_t i = 0; i < per; ++i) {
        const uint32_t gc = base + i;
        c[i] = 0u;
        if (gc < total) {
            while (a.plans && sid + 1u < a.n_streams && a.plans[sid + 1u].chunk_base <= gc) ++sid;
            const StreamPlan& sp = *plan_of(a, sid);
            if (a.lin_count) c[i] = a.lin_count[gc];
            else if (gc - sp.chunk_base < sp.n_chunks) c[i] = chunk_count(sp, a.w, gc - sp.chunk_base);
        }
        v += c[i];
    }
    uint64_t tot;
    const uint64_t incl = wg_incl_scan(v, s_w, &tot);
    if (threadIdx.x < 64u) {                   // wave 0: publish, then look back 64 at a time
        const uint32_t lane = threadIdx.x;
        uint64_t prefix = 0;
        if (blk == 0u) {
            if (lane == 0u)
                __hip_atomic_store(&desc[0], kScanIncl | tot, __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (lane == 0u)
                __hip_atomic_store(&desc[blk], kScanAgg | tot, __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_AGENT);
            // window [lo, hi): descriptor of block hi - 1 - lane
            uint32_t hi = blk;
            bool done = false, ran_out = false;
            while (!done && hi > 0u) {
                const uint32_t j = hi - 1u - lane;        // wraps for lanes below the start
                const bool in = lane < hi;
                uint64_t d = 0;
                for (uint32_t i = 0;; ++i) {
                    d = in ? __hip_atomic_load(&desc[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                           : kScanIncl;                   // past block 0: nothing more
                    const uint64_t missing = __ballot((d >> 62) == 0u);
                    if (!missing) break;
                    if (i >= kScanPolls) { ran_out = true; break; }
                    __builtin_amdgcn_s_sleep(1);
                }
                __atomic_thread_fence(__ATOMIC_ACQUIRE);
                if (ran_out) break;
                // the nearest inclusive prefix in the window (lowest lane with flag 2)
                const uint64_t inclm = __ballot((d >> 62) == 2u);
                const uint32_t stop = inclm ? (uint32_t)__builtin_ctzll(inclm) : 64u;
                uint64_t add = (lane <= stop && in) ? (d & kScanVal) : 0ull;
                for (uint32_t o = 32; o > 0; o >>= 1) add += __shfl_xor(add, o, 64);
                prefix += add;
                if (inclm) done = true;
                hi = hi > 64u ? hi - 64u : 0u;
            }
            if (ran_out) {
                if (lane == 0u) atomicOr(&a.w.status[0], 4u);
            } else if (lane == 0u) {
                __hip_atomic_store(&desc[blk], kScanIncl | (prefix + tot), __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (lane == 0u) s_prefix = prefix;
    }
    __syncthreads();
    uint64_t run = s_prefix + incl - v;
    for (uint32_t i = 0; i < per; ++i) {
        if (base + i < total) a.w.coff[base + i] = run;
        run += c[i];
    }
}

// Sync points of the streams that ask for them (index_out), one lane per chunk (grid
// as k_check), from the proven records (mans::chunk_sync_point).
__global__ __launch_bounds__(kWG) void k_index(KArgs a) {
    __shared__ uint32_t s_stream;
    const uint32_t t = blockIdx.x;
    if (threadIdx.x == 0) s_stream = stream_of_task<false>(a, t);
    __syncthreads();
    const StreamPlan sp = *plan_of(a, __builtin_amdgcn_readfirstlane(s_stream));
    if (!sp.index_out) return;
    const uint32_t c = (t - sp.task_chunk) * blockDim.x + threadIdx.x;   // range = blockDim.x
    if (c < sp.n_chunks) sp.index_out[c] = chunk_sync_point(sp, a.w, c);
}

// Derived table words read from the reference table in HBM (derive_entry per lookup).
struct GlobalTab {
    const uint32_t* t;
    uint32_t L;
    __device__ uint32_t operator[](uint32_t i) const {
        int bad = 0, cyc = 0;
        return derive_entry(load_unit(t, L, i), L, i, &bad, &cyc);
    }
};

// The reference decoder's sync state (mans_ref_sync) of a single stream, from the proven
// checkpoints: lane q decodes checkpoint interval q (its true entry, its scanned offset and
// step count) from HBM and, for each subsequence of ref_S units, writes the count of
// symbols before its first step (where that step lies in the interval) and the unit, bit
// and state of its last step (where that lies in the interval).  Every unit below the
// last holds at least one step start (codewords are at most 15 bits), so whether an
// interval's first (last) step is its subsequence's first (last) follows from where the
// subsequence starts (ends) relative to the interval.  Lanes below ref_nblocks also write
// the per-block flags: 1, or 0 when the fixup rounds did not converge (status bit 1).
// Compatibility output, not on the decode path (no LDS staging).
__global__ __launch_bounds__(kWG) void k_refsync(KArgs a) {
    const StreamPlan& sp = a.one;
    const Work& w = a.w;
    const uint32_t q = blockIdx.x * kWG + threadIdx.x;
    if (q < a.ref_nblocks && a.ref_flags) a.ref_flags[q] = (w.status[0] & 2u) ? 0u : 1u;
    // intervals: the proven checkpoints (K units), or the 32-unit item records of a
    // lineage decode whose records held
    const bool lin = a.lin_rec &&
        __hip_atomic_load(a.lin_fail_word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u;
    const uint32_t n_int = lin ? a.lin_nq : sp.n_ck, K = lin ? kLinQ : sp.K;
    if (q >= n_int) return;
    uint32_t entry, e2;
    uint64_t off, end;
    if (lin) {
        const SyncPoint r = a.lin_rec[q];
        entry = q == 0u ? pack_entry(sp.first_state - sp.L, 32u - sp.first_bit)
                        : pack_entry(r.state, r.bit);
        off = r.before;
        end = q + 1u < n_int ? a.lin_rec[q + 1u].before : sp.n_out;
    } else {
        off = checkpoint_offset(sp, w, q, &entry);
        end = q + 1u < n_int ? checkpoint_offset(sp, w, q + 1u, &e2) : sp.n_out;
    }
    if (end > sp.n_out) end = sp.n_out;
    if (off >= end) return;
    const bool last_interval = q + 1u == n_int;
    const uint32_t S = a.ref_S, nsub = a.ref_nsub;
    const uint64_t k0 = (uint64_t)q * K, k1 = k0 + K;            // interval units [k0, k1)
    GlobalTab tab{sp.table, sp.L};
    BCursor cur;
    bcursor_init(cur, sp.units, sp.n_units, (uint32_t)k0, entry_off(entry));
    uint32_t xi = entry_state(entry);
    uint32_t j = min(cur.u / S, nsub - 1u);
    for (uint64_t i = off; i < end; ++i) {
        const bool first = i == off ? (uint64_t)j * S >= k0 : false;
        if (first) {
            a.ref_sync[4ull * j + 2u] = (uint32_t)i;
            if (a.ref_sizes) a.ref_sizes[j] = (uint32_t)i;
        }
        const uint32_t unit_rel = cur.u - j * S, bit = cur.o, x = xi + sp.L;
        uint32_t crossed;
        bstep(tab, cur, xi, crossed);
        const uint32_t jn = min(cur.u / S, nsub - 1u);
        bool last;
        if (i + 1u == sp.n_out) last = true;
        else if (i + 1u == end) last = last_interval || (uint64_t)(j + 1u) * S <= k1;
        else last = jn != j;
        if (last) {
            a.ref_sync[4ull * j + 0u] = unit_rel;
            a.ref_sync[4ull * j + 1u] = bit;
            a.ref_sync[4ull * j + 3u] = x;
        }
        if (jn != j && i + 1u < end) {          // the next step opens subsequence jn
            a.ref_sync[4ull * jn + 2u] = (uint32_t)(i + 1u);
            if (a.ref_sizes) a.ref_sizes[jn] = (uint32_t)(i + 1u);
        }
        j = jn;
    }
}

// ---------------------------------------------------------------- write pass
// Persistent workgroups; each takes a contiguous range of passes, a pass being 256
// consecutive checkpoint intervals ("items", kR units each) of one stream.  A pass's
// units are staged in LDS like a round of the sync kernels (lane segment = item), the
// next pass's units load while the current pass decodes, every lane decodes its
// item's step count from the item's true entry, and symbols are packed four to a
// dword store (byte stores for the unaligned head and tail).
// A write item spans kWF checkpoint intervals (kWI units): items longer than the
// sync kernels' rounds amortise the per-pass work (offsets, barrier) over more steps.
constexpr uint32_t kWF = kR >= 32u ? 1u : 32u / kR;
constexpr uint32_t kWI = kR * kWF;          // units per item
constexpr uint32_t kWPieces = kWI / 4u + 1u;  // 16-byte loads per lane (item + lookahead)
// LDS stride per lane: the item plus the two units a step may look ahead (8-byte
// aligned, stored as 8-byte halves; 2 mod 4, which spreads lanes at the same offset over
// twice the banks a 16-byte multiple would); the shorter stride lets three workgroups
// share a CU
constexpr uint32_t kWL = kWI + 2u;

__host__ __device__ __forceinline__ uint32_t n_write_items(uint32_t n_ck) {
    return (n_ck + kWF - 1u) / kWF;
}

struct WPrefetch { uint4 v[kWPieces]; };

__device__ __forceinline__ void pass_load(const uint32_t* units, uint32_t n, uint32_t u_first,
                                          WPrefetch& pf) {
#pragma unroll
    for (uint32_t j = 0; j < kWPieces; ++j) {
        const uint32_t idx = threadIdx.x + kWG * j;
        const uint32_t lane = idx / kWPieces, piece = idx - lane * kWPieces;
        const U4 u = load_block_raw(units, n, u_first + lane * kWI + 4u * piece);
        pf.v[j] = make_uint4(u.x, u.y, u.z, u.w);
    }
}

__device__ __forceinline__ void pass_store(uint32_t* buf, const WPrefetch& pf) {
#pragma unroll
    for (uint32_t j = 0; j < kWPieces; ++j) {
        const uint32_t idx = threadIdx.x + kWG * j;
        const uint32_t lane = idx / kWPieces, piece = idx - lane * kWPieces;
        uint2* d = reinterpret_cast<uint2*>(buf + lane * kWL + 4u * piece);
        d[0] = make_uint2(pf.v[j].x, pf.v[j].y);
        if (piece + 1u < kWPieces) d[1] = make_uint2(pf.v[j].z, pf.v[j].w);
    }
}

__device__ __forceinline__ void item_write(const LdsTab& tab, const uint32_t* lbuf, uint32_t entry,
                                           uint8_t* dst_generic, uint32_t cnt) {
    // global (not flat) stores: flat stores also count on the LDS counter and would
    // force a full LDS wait every step
    typedef uint8_t __attribute__((address_space(1)))* gptr_u8;
    typedef uint32_t __attribute__((address_space(1)))* gptr_w;
    gptr_u8 dst = (gptr_u8)dst_generic;
    // p: LDS bit address in the item's staged units (at most kWI * 32 + 15 past their
    // start after the last step, so the window read stays inside the kWI + 2 units)
    uint32_t p = lds_addr(lbuf) * 8u + entry_off(entry);
    uint32_t lo, hi;
    lds_window_load(p, lo, hi);
    uint32_t w = lds_word((entry_state(entry) + tab.tb) << 2);
    // one step; the symbol is byte 1 of the result
    auto step1 = [&]() -> uint32_t {
        const uint32_t cur = w;
        const uint32_t win = funnel_shr(hi, lo, p);
        const uint32_t i1 = (w >> 16) + low_bits(win, w);
        p += w & 0xFFu;
        w = lds_word(i1 << 2);
        lds_window_load(p, lo, hi);
        return cur;
    };
    auto next_sym = [&]() -> uint32_t { return (step1() >> 8) & 0xFFu; };
    // two steps from one window (as in round_decode); symbols in bytes 0 and 1
    auto step2 = [&]() -> uint32_t {
        const uint32_t win = funnel_shr(hi, lo, p);
        const uint32_t i1 = (w >> 16) + low_bits(win, w);
        const uint32_t w1 = lds_word(i1 << 2);
        const uint32_t p1 = p + (w & 0xFFu);
        const uint32_t i2 = (w1 >> 16) + __builtin_amdgcn_ubfe(win, w, w1);
        const uint32_t two = __builtin_amdgcn_perm(w1, w, 0x0c0c0501u);
        w = lds_word(i2 << 2);
        p = p1 + (w1 & 0xFFu);
        lds_window_load(p, lo, hi);
        return two;
    };
    // Symbols are packed four to a dword and stored 16 bytes at a time: the lanes of a
    // wave write 64 separate regions, so every store instruction costs one cache-line
    // request per lane in the address unit, whatever its width.  Bytes, then dwords up
    // to 16-byte alignment, then 16-byte stores, then the tail.
    auto next_word = [&]() -> uint32_t {
        const uint32_t a = step2();
        const uint32_t b = step2();
        return __builtin_amdgcn_perm(b, a, 0x05040100u);
    };
    uint32_t i = 0;
    const uint32_t head = (uint32_t)((4u - ((uintptr_t)dst & 3u)) & 3u);
    const uint32_t h = head < cnt ? head : cnt;
    for (; i < h; ++i) dst[i] = (uint8_t)next_sym();
    gptr_w dw = (gptr_w)(dst + i);
    uint32_t nd = (cnt - i) >> 2;                  // whole dwords left
    const uint32_t wa = (uint32_t)(((16u - ((uintptr_t)(dst + i) & 15u)) & 15u) >> 2);
    const uint32_t ha = wa < nd ? wa : nd;
    for (uint32_t d = 0; d < ha; ++d) dw[d] = next_word();
    dw += ha;
    nd -= ha;
    typedef uint32_t v4u __attribute__((ext_vector_type(4)));
    typedef v4u __attribute__((address_space(1)))* gptr_q;
    gptr_q dq = (gptr_q)dw;
    const uint32_t nq = nd >> 2;
    for (uint32_t d = 0; d < nq; ++d) {
        v4u v;
        v.x = next_word();
        v.y = next_word();
        v.z = next_word();
        v.w = next_word();
        dq[d] = v;
    }
    dw += nq << 2;
    for (uint32_t d = 0; d < (nd & 3u); ++d) dw[d] = next_word();
    i += (ha + nd) << 2;
    for (; i < cnt; ++i) dst[i] = (uint8_t)next_sym();
}

// Trusted item record `it` (MANS_F_TRUST_SYNC_POINTS: one record per write item, 32
// units): its entry (item 0: the stream start) and its output offset, counted from
// record 0 and kept inside [0, n_out].
__device__ __forceinline__ uint64_t item_record(const StreamPlan& sp, uint32_t it, uint32_t* entry) {
    const SyncPoint q = sp.index_in[it];
    const uint64_t base = sp.index_in[0].before;
    if (it == 0u) {
        *entry = pack_entry(sp.first_state - sp.L, 32u - sp.first_bit);
        return 0u;
    }
    *entry = pack_entry(q.state & (sp.L - 1u), q.bit & 31u);
    const uint64_t v = q.before > base ? q.before - base : 0u;
    return v < sp.n_out ? v : sp.n_out;
}

__global__ __launch_bounds__(kWG) void k_write(KArgs a, uint32_t n_passes, uint32_t per_wg) {
    if (gated_off(a)) return;
    // Persistent workgroups take blocks of per_wg consecutive passes from a device
    // counter (ctick[0], zeroed by k_init), so slow passes (high symbols per unit) do
    // not pile up on a few workgroups.
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t s_have, s_next, s_skip;
    __shared__ uint64_t s_off[kWG + 1];
    if (threadIdx.x == 0) {
        s_have = kNone;
        s_next = atomicAdd(&a.w.ctick[0], per_wg);
    }
    __syncthreads();
    uint32_t t = s_next;
    if (t >= n_passes) return;
    uint32_t t_end = min(n_passes, t + per_wg);
    // The stream of pass t (sid) and the first pass of the stream after it (next_first),
    // uniform: a pass usually follows the previous one in the same stream or the next,
    // so the stream search runs only where a block of passes starts elsewhere.
    auto first_pass = [&](uint32_t s) -> uint32_t {
        return a.plans && s < a.n_streams
                   ? __builtin_amdgcn_readfirstlane(a.plans[s].task_item) : 0xFFFFFFFFu;
    };
    uint32_t sid = __builtin_amdgcn_readfirstlane(stream_of_task<true>(a, t));
    uint32_t next_first = first_pass(sid + 1u);
    // thread 0: the item that ended the last pass (stream, index, offset, entry)
    uint32_t carry_sid = kNone, carry_it = 0u, carry_entry = 0u;
    uint64_t carry_off = 0u;
    // prefetch the first pass
    WPrefetch pf;
    {
        const StreamPlan* sp0 = plan_of(a, sid);
        pass_load(sp0->units, sp0->n_units, (t - sp0->task_item) * kWG * kWI, pf);
    }
    while (t < n_passes) {
        __syncthreads();
        if (threadIdx.x == 0 && t + 1u == t_end) s_next = atomicAdd(&a.w.ctick[0], per_wg);
        const StreamPlan sp = *plan_of(a, sid);
        uint32_t* buf = lds + table_words(sp.L);
        if (s_have != sid) {
            const uint32_t sf = stage_table(sp, lds, a.w.status);    // ends with a barrier
            if (threadIdx.x == 0) {
                s_have = sid;
                s_skip = sf & 2u;                  // rejected table: nothing to write
            }
        }
        // items i0 .. i0+nitems-1 of the stream; item i starts at checkpoint i*kWF
        const uint32_t i0 = (t - sp.task_item) * kWG;
        const uint32_t nitems = min((uint32_t)kWG, n_write_items(sp.n_ck) - i0);
        pass_store(buf, pf);
        uint32_t entry = 0;
        const uint32_t n_items = n_write_items(sp.n_ck);
        if (threadIdx.x < nitems) {
            const uint32_t it = i0 + threadIdx.x;
            if (carry_sid == sid && carry_it == it) {   // thread 0: the last pass's end
                s_off[threadIdx.x] = carry_off;
                entry = carry_entry;
            } else if (a.direct) {              // trusted records, one per item
                s_off[threadIdx.x] = item_record(sp, it, &entry);
            } else {
                s_off[threadIdx.x] = checkpoint_offset(sp, a.w, it * kWF, &entry);
            }
            if (sp.items_out) {
                SyncPoint q;
                q.state = entry_state(entry);
                q.bit = entry_off(entry);
                q.before = s_off[threadIdx.x];
                sp.items_out[it] = q;
            }
        }
        if (threadIdx.x == 0) {
            // the pass's end: the next item's offset, kept (with its entry) for the next
            // pass when that pass starts at this item
            uint32_t e2 = 0;
            const uint32_t ie = i0 + nitems;
            const uint64_t oe = ie >= n_items ? sp.n_out
                              : a.direct ? item_record(sp, ie, &e2)
                                         : checkpoint_offset(sp, a.w, ie * kWF, &e2);
            s_off[nitems] = oe;
            carry_sid = ie < n_items ? sid : kNone;
            carry_it = ie;
            carry_off = oe;
            carry_entry = e2;
        }
        __syncthreads();
        const uint32_t tn = t + 1u < t_end ? t + 1u : __builtin_amdgcn_readfirstlane(s_next);
        uint32_t nsid = sid, nnext = next_first;
        if (tn < n_passes) {
            if (tn >= next_first) {                // the next pass belongs to another stream
                const uint32_t f2 = tn == t + 1u ? first_pass(sid + 2u) : 0u;
                if (tn < f2) {
                    nsid = sid + 1u;
                    nnext = f2;
                } else {
                    nsid = __builtin_amdgcn_readfirstlane(stream_of_task<true>(a, tn));
                    nnext = first_pass(nsid + 1u);
                }
            }
            const StreamPlan* np = plan_of(a, nsid);
            pass_load(np->units, np->n_units, (tn - np->task_item) * kWG * kWI, pf);
        }
        if (threadIdx.x < nitems && !s_skip) {
            uint64_t off = s_off[threadIdx.x], end = s_off[threadIdx.x + 1];
            if (end > sp.n_out) end = sp.n_out;
            if (off < end) {
                LdsTab tab{lds_addr(lds) >> 2};
                item_write(tab, buf + threadIdx.x * kWL, entry, sp.out + off, (uint32_t)(end - off));
            }
        }
        if (t + 1u == t_end) t_end = min(n_passes, tn + per_wg);
        t = tn;
        sid = nsid;
        next_first = nnext;
    }
}

// ------------------------------------------------------------ lineage synchronisation
// Single-stream decodes that a lineage rule selects find the true path with lineages
// (mans_lineage.h: a guessed-state decode starts every D units and runs until it meets a
// path another lineage recorded) instead of k_sync / k_check / k_fixup, then decode it with
// the trusted form of k_write.  The chunked pipeline is enqueued after the lineage kernels
// with a gate on the failure word, so it runs only when the lineage records do not hold.

struct LinArgs {
    LinWork lw;
    uint32_t D;               // units between lineage starts (multiple of kLinQ)
    uint32_t n_lin;           // lineages: ceil(n_units / D)
    uint32_t cap;             // checkpoints one lineage may pass
    uint32_t n_scan_blocks;   // k_scan descriptors to clear
    uint32_t* status;         // status words (offset 0 of the workspace)
    uint32_t* ctick;
    uint64_t* bsum;
};

__device__ __forceinline__ uint32_t lin_ld(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A failure also exhausts the lineage ticket, so lanes stop taking lineages at their
// next grab (no lane polls the failure word: one word read by every lane at every
// checkpoint serialises on its cache line).
__device__ __forceinline__ void lin_fail(const LinArgs& la, uint32_t reason) {
    atomicOr(&la.lw.ctl[2], reason);
    __hip_atomic_store(&la.lw.ctl[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    atomicMax(&la.lw.ctl[0], 0x40000000u);
}

// Clears the status and control words, the scan's tickets and descriptors, the slot
// counts and slot values, and records the stream start in slot 0 of checkpoint 0.
__global__ void k_lin_init(KArgs a, LinArgs la) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const LinWork& lw = la.lw;
    if (i < 128u) la.status[i] = 0u;
    if (i < 8u) la.ctick[i] = 0u;
    if (i < la.n_scan_blocks) la.bsum[i] = 0ull;
    if (i < lw.n_q) {
        uint4* v = reinterpret_cast<uint4*>(lw.val + (uint64_t)i * kLinSlots);
        const uint4 e = make_uint4(kLinEmpty, kLinEmpty, kLinEmpty, kLinEmpty);
#pragma unroll
        for (uint32_t k = 0; k < kLinSlots / 4u; ++k) v[k] = e;
        lw.cnt[i] = 0u;
        if (i == 0u) {
            lw.val[0] = pack_entry(a.one.first_state - a.one.L, 32u - a.one.first_bit);
            lw.cnt[0] = 1u;
        }
    }
}

// A lane's own round of units (kRL from unit rs, a multiple of 4), and its store into the
// lane's LDS segment (kLS units, 8-byte aligned).
__device__ __forceinline__ void lane_load(const uint32_t* units, uint32_t n, uint32_t rs,
                                          Prefetch& pf) {
#pragma unroll
    for (uint32_t j = 0; j < kPieces; ++j) {
        const U4 u = load_block_raw(units, n, rs + 4u * j);
        pf.v[j] = make_uint4(u.x, u.y, u.z, u.w);
    }
}

__device__ __forceinline__ void lane_store(uint32_t* seg, const Prefetch& pf) {
#pragma unroll
    for (uint32_t j = 0; j < kPieces; ++j) {
        uint2* d = reinterpret_cast<uint2*>(seg + 4u * j);
        if (4u * j < kLS) d[0] = make_uint2(pf.v[j].x, pf.v[j].y);
        if (4u * j + 2u < kLS) d[1] = make_uint2(pf.v[j].z, pf.v[j].w);
    }
}

// Persistent lineage kernel: every lane takes lineages from a ticket (the last start
// first, lineage 0 from the true start last) and runs each in rounds of kR units from
// its own LDS segment (the count-only round engine of k_sync, zero-bit runs folded),
// loading the next round's units while it decodes the current one.  At each checkpoint
// (two rounds) it searches the checkpoint's slots for its entry: found, it links its last
// slot there and ends; not found, it takes a slot, links to it and goes on.  No barriers
// after the table is staged: lanes are independent.
__global__ __launch_bounds__(kWG) void k_lin(KArgs a, LinArgs la) {
    extern __shared__ uint32_t lds[];
    const StreamPlan& sp = a.one;
    const LinWork& lw = la.lw;
    uint32_t* buf = lds + table_words(sp.L);
    const uint32_t tf = stage_table(sp, lds, la.status, true);
    if ((tf & 2u) || ((tf & 1u) && lw.n_q > 1u)) {       // uniform: leave it to k_sync
        if (threadIdx.x == 0) lin_fail(la, 16u);
        return;
    }
    fold_zero_bit_runs<kWG>(lds, sp.L, buf, kWG * kLS);
    __syncthreads();
    const LdsTab tab{lds_addr(lds) >> 2};
    uint32_t* seg = buf + threadIdx.x * kLS;
    RLane L;
    L.active = 0u;
    L.u_stop = sp.n_units;
    uint32_t q = 0, slot = kLinEmpty, half = 0, passed = 0, nlin = 0, nck = 0;
    const uint32_t lane = __lane_id();
    const uint32_t qstep = la.D / kLinQ;
    // Tickets: a wave takes 64 at a time (one atomic) and hands them to its lanes as they
    // need them, in rank order (the rest of a block first); every lineage is run (lineage 0,
    // the true path, last).  Each lane holds its next lineage's ticket (nt, kNone when none
    // is left) and has that lineage's first round of units in flight (pn) while it decodes
    // the current one.  take() is called by every lane still in the loop (wave-uniform).
    uint32_t pool = 0u, pool_left = 0u;
    auto take = [&](bool want) -> uint32_t {
        const uint64_t m = __ballot(want);
        if (!m) return kNone;
        const uint32_t n = (uint32_t)__popcll(m);
        const uint32_t r = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        uint32_t b = pool;
        if (n > pool_left) {
            const uint32_t lead = (uint32_t)__builtin_ctzll(__ballot(true));
            uint32_t nb = 0u;
            if (lane == lead) nb = atomicAdd(&lw.ctl[0], 64u);
            b = __shfl(nb, (int)lead, 64);
        }
        const uint32_t t = r < pool_left ? pool + r : b + (r - pool_left);
        if (n > pool_left) {
            pool = b + (n - pool_left);
            pool_left = 64u - (n - pool_left);
        } else {
            pool += n;
            pool_left -= n;
        }
        return want && t < la.n_lin ? t : kNone;
    };
    auto first_unit = [&](uint32_t t) -> uint32_t {
        return (la.n_lin - 1u - t) * qstep * kLinQ;
    };
    Prefetch pf, pn;
    uint32_t nt = take(true);
    if (nt != kNone) lane_load(sp.units, sp.n_units, first_unit(nt), pn);
    constexpr uint32_t kAhead = 8u;                  // slots read ahead (the most used: 7-9)
    uint32_t cv[kAhead], cu = 0u;
    for (;;) {
        // lanes without a lineage start the held one (its first round is in pn) and take
        // the next ticket; a lane with neither leaves
        const bool need = !L.active;
        const uint32_t t = need ? nt : kNone;
        const uint32_t t2 = take(need && t != kNone);
        if (need) {
            if (t == kNone) break;
            pf = pn;
            nt = t2;
            if (nt != kNone) lane_load(sp.units, sp.n_units, first_unit(nt), pn);
            const uint32_t j = la.n_lin - 1u - t;
            q = j * qstep;
            if (q + 1u >= lw.n_q) continue;           // no checkpoint after it
            ++nlin;
            passed = 0u;
            half = 0u;
            if (j == 0u) {
                slot = 0u;
                L.xi = sp.first_state - sp.L;
                const uint32_t b0 = 32u - sp.first_bit;
                L.u = b0 >> 5;
                L.o = b0 & 31u;
            } else {
                slot = kLinEmpty;
                L.xi = lin_guess(j, sp.L);
                L.u = q * kLinQ;
                L.o = 0u;
            }
            L.rbase = q * kLinQ;
            L.cnt = 0u;
            L.active = 1u;
        }
        lane_store(seg, pf);
        lane_load(sp.units, sp.n_units, L.rbase + kR, pf);  // the next round, in flight
        if (half) {
            // the slots of the checkpoint this round ends at, read before decoding it (a
            // value written meanwhile is only missed: the lineage then takes a slot of its
            // own, the same path twice)
            const uint64_t base = (uint64_t)(q + 1u) * kLinSlots;
            cu = lin_slots_used(lin_ld(&lw.cnt[q + 1u]));
#pragma unroll
            for (uint32_t s = 0; s < kAhead; ++s) cv[s] = lin_ld(&lw.val[base + s]);
        }
        round_decode(tab, seg, L, 0u);
        L.rbase += kR;
        if (!half) {
            half = 1u;
            continue;
        }
        // checkpoint q + 1 (unit L.rbase): the lane is at its entry
        half = 0u;
        ++q;
        ++nck;
        const uint32_t steps = L.cnt;
        L.cnt = 0u;
        const uint32_t v = pack_entry(L.xi, L.o);
        const uint64_t base = (uint64_t)q * kLinSlots;
        const uint64_t prev = (uint64_t)(q - 1u) * kLinSlots + slot;
        uint32_t found = kLinEmpty;
#pragma unroll
        for (uint32_t s = kAhead; s-- > 0;)
            if (s < cu && cv[s] == v) found = s;
        if (found == kLinEmpty && cu > kAhead) {              // rare: more slots in use
            for (uint32_t s = kAhead; s < cu; ++s)
                if (lin_ld(&lw.val[base + s]) == v) { found = s; break; }
        }
        if (found != kLinEmpty) {                             // merged
            if (slot != kLinEmpty) lw.inf[prev] = lin_link(found, steps);
            L.active = 0u;
            continue;
        }
        const uint32_t ns = atomicAdd(&lw.cnt[q], 1u);
        if (ns >= kLinSlots) {                                // slot list full
            if (slot != kLinEmpty) lw.inf[prev] = lin_link(kLinBroken, 0u);
            lin_fail(la, 1u);
            L.active = 0u;
            continue;
        }
        __hip_atomic_store(&lw.val[base + ns], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (slot != kLinEmpty) lw.inf[prev] = lin_link(ns, steps);
        slot = ns;
        if (q + 1u == lw.n_q) {                               // the last checkpoint
            lw.inf[base + ns] = lin_link(kLinEnd, 0u);
            L.active = 0u;
            continue;
        }
        if (++passed >= la.cap) {                             // longer than any merge
            lw.inf[base + ns] = lin_link(kLinBroken, 0u);
            lin_fail(la, 2u);
            L.active = 0u;
            continue;
        }
    }
    // statistics: one atomic per wave
    for (uint32_t o = 32; o > 0; o >>= 1) {
        nlin += __shfl_xor(nlin, (int)o, 64);
        nck += __shfl_xor(nck, (int)o, 64);
    }
    if (lane == 0u) {
        atomicAdd(&lw.ctl[3], nlin);
        atomicAdd(&lw.ctl[4], nck);
    }
}

__global__ void k_lin_map(LinArgs la) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= la.lw.n_c * kLinSlots || lin_ld(&la.lw.ctl[1])) return;
    lin_map_lane(la.lw, i / kLinSlots, i % kLinSlots);
}

__global__ void k_lin_resolve(LinArgs la) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= la.lw.n_c || lin_ld(&la.lw.ctl[1])) return;
    const uint32_t r = lin_resolve_lane(la.lw, c);
    if (r) lin_fail(la, r);
}

__global__ void k_lin_trace(KArgs a, LinArgs la) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= la.lw.n_c || lin_ld(&la.lw.ctl[1])) return;
    const uint32_t r = lin_trace_lane(la.lw, c, a.w.coff[c]);
    if (r) lin_fail(la, r);
}

}  // namespace

// ------------------------------------------------------------------ host side
namespace mans {

static uint32_t ilog2(uint32_t v) { uint32_t r = 0; while ((1u << (r + 1)) <= v) ++r; return r; }

// Chunk/checkpoint lengths for one stream.  Sync chunks must be much longer than
// the typical distance over which a guessed-state decode merges with the true
// path (tens to hundreds of units at L = 4096), so the spec and walk passes stay
// close to one decode of the stream.
static std::atomic<uint32_t> g_wait_polls{MANS_WAIT_POLLS_DEFAULT};
static std::atomic<uint32_t> g_round_limit{0u};
static std::mutex g_rules_mu;
static std::vector<mans_geometry_rule> g_rules;

// Built-in rules: the autotuner's table for MI355X (multians_amd/tune.py); empty until
// a tuning run has filled it.
static const mans_geometry_rule kTunedRules[] = {
#include "tuned_rules.inc"
    {0u, 0u, 0.0, 0ull, 0u, 0u}};

// Smallest chunk length the rules may choose (the workspace bound of mans_workspace_bytes
// assumes it).
constexpr uint32_t kMinRuleChunk = 32;

struct RuleHit {
    uint32_t S;     // 0 = no rule
    uint32_t W;     // MANS_WARM_AUTO = built-in warm-up
};

static RuleHit rule_geometry(uint64_t n_units, uint64_t n_out, uint32_t L, uint32_t scope) {
    const double spu = n_units ? (double)n_out / (double)n_units : 4.0;
    auto match = [&](const mans_geometry_rule& r) {
        return r.chunk_units && L <= r.L_max && spu <= r.spu_max && n_units <= r.units_max &&
               (r.scope == 0u || (r.scope & scope));
    };
    {
        std::lock_guard<std::mutex> lk(g_rules_mu);
        for (const mans_geometry_rule& r : g_rules)
            if (match(r)) return {r.chunk_units, r.warm_units};
        if (!g_rules.empty()) return {0u, MANS_WARM_AUTO};
    }
    for (const mans_geometry_rule& r : kTunedRules)
        if (match(r)) return {r.chunk_units, r.warm_units};
    return {0u, MANS_WARM_AUTO};
}

// Lineage spacing D (units) for a single-stream decode, from the first rule of scope
// MANS_RULE_LINEAGE that matches (rules of other scopes never select it); 0 = none, or a
// matching rule with chunk_units 0 (the chunked pipeline for that class of streams).
static uint32_t rule_lineage(uint64_t n_units, uint64_t n_out, uint32_t L) {
    const double spu = n_units ? (double)n_out / (double)n_units : 4.0;
    auto match = [&](const mans_geometry_rule& r) {
        return r.scope == MANS_RULE_LINEAGE && L <= r.L_max && spu <= r.spu_max &&
               n_units <= r.units_max;
    };
    uint32_t D = 0;
    {
        std::lock_guard<std::mutex> lk(g_rules_mu);
        if (!g_rules.empty()) {
            for (const mans_geometry_rule& r : g_rules)
                if (match(r)) { D = r.chunk_units; break; }
            return D / kLinQ * kLinQ;
        }
    }
    for (const mans_geometry_rule& r : kTunedRules)
        if (match(r)) { D = r.chunk_units; break; }
    return D / kLinQ * kLinQ;
}

// Spec warm-up (units decoded before each chunk, a multiple of K).  Built-in choice (no
// rule gives one): a stream of one 256-chunk range gets 512 units, larger ones none.  In
// a stream of one range the walks then mostly merge at the chunk start, and the range's
// walk phase (which lasts as long as its longest walk) shrinks more than the spec pass
// grows (10,000-stream mix: 43.4 -> 40.6 ms).  The autotuner's rules set it per bucket.
static uint32_t choose_warm(uint64_t n_units, uint32_t S, uint32_t K, bool index_in,
                            uint32_t rule_W) {
    if (index_in) return 0u;                  // the specs start at recorded sync points
    if (rule_W != MANS_WARM_AUTO) return std::min<uint32_t>(rule_W, 1u << 16) / K * K;
    const uint64_t chunks = (n_units + S - 1u) / S;
    return chunks <= 256u ? (512u / K) * K : 0u;
}

// Chunk/checkpoint lengths of one stream.  A stream with sync points uses the chunk
// length they were recorded with, as given (one record per chunk of that length).
static int stream_geometry(uint64_t n_units, uint64_t n_out, uint32_t L, uint32_t chunk_units,
                           uint32_t checkpoint_units, const void* index_in,
                           const void* index_out, uint32_t scope, uint32_t* S, uint32_t* K,
                           uint32_t* W);

// scope: MANS_RULE_SINGLE or MANS_RULE_BATCH (which rules apply); *W: the rule's warm-up
// (MANS_WARM_AUTO when none), may be NULL.
void choose_geometry(uint64_t n_units, uint64_t n_out, uint32_t L, uint32_t req_S,
                     uint32_t req_K, uint32_t* S, uint32_t* K, uint32_t scope = MANS_RULE_BATCH,
                     uint32_t* W = nullptr) {
    // checkpoints sit at round boundaries of the sync kernels: K is the round length
    (void)req_K;
    const uint32_t k = kR;
    uint32_t s = req_S;
    uint32_t w = MANS_WARM_AUTO;
    if (!s) {
        const RuleHit h = rule_geometry(n_units, n_out, L, scope);
        s = h.S;
        if (s) w = h.W;
    }
    if (!s) {
        // s_safe: long against the distance over which a guessed-state decode merges
        // with the true path, so that walks rarely fail.  Measured at L = 4096 (units,
        // 99.9th percentile): 851 at 4.3 symbols/unit, 1832 at 5.5, ~4000-4600 at
        // 8.5-21, ~600 at 100+.
        const double spu = n_units ? (double)n_out / (double)n_units : 4.0;
        uint32_t s_safe = spu <= 4.6 ? 2048u : (spu <= 6.5 ? 4096u : (spu <= 40.0 ? 8192u : 2048u));
        if (L <= 1024u) s_safe >>= 1;
        // s_fill: enough chunks to fill one 256-lane workgroup; a short stream trades
        // failed walks (fixup rounds) for lanes, which lowers its latency
        uint64_t fill = (n_units + 255u) / 256u;
        uint32_t s_fill = 1u;
        while (s_fill < fill && s_fill < (1u << 30)) s_fill <<= 1;
        // s_floor: below about the mean merge distance, failed walks chain through
        // fixup rounds and cost more than the lanes they free
        uint32_t s_floor = spu <= 4.6 ? 128u : (spu <= 6.5 ? 256u : 512u);
        s = std::max<uint32_t>(s_floor, std::min(s_safe, s_fill));
    }
    // a chunk longer than the stream gains nothing (and would waste checkpoint slots)
    uint32_t cap = k;
    while (cap < n_units && cap < (1u << 30)) cap <<= 1;
    if (s > cap) s = cap;
    s = 1u << ilog2(s < k ? k : s);
    *S = s;
    *K = k;
    if (W) *W = w;
}

static int stream_geometry(uint64_t n_units, uint64_t n_out, uint32_t L, uint32_t chunk_units,
                           uint32_t checkpoint_units, const void* index_in,
                           const void* index_out, uint32_t scope, uint32_t* S, uint32_t* K,
                           uint32_t* W) {
    *W = MANS_WARM_AUTO;
    if (!index_in && !index_out) {
        choose_geometry(n_units, n_out, L, chunk_units, checkpoint_units, S, K, scope, W);
        return MANS_OK;
    }
    if (chunk_units < kR || chunk_units > (1u << 24) || (chunk_units & (chunk_units - 1u)))
        return set_error(MANS_EINVAL,
                         "chunk_units=%u: sync points need a power of two in [%u, 2^24]",
                         chunk_units, kR);
    if ((((uintptr_t)index_in) | ((uintptr_t)index_out)) & 7u)
        return set_error(MANS_EINVAL, "sync point arrays must be 8-byte aligned");
    *S = chunk_units;
    *K = kR;
    return MANS_OK;
}

// Checkpoint slots of a stream (see ck_slot): n_chunks * S / K.
static uint64_t ck_slots(uint64_t n_chunks, uint32_t S, uint32_t K) { return n_chunks * (S / K); }

// Workspace layout for a set of plans (shared by single and batch paths).
struct Layout {
    uint64_t chunks = 0, cks = 0, bytes = 0;
    uint64_t o_exitv, o_own, o_went, o_wexit, o_wtot, o_mq, o_wcm, o_ff, o_coff;
    uint64_t o_ck, o_ckc, o_wb, o_wbc, o_status, o_fl, o_fl2, o_tick, o_plans, o_bsum, o_fcnt;
};

static uint64_t align256(uint64_t v) { return (v + 255u) & ~255ull; }

Layout make_layout(uint64_t chunks, uint64_t cks, uint64_t n_plans, uint32_t R = kWG) {
    Layout l;
    l.chunks = chunks;
    l.cks = cks;
    uint64_t o = 0;
    auto take = [&](uint64_t bytes) { uint64_t r = o; o = align256(o + bytes); return r; };
    // first: mans_decode_status reads offset 0; words 0-63 are cleared by k_init, the
    // lineage control words (kLinCtl..) follow and survive it
    l.o_status = take(4 * 128);
    l.o_exitv = take(4 * chunks);
    l.o_own = take(4 * chunks);
    l.o_went = take(4 * chunks);
    l.o_wexit = take(4 * chunks);
    l.o_wtot = take(4 * chunks);
    l.o_mq = take(4 * chunks);
    l.o_wcm = take(4 * chunks);
    l.o_ff = take(4 * chunks);
    l.o_coff = take(8 * chunks);
    l.o_ck = take(4 * cks);
    l.o_ckc = take(4 * cks);
    l.o_wb = take(4 * cks);
    l.o_wbc = take(4 * cks);
    l.o_fl = take(4 * chunks + 4);
    l.o_fl2 = take(4 * chunks + 4);
    l.o_tick = take(4 * 8);
    l.o_bsum = take(8 * (chunks / 2048 + 2));
    l.o_fcnt = take(4 * (n_plans + 1));
    // plans, then the chunk-task order (at most chunks/R + n_plans tasks, R chunks a task)
    l.o_plans = take(sizeof(StreamPlan) * n_plans + 4 * (chunks / R + 2 * n_plans + 1));
    l.bytes = o;
    return l;
}

Work make_work(void* base, const Layout& l) {
    char* b = (char*)base;
    Work w;
    w.exitv = (uint32_t*)(b + l.o_exitv);
    w.own_total = (uint32_t*)(b + l.o_own);
    w.wentry = (uint32_t*)(b + l.o_went);
    w.wexit = (uint32_t*)(b + l.o_wexit);
    w.wtotal = (uint32_t*)(b + l.o_wtot);
    w.mq = (uint32_t*)(b + l.o_mq);
    w.wcm = (uint32_t*)(b + l.o_wcm);
    w.fflag = (uint32_t*)(b + l.o_ff);
    w.coff = (uint64_t*)(b + l.o_coff);
    w.ck = (uint32_t*)(b + l.o_ck);
    w.ckc = (uint32_t*)(b + l.o_ckc);
    w.wb = (uint32_t*)(b + l.o_wb);
    w.wbc = (uint32_t*)(b + l.o_wbc);
    w.status = (uint32_t*)(b + l.o_status);
    w.flist = (uint32_t*)(b + l.o_fl);
    w.flist2 = (uint32_t*)(b + l.o_fl2);
    w.ctick = (uint32_t*)(b + l.o_tick);
    w.bsum = (uint64_t*)(b + l.o_bsum);
    w.fcnt = (uint32_t*)(b + l.o_fcnt);
    return w;
}

// Workspace layout of a lineage decode.  The status block sits at offset 0 as in the
// chunked layout (same words, same place), so the gate word (a control word after the
// 64 words k_init clears) is the same word for both pipelines.
struct LinLayout {
    uint64_t n_q = 0, n_c = 0, bytes = 0;
    uint64_t o_tick, o_bsum, o_coff, o_count, o_entry, o_mslot, o_msteps, o_cnt, o_val, o_inf,
        o_rec;
};

static LinLayout lin_layout(uint64_t n_units) {
    LinLayout l;
    l.n_q = (n_units + kLinQ - 1u) / kLinQ;
    l.n_c = (l.n_q + kLinChunk - 1u) / kLinChunk;
    uint64_t o = 0;
    auto take = [&](uint64_t bytes) { uint64_t r = o; o = align256(o + bytes); return r; };
    take(4 * 128);                          // status and control words
    l.o_tick = take(4 * 8);
    l.o_bsum = take(8 * (l.n_c / kScanBlock + 2));
    l.o_coff = take(8 * l.n_c);
    l.o_count = take(8 * l.n_c);
    l.o_entry = take(4 * l.n_c);
    l.o_mslot = take(4 * l.n_c * kLinSlots);
    l.o_msteps = take(8 * l.n_c * kLinSlots);
    l.o_cnt = take(4 * l.n_q);
    l.o_val = take(4 * l.n_q * kLinSlots);
    l.o_inf = take(4 * l.n_q * kLinSlots);
    l.o_rec = take(sizeof(SyncPoint) * l.n_q);
    l.bytes = o;
    return l;
}

static int launch_pipeline(KArgs a, uint32_t n_chunk_tasks, uint32_t n_item_tasks,
                           uint32_t total_chunks, uint32_t maxL, uint32_t maxK, hipStream_t st,
                           void** ev = nullptr) {
    size_t lds = (size_t)table_words(maxL) * 4u;
    size_t lds_s = lds + (size_t)kWG * kLS * 4u;   // table + round buffer
    const uint32_t R = range_of(a);                // chunks per sync task = k_sync's lanes
    const size_t lds_r = lds + (size_t)R * kLS * 4u;
    auto mark = [&](int i) { if (ev) (void)hipEventRecord((hipEvent_t)ev[i], st); };
    a.wait_polls = g_wait_polls.load();
    a.round_limit = g_round_limit.load();
    mark(0);
    a.n_scan_blocks = (total_chunks + kScanBlock - 1) / kScanBlock;
    hipLaunchKernelGGL(k_init,
                       dim3((std::max({a.n_streams, a.n_tasks, a.n_scan_blocks, 64u}) + 255u) / 256u),
                       dim3(256), 0, st, a, 64u);
    mark(1);
    if (a.direct) {                         // trusted item records: the write pass alone
        for (int i = 2; i <= 5; ++i) mark(i);
        const uint32_t per = kWritePer;
        const uint32_t nwg = std::min<uint32_t>((n_item_tasks + per - 1) / per, 2048u);
        const size_t lds_w = lds + (size_t)kWG * kWL * 4u;
        hipLaunchKernelGGL(k_write, dim3(nwg), dim3(kWG), lds_w, st, a, n_item_tasks, per);
        mark(6);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return set_error(MANS_EHIP, "kernel launch: %s", hipGetErrorString(e));
        return MANS_OK;
    }
    if (R == 64u)
        hipLaunchKernelGGL(k_sync<64>, dim3(n_chunk_tasks), dim3(64), lds_r, st, a);
    else if (R == 128u)
        hipLaunchKernelGGL(k_sync<128>, dim3(n_chunk_tasks), dim3(128), lds_r, st, a);
    else
        hipLaunchKernelGGL(k_sync<kWG>, dim3(n_chunk_tasks), dim3(kWG), lds_s, st, a);
    mark(2);
    hipLaunchKernelGGL(k_check, dim3(n_chunk_tasks), dim3(R), 0, st, a);
    mark(3);
    hipLaunchKernelGGL(k_fixup, dim3(a.n_streams), dim3(kWG), lds_s, st, a);
    mark(4);
    {
        const uint32_t nb = (total_chunks + kScanBlock - 1) / kScanBlock;
        hipLaunchKernelGGL(k_scan, dim3(nb), dim3(kWG), 0, st, a, total_chunks, a.w.bsum);
    }
    if (a.any_index_out) hipLaunchKernelGGL(k_index, dim3(n_chunk_tasks), dim3(R), 0, st, a);
    if (a.ref_sync) {
        const uint32_t lanes = std::max(a.one.n_ck, a.ref_nblocks);
        hipLaunchKernelGGL(k_refsync, dim3((lanes + kWG - 1u) / kWG), dim3(kWG), 0, st, a);
    }
    mark(5);
    {
        // persistent write pass: workgroups take blocks of kWritePer passes from a device
        // counter
        const uint32_t per = kWritePer;
        uint32_t nwg = std::min<uint32_t>((n_item_tasks + per - 1) / per, 2048u);
        (void)maxK;
        const size_t lds_w = lds + (size_t)kWG * kWL * 4u;
        hipLaunchKernelGGL(k_write, dim3(nwg), dim3(kWG), lds_w, st, a, n_item_tasks, per);
    }
    mark(6);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(MANS_EHIP, "kernel launch: %s", hipGetErrorString(e));
    return MANS_OK;
}

static int check_stream_args(const uint32_t* units, uint64_t n_units, const uint8_t* out,
                             uint64_t n_out, const uint32_t* table, uint32_t L,
                             uint32_t first_state, uint32_t first_bit) {
    if (!table) return set_error(MANS_EINVAL, "table is NULL");
    if (L < 2u || L > 16384u || (L & (L - 1u)))
        return set_error(MANS_EINVAL, "L=%u is not a power of two in [2, 16384]", L);
    if (n_out && (!units || !out)) return set_error(MANS_EINVAL, "units/out is NULL");
    if (n_out && n_units == 0) return set_error(MANS_EINVAL, "n_units is 0 but n_out > 0");
    if (((uintptr_t)units & 15u) != 0u)
        return set_error(MANS_EINVAL, "units must be 16-byte aligned (device allocations are)");
    // units must also be readable up to n_units rounded up to 4 (whole 16-byte loads)
    if (n_units >= 0xFFFFFF00ull) return set_error(MANS_EINVAL, "n_units too large");
    if (first_bit > 32u) return set_error(MANS_EINVAL, "first_bit=%u > 32", first_bit);
    if (n_out && (first_state < L || first_state >= 2u * L))
        return set_error(MANS_ECORRUPT, "first_state %u outside [L, 2L)", first_state);
    return MANS_OK;
}

}  // namespace mans

extern "C" int mans_workspace_bytes(uint64_t n_units, uint32_t L, uint64_t* bytes) {
    (void)L;
    if (!bytes) return set_error(MANS_EINVAL, "bytes is NULL");
    // smallest geometry the automatic choice uses: S = kMinRuleChunk; K (the round
    // length) >= 8
    uint64_t chunks = n_units / kMinRuleChunk + 2u, cks = n_units / 8u + 2u;
    *bytes = std::max(make_layout(chunks, cks, 1).bytes, lin_layout(n_units).bytes);
    return MANS_OK;
}

extern "C" int mans_workspace_bytes_geometry(uint64_t n_units, uint32_t chunk_units,
                                             uint32_t checkpoint_units, uint64_t* bytes) {
    if (!bytes || !chunk_units || !checkpoint_units) return set_error(MANS_EINVAL, "bad geometry");
    // the geometry mans_decode will use for these requests (checkpoints always sit at
    // round boundaries, kR units, whatever K is asked for)
    uint32_t S, K;
    choose_geometry(n_units, 1u, 4096u, chunk_units, checkpoint_units, &S, &K);
    uint64_t chunks = (n_units + S - 1u) / S + 2u, cks = ck_slots(chunks, S, K) + 2u;
    *bytes = make_layout(chunks, cks, 1).bytes;
    if (chunk_units >= kR && !(chunk_units & (chunk_units - 1u))) {
        // a decode with sync points uses chunk_units as given (not capped)
        chunks = (n_units + chunk_units - 1u) / chunk_units + 2u;
        cks = ck_slots(chunks, chunk_units, kR) + 2u;
        *bytes = std::max(*bytes, make_layout(chunks, cks, 1).bytes);
    }
    return MANS_OK;
}

// Plan of a single-stream decode (shared by mans_decode and mans_ref_sync_fill): the
// kernel arguments over the caller's workspace and the task counts.
static int single_plan(const mans_decode_args* d, KArgs* out, uint32_t* nct, uint32_t* nit) {
    int rc = check_stream_args(d->units, d->n_units, d->out, d->n_out, d->table, d->L,
                               d->first_state, d->first_bit);
    if (rc) return rc;
    uint32_t S, K, W;
    rc = stream_geometry(d->n_units, d->n_out, d->L, d->chunk_units, d->checkpoint_units,
                         d->index_in, d->index_out, MANS_RULE_SINGLE, &S, &K, &W);
    if (rc) return rc;
    if (K < 8u && !d->checkpoint_units) K = 8u;
    StreamPlan sp;
    std::memset(&sp, 0, sizeof(sp));
    sp.units = d->units;
    sp.out = d->out;
    sp.table = d->table;
    sp.n_out = d->n_out;
    sp.n_units = (uint32_t)d->n_units;
    sp.L = d->L;
    sp.first_state = d->first_state;
    sp.first_bit = d->first_bit;
    sp.S = S;
    sp.K = K;
    sp.log2K = ilog2(K);
    sp.warm = choose_warm(d->n_units, S, K, d->index_in != nullptr, W);
    sp.index_in = (const SyncPoint*)d->index_in;
    sp.index_out = (SyncPoint*)d->index_out;
    sp.items_out = (SyncPoint*)d->items_out;
    sp.n_chunks = (uint32_t)((d->n_units + S - 1) / S);
    sp.n_ck = (uint32_t)((d->n_units + K - 1) / K);
    sp.chunk_base = 0;
    sp.ck_base = 0;
    sp.task_chunk = 0;
    sp.task_item = 0;
    Layout l = make_layout(sp.n_chunks, ck_slots(sp.n_chunks, S, K), 1);
    if (!d->workspace || d->workspace_bytes < l.bytes)
        return set_error(MANS_ENOMEM, "workspace of %llu bytes < %llu needed",
                         (unsigned long long)d->workspace_bytes, (unsigned long long)l.bytes);
    if ((d->flags & MANS_F_TRUST_SYNC_POINTS) &&
        (!d->index_in || d->index_out || d->chunk_units != kWI))
        return set_error(MANS_EINVAL, "MANS_F_TRUST_SYNC_POINTS needs index_in at chunk_units "
                                      "%u (and no index_out)", kWI);
    if (((uintptr_t)d->items_out) & 7u)
        return set_error(MANS_EINVAL, "items_out must be 8-byte aligned");
    KArgs& a = *out;
    std::memset(&a, 0, sizeof(a));
    a.one = sp;
    a.plans = nullptr;
    a.n_streams = 1;
    a.any_index_out = d->index_out ? 1u : 0u;
    a.direct = (d->flags & MANS_F_TRUST_SYNC_POINTS) ? 1u : 0u;
    a.w = make_work(d->workspace, l);
    if (d->ref_sync) {
        const mans_ref_sync& r = *d->ref_sync;
        if (!r.sync_points || !r.subsequence_units || !r.threads || a.direct)
            return set_error(MANS_EINVAL, "ref_sync needs sync_points, subsequence_units and "
                                          "threads (and a proving decode)");
        if ((((uintptr_t)r.sync_points) | ((uintptr_t)r.output_sizes)) & 3u)
            return set_error(MANS_EINVAL, "ref_sync arrays must be 4-byte aligned");
        a.ref_sync = r.sync_points;
        a.ref_sizes = r.output_sizes;
        a.ref_flags = r.synced_flags;
        a.ref_S = r.subsequence_units;
        a.ref_nsub = (uint32_t)((d->n_units + r.subsequence_units - 1u) / r.subsequence_units);
        a.ref_nblocks = (a.ref_nsub + r.threads - 1u) / r.threads;
    }
    *nct = (sp.n_chunks + kWG - 1) / kWG;
    *nit = (n_write_items(sp.n_ck) + kWG - 1) / kWG;
    a.n_tasks = *nct;
    return MANS_OK;
}

// ------------------------------------------------------------ lineage decodes
static std::atomic<uint32_t> g_lin_cap{2048u};

// Lineage spacing for this decode, or 0 for the chunked pipeline: a lineage rule matches,
// the geometry is automatic, nothing asks for chunk-based sync state, and the workspace
// holds the lineage layout.
static uint32_t lineage_choice(const mans_decode_args* d) {
    if (d->chunk_units || d->index_in || d->index_out || d->ref_sync ||
        (d->flags & MANS_F_TRUST_SYNC_POINTS))
        return 0u;
    const uint64_t n_q = (d->n_units + kLinQ - 1u) / kLinQ;
    if (n_q < 2u) return 0u;
    const uint32_t D = rule_lineage(d->n_units, d->n_out, d->L);
    if (!D || D > (1u << 24)) return 0u;
    if (!d->workspace || d->workspace_bytes < lin_layout(d->n_units).bytes) return 0u;
    return D;
}

// The lineage kernels, then the chunked pipeline gated on the failure word, then the
// trusted write pass gated on success.  a: the chunked plan of the same decode.
static int launch_lineage(const mans_decode_args* d, const KArgs& a, uint32_t nct, uint32_t nit,
                          uint32_t D, hipStream_t st, void** ev) {
    auto mark = [&](int i) { if (ev) (void)hipEventRecord((hipEvent_t)ev[i], st); };
    const LinLayout l = lin_layout(d->n_units);
    char* b = (char*)d->workspace;
    LinArgs la;
    std::memset(&la, 0, sizeof(la));
    la.lw.val = (uint32_t*)(b + l.o_val);
    la.lw.inf = (uint32_t*)(b + l.o_inf);
    la.lw.cnt = (uint32_t*)(b + l.o_cnt);
    la.lw.mslot = (uint32_t*)(b + l.o_mslot);
    la.lw.msteps = (uint64_t*)(b + l.o_msteps);
    la.lw.entry = (uint32_t*)(b + l.o_entry);
    la.lw.count = (uint64_t*)(b + l.o_count);
    la.lw.rec = (SyncPoint*)(b + l.o_rec);
    la.status = (uint32_t*)b;
    la.lw.ctl = la.status + kLinCtl;
    la.lw.n_q = (uint32_t)l.n_q;
    la.lw.n_c = (uint32_t)l.n_c;
    la.D = D;
    la.n_lin = (uint32_t)((d->n_units + D - 1u) / D);
    la.cap = std::max<uint32_t>(1u, g_lin_cap.load());
    la.n_scan_blocks = (uint32_t)((l.n_c + kScanBlock - 1u) / kScanBlock);
    la.ctick = (uint32_t*)(b + l.o_tick);
    la.bsum = (uint64_t*)(b + l.o_bsum);
    // the lineage path's own arguments for k_scan and the trusted k_write
    KArgs k = a;
    std::memset(&k.w, 0, sizeof(k.w));
    k.w.status = la.status;
    k.w.ctick = la.ctick;
    k.w.bsum = la.bsum;
    k.w.coff = (uint64_t*)(b + l.o_coff);
    k.one.index_in = la.lw.rec;
    k.direct = 1u;
    k.any_index_out = 0u;
    k.ref_sync = nullptr;
    k.lin_count = la.lw.count;
    k.gate = nullptr;
    k.n_scan_blocks = la.n_scan_blocks;
    mark(0);
    {
        const uint64_t nthr = std::max<uint64_t>({l.n_q, 128u, la.n_scan_blocks});
        hipLaunchKernelGGL(k_lin_init, dim3((uint32_t)((nthr + 255u) / 256u)), dim3(256), 0, st,
                           k, la);
    }
    mark(1);
    {
        const size_t lds_s = (size_t)table_words(a.one.L) * 4u + (size_t)kWG * kLS * 4u;
        const uint32_t nwg = std::min<uint32_t>((la.n_lin + kWG - 1u) / kWG, 1024u);
        hipLaunchKernelGGL(k_lin, dim3(nwg), dim3(kWG), lds_s, st, k, la);
    }
    mark(2);
    hipLaunchKernelGGL(k_lin_map, dim3((uint32_t)((l.n_c * kLinSlots + 255u) / 256u)), dim3(256),
                       0, st, la);
    hipLaunchKernelGGL(k_lin_resolve, dim3((uint32_t)((l.n_c + 255u) / 256u)), dim3(256), 0, st,
                       la);
    mark(3);
    hipLaunchKernelGGL(k_scan, dim3(la.n_scan_blocks), dim3(kWG), 0, st, k, (uint32_t)l.n_c,
                       la.bsum);
    hipLaunchKernelGGL(k_lin_trace, dim3((uint32_t)((l.n_c + 255u) / 256u)), dim3(256), 0, st, k,
                       la);
    mark(4);
    // the chunked pipeline, run only if the lineage records do not hold
    KArgs g = a;
    g.gate = la.lw.ctl + 1;
    g.gate_val = 1u;
    int rc = launch_pipeline(g, nct, nit, a.one.n_chunks, a.one.L, a.one.K, st, nullptr);
    if (rc) return rc;
    mark(5);
    {
        k.gate = la.lw.ctl + 1;
        k.gate_val = 0u;
        const uint32_t per = kWritePer;
        const uint32_t nwg = std::min<uint32_t>((nit + per - 1) / per, 2048u);
        const size_t lds_w = (size_t)table_words(a.one.L) * 4u + (size_t)kWG * kWL * 4u;
        hipLaunchKernelGGL(k_write, dim3(nwg), dim3(kWG), lds_w, st, k, nit, per);
    }
    mark(6);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(MANS_EHIP, "kernel launch: %s", hipGetErrorString(e));
    return MANS_OK;
}

extern "C" int mans_lineage_spacing(uint64_t n_units, uint64_t n_out, uint32_t L, uint32_t* D) {
    if (!D) return set_error(MANS_EINVAL, "D is NULL");
    *D = rule_lineage(n_units, n_out, L);
    return MANS_OK;
}

extern "C" int mans_set_lineage_cap(uint32_t checkpoints) {
    g_lin_cap.store(checkpoints ? checkpoints : 2048u);
    return MANS_OK;
}

extern "C" int mans_lineage_stats(void* workspace, void* stream, uint32_t* out4) {
    if (!workspace || !out4) return set_error(MANS_EINVAL, "NULL argument");
    const uint32_t* ctl = (const uint32_t*)workspace + kLinCtl;
    hipError_t e = hipMemcpyAsync(out4, ctl + 1, 16, hipMemcpyDeviceToHost, (hipStream_t)stream);
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) return set_error(MANS_EHIP, "lineage stats read: %s", hipGetErrorString(e));
    return MANS_OK;
}

extern "C" int mans_decode(const mans_decode_args* d) {
    if (!d) return set_error(MANS_EINVAL, "args is NULL");
    if (d->n_out == 0)
        return check_stream_args(d->units, d->n_units, d->out, d->n_out, d->table, d->L,
                                 d->first_state, d->first_bit);
    KArgs a;
    uint32_t nct, nit;
    int rc = single_plan(d, &a, &nct, &nit);
    if (rc) return rc;
    const uint32_t D = lineage_choice(d);
    if (D)
        rc = launch_lineage(d, a, nct, nit, D, (hipStream_t)d->stream, d->kernel_events);
    else
        rc = launch_pipeline(a, nct, nit, a.one.n_chunks, a.one.L, a.one.K,
                             (hipStream_t)d->stream, d->kernel_events);
    if (rc) return rc;
    if (d->flags & MANS_F_SYNC_ERRORS) {
        int status = 0;
        rc = mans_decode_status(d->workspace, d->stream, &status);
        if (rc) return rc;
        if (status) return set_error(status, "device reported status %d", status);
    }
    return MANS_OK;
}

// ------------------------------------------------------------ no-allocation decode
// mans_decode_noalloc: mans_decode with the workspace taken from a static device array
// instead of the caller's, for callers that cannot allocate.  The array (`scratch`, one
// copy per GPU, held until the process ends) is addressed from the host only through the
// pointer hipGetSymbolAddress returns for the current device (kept per device); the
// symbol itself is a host-side stand-in and is never passed to a kernel.  Every call on
// a GPU shares that GPU's array: calls that overlap on one GPU race.  k_init clears what
// the kernels expect cleared (status words, tickets, fixup counters) at the start of each
// call, as it does for a caller's workspace.
constexpr uint64_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static std::vector<void*> g_scratch_ptr;    // per device, nullptr until looked up

static int scratch_ptr(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_error(MANS_EHIP, "get device: %s", hipGetErrorString(e));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((size_t)dev >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(MANS_EHIP, "scratch address on device %d: %s", dev,
                             hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return MANS_OK;
}

extern "C" int mans_decode_noalloc(const mans_decode_args* d) {
    if (!d) return set_error(MANS_EINVAL, "args is NULL");
    if (d->workspace || d->workspace_bytes)
        return set_error(MANS_EINVAL, "mans_decode_noalloc uses its own scratch: pass no "
                                      "workspace");
    if (d->n_out == 0)
        return check_stream_args(d->units, d->n_units, d->out, d->n_out, d->table, d->L,
                                 d->first_state, d->first_bit);
    mans_decode_args a = *d;
    // the size this stream's decode needs (the same layout mans_decode checks)
    a.workspace = (void*)(uintptr_t)256u;
    a.workspace_bytes = ~0ull;
    KArgs k;
    uint32_t nct, nit;
    int rc = single_plan(&a, &k, &nct, &nit);
    if (rc) return rc;
    uint64_t need = 0;
    {
        const StreamPlan& sp = k.one;
        need = make_layout(sp.n_chunks, ck_slots(sp.n_chunks, sp.S, sp.K), 1).bytes;
    }
    if (need > kScratchBytes)
        return set_error(MANS_ENOMEM, "mans_decode_noalloc: this stream needs %llu bytes of "
                                      "scratch, the static scratch holds %llu (use mans_decode "
                                      "with a workspace)",
                         (unsigned long long)need, (unsigned long long)kScratchBytes);
    rc = scratch_ptr(&a.workspace);
    if (rc) return rc;
    a.workspace_bytes = kScratchBytes;
    rc = single_plan(&a, &k, &nct, &nit);
    if (rc) return rc;
    rc = launch_pipeline(k, nct, nit, k.one.n_chunks, k.one.L, k.one.K, (hipStream_t)d->stream,
                         d->kernel_events);
    if (rc) return rc;
    if (d->flags & MANS_F_SYNC_ERRORS) {
        int status = 0;
        rc = mans_decode_status(a.workspace, d->stream, &status);
        if (rc) return rc;
        if (status) return set_error(status, "device reported status %d", status);
    }
    return MANS_OK;
}

extern "C" int mans_ref_sync_fill(const mans_decode_args* d) {
    if (!d || !d->ref_sync) return set_error(MANS_EINVAL, "args or ref_sync is NULL");
    if (d->n_out == 0) return MANS_OK;
    KArgs a;
    uint32_t nct, nit;
    int rc = single_plan(d, &a, &nct, &nit);
    if (rc) return rc;
    // a decode with these arguments (and no ref_sync) ran by lineages when the rules chose
    // them; its item records hold unless it fell back (read on the device)
    mans_decode_args dd = *d;
    dd.ref_sync = nullptr;
    if (lineage_choice(&dd)) {
        const LinLayout l = lin_layout(d->n_units);
        a.lin_rec = (const SyncPoint*)((const char*)d->workspace + l.o_rec);
        a.lin_fail_word = (const uint32_t*)d->workspace + kLinCtl + 1u;
        a.lin_nq = (uint32_t)l.n_q;
    }
    const uint32_t lanes = std::max({a.one.n_ck, a.ref_nblocks, a.lin_nq});
    hipLaunchKernelGGL(k_refsync, dim3((lanes + kWG - 1u) / kWG), dim3(kWG), 0,
                       (hipStream_t)d->stream, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(MANS_EHIP, "kernel launch: %s", hipGetErrorString(e));
    return MANS_OK;
}

extern "C" int mans_decode_status(void* workspace, void* stream, int* status) {
    if (!workspace || !status) return set_error(MANS_EINVAL, "NULL argument");
    // status words sit at offset 0 of every layout
    const uint32_t* st = (const uint32_t*)workspace;
    uint32_t v = 0;
    hipError_t e = hipMemcpyAsync(&v, st, 4, hipMemcpyDeviceToHost, (hipStream_t)stream);
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) return set_error(MANS_EHIP, "status read: %s", hipGetErrorString(e));
    // bit 0: a rejected table; bit 1: sync rounds that did not converge (k_fixup); bit 2:
    // a scan look-back wait that ran out
    *status = (v & 1u) ? MANS_ETABLE : (v & 6u) ? MANS_ECORRUPT : 0;
    return MANS_OK;
}

// --------------------------------------------------------------------- batch
namespace mans {

// Plans for a batch: per-stream geometry, global chunk/checkpoint bases, task bases.
// R: chunks per sync task (range) of this pipeline.
static int build_plans(const mans_stream_desc* d, size_t n, std::vector<StreamPlan>& plans,
                       uint64_t* chunks, uint64_t* cks, uint32_t* chunk_tasks,
                       uint32_t* item_tasks, uint32_t* maxL, uint32_t* maxK, uint64_t* units,
                       uint64_t* syms, uint32_t R = kWG) {
    plans.resize(n);
    uint64_t cb = 0, qb = 0, tc = 0, ti = 0, tu = 0, ts = 0;
    uint32_t ml = 2, mk = 1;
    for (size_t i = 0; i < n; ++i) {
        int rc = check_stream_args(d[i].units, d[i].n_units, d[i].out, d[i].n_out, d[i].table,
                                   d[i].L, d[i].first_state, d[i].first_bit);
        if (rc) return set_error(rc, "stream %zu: %s", i, mans_last_error());
        StreamPlan& sp = plans[i];
        std::memset(&sp, 0, sizeof(sp));
        uint32_t S, K, W;
        rc = stream_geometry(d[i].n_units, d[i].n_out, d[i].L, d[i].chunk_units, 0,
                             d[i].index_in, d[i].index_out, MANS_RULE_BATCH, &S, &K, &W);
        if (rc) return set_error(rc, "stream %zu: %s", i, mans_last_error());
        if (K < 8u) K = 8u;
        // a small-stream part (R < kWG): chunks lengthened until the stream is one range
        if (R < (uint32_t)kWG && !d[i].index_in && !d[i].chunk_units)
            while ((d[i].n_units + S - 1u) / S > R) S *= 2u;
        sp.units = d[i].units;
        sp.out = d[i].out;
        sp.table = d[i].table;
        sp.n_out = d[i].n_out;
        sp.n_units = (uint32_t)d[i].n_units;
        sp.L = d[i].L;
        sp.first_state = d[i].first_state;
        sp.first_bit = d[i].first_bit;
        sp.S = S;
        sp.K = K;
        sp.warm = choose_warm(d[i].n_units, S, K, d[i].index_in != nullptr, W);
        sp.index_in = (const SyncPoint*)d[i].index_in;
        sp.index_out = (SyncPoint*)d[i].index_out;
        sp.items_out = (SyncPoint*)d[i].items_out;
        if (((uintptr_t)d[i].items_out) & 7u)
            return set_error(MANS_EINVAL, "stream %zu: items_out must be 8-byte aligned", i);
        sp.log2K = ilog2(K);
        // streams with nothing to decode keep one idle chunk/item so tasks stay valid
        uint64_t nu = d[i].n_out ? d[i].n_units : 1;
        sp.n_chunks = (uint32_t)((nu + S - 1) / S);
        sp.n_ck = d[i].n_out ? (uint32_t)((nu + K - 1) / K) : 0u;
        sp.chunk_base = (uint32_t)cb;
        sp.ck_base = qb;
        sp.task_chunk = (uint32_t)tc;
        sp.task_item = (uint32_t)ti;
        if (!d[i].n_out) sp.n_chunks = 0;
        cb += std::max<uint32_t>(sp.n_chunks, 1u);
        qb += std::max<uint64_t>(ck_slots(sp.n_chunks, S, K), 1u);
        tc += std::max<uint64_t>((sp.n_chunks + R - 1) / R, 1);
        ti += std::max<uint64_t>((n_write_items(sp.n_ck) + kWG - 1) / kWG, 1);
        tu += d[i].n_units;
        ts += d[i].n_out;
        ml = std::max(ml, sp.L);
        mk = std::max(mk, sp.K);
    }
    if (cb >= 0xFFFFFFF0ull || tc >= 0x7FFFFFFFull || ti >= 0x7FFFFFFFull)
        return set_error(MANS_EINVAL, "batch too large");
    *chunks = cb;
    *cks = qb;
    *chunk_tasks = (uint32_t)tc;
    *item_tasks = (uint32_t)ti;
    *maxL = ml;
    *maxK = mk;
    *units = tu;
    *syms = ts;
    return MANS_OK;
}

// Plans plus the chunk-task order (longest expected first) as one host image for the
// workspace's plan region.  A task's expected length is its chunks' serial decode plus a
// merge tail, in steps: (S + 2048) * symbols per unit.
static std::vector<char> plan_image(const std::vector<StreamPlan>& plans, uint32_t n_tasks) {
    const size_t n = plans.size();
    std::vector<std::pair<double, uint32_t>> cost;
    cost.reserve(n_tasks);
    for (size_t i = 0; i < n; ++i) {
        const StreamPlan& p = plans[i];
        const uint32_t nt = i + 1 < n ? plans[i + 1].task_chunk - p.task_chunk
                                      : n_tasks - p.task_chunk;
        const double spu = p.n_units ? (double)p.n_out / (double)p.n_units : 0.0;
        const double c = ((double)p.S + 2048.0) * spu;
        for (uint32_t k = 0; k < nt; ++k) cost.push_back({c, p.task_chunk + k});
    }
    std::stable_sort(cost.begin(), cost.end(),
                     [](const std::pair<double, uint32_t>& x, const std::pair<double, uint32_t>& y) {
                         return x.first > y.first;
                     });
    std::vector<char> img(sizeof(StreamPlan) * n + 4 * ((size_t)n_tasks + n));
    std::memcpy(img.data(), plans.data(), sizeof(StreamPlan) * n);
    uint32_t* ord = (uint32_t*)(img.data() + sizeof(StreamPlan) * n);
    for (uint32_t k = 0; k < n_tasks; ++k) ord[k] = cost[k].second;
    // streams in the same order (for the fixup kernel)
    std::vector<std::pair<double, uint32_t>> sc(n);
    for (size_t i = 0; i < n; ++i) {
        const StreamPlan& p = plans[i];
        const double spu = p.n_units ? (double)p.n_out / (double)p.n_units : 0.0;
        sc[i] = {((double)p.S + 2048.0) * spu, (uint32_t)i};
    }
    std::stable_sort(sc.begin(), sc.end(),
                     [](const std::pair<double, uint32_t>& x, const std::pair<double, uint32_t>& y) {
                         return x.first > y.first;
                     });
    for (size_t i = 0; i < n; ++i) ord[n_tasks + i] = sc[i].second;
    return img;
}

}  // namespace mans

// ------------------------------------------------------------ batch parts
// A batch of many streams is decoded as up to kMaxParts independent pipelines on
// concurrent HIP streams: each part (a share of the streams, balanced by symbol count)
// has its own plans, tickets and workspace region.  While one part's sync pass drains
// its last long workgroups, another part's write pass uses the free CUs.  A part's
// status words are folded into the first part's (read by mans_decode_status).
constexpr size_t kMaxParts = 3;
constexpr size_t kStreamsPerPart = 256;   // fewer streams per part: one pipeline

struct BatchPart {
    std::vector<mans_stream_desc> descs;
    std::vector<size_t> orig;            // caller index of each of descs
    std::vector<StreamPlan> plans;
    uint64_t chunks = 0, cks = 0, units = 0, syms = 0, off = 0;
    uint32_t ct = 0, it = 0, ml = 0, mk = 0;
    uint32_t R = kWG;                    // chunks per sync task (k_sync's workgroup size)
    uint32_t any_index_out = 0;
    uint32_t direct = 0;
    Layout l;
};

// Streams of at most 2 kSmallRange chunks (at their rule's chunk length) sync in
// workgroups of kSmallRange lanes, one range each (with chunks twice as long where
// needed): a workgroup then holds an LDS table and kSmallRange round segments instead of
// kWG, so more such streams sync at once on a CU.  0 = off.
#ifndef MANS_SMALL_RANGE
#define MANS_SMALL_RANGE 64
#endif
constexpr uint32_t kSmallRange = MANS_SMALL_RANGE;
constexpr size_t kSmallPartMin = 256;     // fewer small streams: no part of their own

static int plan_parts(const mans_stream_desc* d, size_t n, size_t k, std::vector<BatchPart>& parts,
                      uint64_t* bytes) {
    parts.assign(k, BatchPart());
    if (k == 1) {
        parts[0].descs.assign(d, d + n);
        for (size_t i = 0; i < n; ++i) parts[0].orig.push_back(i);
    } else {
        // the small streams (see kSmallRange), when there are enough of them, form part 0
        std::vector<char> small(n, 0);
        size_t n_small = 0;
        if (kSmallRange) {
            for (size_t i = 0; i < n; ++i) {
                uint32_t S = 0, K = 0, W = 0;
                if (!d[i].n_out || stream_geometry(d[i].n_units, d[i].n_out, d[i].L,
                                                   d[i].chunk_units, 0, d[i].index_in,
                                                   d[i].index_out, MANS_RULE_BATCH, &S, &K, &W))
                    continue;
                // at most twice as many chunks as a range: twice as long chunks make it one
                // (streams whose chunk length is given keep it)
                const uint32_t lim = d[i].index_in || d[i].chunk_units ? kSmallRange
                                                                        : 2u * kSmallRange;
                if ((d[i].n_units + S - 1u) / S <= lim) {
                    small[i] = 1;
                    ++n_small;
                }
            }
        }
        const size_t p0 = n_small >= kSmallPartMin && n - n_small >= kSmallPartMin ? 1 : 0;
        if (!p0) std::fill(small.begin(), small.end(), 0);
        else parts[0].R = kSmallRange;
        // longest processing time first: each other stream to the part with the fewest
        // symbols
        std::vector<size_t> idx(n);
        for (size_t i = 0; i < n; ++i) idx[i] = i;
        std::stable_sort(idx.begin(), idx.end(),
                         [&](size_t x, size_t y) { return d[x].n_out > d[y].n_out; });
        std::vector<uint64_t> load(k, 0);
        std::vector<std::vector<size_t>> mine(k);
        for (size_t i : idx) {
            if (small[i]) {
                mine[0].push_back(i);
                continue;
            }
            size_t p = p0;
            for (size_t q = p0 + 1; q < k; ++q)
                if (load[q] < load[p]) p = q;
            load[p] += d[i].n_out + 1u;
            mine[p].push_back(i);
        }
        for (size_t p = 0; p < k; ++p) {
            std::sort(mine[p].begin(), mine[p].end());
            for (size_t i : mine[p]) {
                parts[p].descs.push_back(d[i]);
                parts[p].orig.push_back(i);
            }
        }
    }
    uint64_t o = 0;
    for (BatchPart& bp : parts) {
        int rc = build_plans(bp.descs.data(), bp.descs.size(), bp.plans, &bp.chunks, &bp.cks,
                             &bp.ct, &bp.it, &bp.ml, &bp.mk, &bp.units, &bp.syms, bp.R);
        if (rc) return rc;
        for (const mans_stream_desc& x : bp.descs)
            if (x.index_out) bp.any_index_out = 1u;
        bp.l = make_layout(bp.chunks, bp.cks, bp.descs.size(), bp.R);
        bp.off = o;
        o = align256(o + bp.l.bytes);
    }
    *bytes = o;
    return MANS_OK;
}

// MANS_BATCH_TRUST_SYNC_POINTS: every stream must bring its sync points (and write none).
static int mark_direct(std::vector<BatchPart>& parts, const mans_stream_desc* d, size_t n,
                       uint32_t flags) {
    if (!(flags & MANS_BATCH_TRUST_SYNC_POINTS)) return MANS_OK;
    for (size_t i = 0; i < n; ++i)
        if (d[i].n_out && (!d[i].index_in || d[i].index_out || d[i].chunk_units != kWI))
            return set_error(MANS_EINVAL, "stream %zu: MANS_BATCH_TRUST_SYNC_POINTS needs "
                                          "index_in at chunk_units %u (and no index_out)", i, kWI);
    for (BatchPart& bp : parts) bp.direct = 1u;
    return MANS_OK;
}

// opts->stream_status (optional, device, one u32 per caller stream): each stream's plan
// points at its word (k_init zeroes it; the kernels OR the stream's errors into it).
static void attach_status(std::vector<BatchPart>& parts, const mans_batch_opts* opts) {
    if (!opts->stream_status) return;
    for (BatchPart& bp : parts)
        for (size_t i = 0; i < bp.plans.size(); ++i)
            bp.plans[i].status_out = opts->stream_status + bp.orig[i];
}

static size_t n_parts(size_t n) {
    return std::max<size_t>(1, std::min(kMaxParts, n / kStreamsPerPart));
}

// Fold a later part's status words into the first part's.
__global__ void k_status_merge(uint32_t* dst, const uint32_t* src) {
    if (threadIdx.x == 0) {
        dst[0] |= src[0];
        dst[3] += src[3];
        dst[4] = max(dst[4], src[4]);
        dst[5] += src[5];
        dst[6] = max(dst[6], src[6]);
        dst[7] += src[7];
    }
}

// Internal streams (default priority) and fork/join events for the parts of batches
// enqueued on one caller stream, and a ring of two pinned staging buffers for their plan
// uploads, created once per (caller stream, device).  With pinned staging the upload is
// an ordinary stream-ordered copy: the host does not wait for the previous decode.
struct PartStreams {
    hipStream_t caller = nullptr;
    int dev = 0;
    std::mutex mu;                        // one enqueue at a time per caller stream
    hipStream_t s[kMaxParts] = {};
    hipEvent_t fork = nullptr, join[kMaxParts] = {};
    char* pin[2] = {};
    size_t cap[2] = {};
    hipEvent_t copied[2] = {};
    int next = 0;
};
static std::mutex g_parts_mu;
static std::vector<std::unique_ptr<PartStreams>> g_parts;

// Creates the part streams of ps on first use: a decode that runs as one pipeline never
// needs them, and every stream takes a share of the process's hardware queues.
static int part_streams_ready(PartStreams* ps) {
    if (ps->s[0]) return MANS_OK;
    hipError_t e = hipSuccess;
    for (size_t p = 0; p < kMaxParts && e == hipSuccess; ++p) {
        e = hipStreamCreateWithFlags(&ps->s[p], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ps->join[p], hipEventDisableTiming);
    }
    if (e != hipSuccess) return set_error(MANS_EHIP, "part streams: %s", hipGetErrorString(e));
    return MANS_OK;
}

static int part_streams(hipStream_t caller, PartStreams** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_error(MANS_EHIP, "get device: %s", hipGetErrorString(e));
    std::lock_guard<std::mutex> lk(g_parts_mu);
    for (auto& ps : g_parts)
        if (ps->caller == caller && ps->dev == dev) {
            *out = ps.get();
            return MANS_OK;
        }
    std::unique_ptr<PartStreams> ps(new PartStreams());
    ps->caller = caller;
    ps->dev = dev;
    e = hipEventCreateWithFlags(&ps->fork, hipEventDisableTiming);
    for (int b = 0; b < 2 && e == hipSuccess; ++b)
        e = hipEventCreateWithFlags(&ps->copied[b], hipEventDisableTiming);
    if (e != hipSuccess) return set_error(MANS_EHIP, "part streams: %s", hipGetErrorString(e));
    *out = ps.get();
    g_parts.push_back(std::move(ps));
    return MANS_OK;
}

static KArgs part_args(const BatchPart& bp, void* workspace) {
    char* base = (char*)workspace + bp.off;
    StreamPlan* dplans = (StreamPlan*)(base + bp.l.o_plans);
    KArgs a;
    std::memset(&a, 0, sizeof(a));
    a.plans = dplans;
    a.order = (const uint32_t*)(dplans + bp.descs.size());
    a.sorder = a.order + bp.ct;
    a.n_streams = (uint32_t)bp.descs.size();
    a.n_tasks = bp.ct;
    a.any_index_out = bp.any_index_out;
    a.direct = bp.direct;
    a.range = bp.R;
    a.w = make_work(base, bp.l);
    return a;
}

// Enqueue the parts' pipelines: forked from `st` onto `ps`'s streams and joined back,
// then the status merge on `st`.
static int launch_parts(const std::vector<BatchPart>& parts, void* workspace, hipStream_t st,
                        const PartStreams& ps) {
    hipError_t e = hipEventRecord(ps.fork, st);
    // the last part first: the large streams' long sync tasks reach the device before the
    // small streams' part (part 0), whose short workgroups then fill in around them
    // (0.6% faster than the other order on the benchmark batch, seven runs each)
    for (size_t q = parts.size(); q-- > 0 && e == hipSuccess;) {
        const size_t p = q;
        const BatchPart& bp = parts[p];
        e = hipStreamWaitEvent(ps.s[p], ps.fork, 0);
        if (e != hipSuccess) break;
        int rc = launch_pipeline(part_args(bp, workspace), bp.ct, bp.it, (uint32_t)bp.chunks, bp.ml,
                                 bp.mk, ps.s[p], nullptr);
        if (rc) return rc;
        e = hipEventRecord(ps.join[p], ps.s[p]);
        if (e == hipSuccess) e = hipStreamWaitEvent(st, ps.join[p], 0);
    }
    uint32_t* s0 = (uint32_t*)((char*)workspace + parts[0].off + parts[0].l.o_status);
    for (size_t p = 1; p < parts.size() && e == hipSuccess; ++p) {
        const uint32_t* sp = (const uint32_t*)((char*)workspace + parts[p].off + parts[p].l.o_status);
        hipLaunchKernelGGL(k_status_merge, dim3(1), dim3(64), 0, st, s0, sp);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return set_error(MANS_EHIP, "part launch: %s", hipGetErrorString(e));
    return MANS_OK;
}

// Plan images of all parts, uploaded into their workspace regions: synchronously (ps ==
// nullptr) or as copies on st from the next pinned staging buffer of ps (reused once the
// copies from its previous use have run).
static int upload_parts(const std::vector<BatchPart>& parts, void* workspace, hipStream_t st,
                        PartStreams* ps) {
    std::vector<std::vector<char>> imgs;
    size_t total = 0;
    for (const BatchPart& bp : parts) {
        imgs.push_back(plan_image(bp.plans, bp.ct));
        total += (imgs.back().size() + 255u) & ~(size_t)255u;
    }
    hipError_t e = hipSuccess;
    if (!ps) {
        for (size_t p = 0; p < parts.size() && e == hipSuccess; ++p)
            e = hipMemcpy((char*)workspace + parts[p].off + parts[p].l.o_plans, imgs[p].data(),
                          imgs[p].size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) return set_error(MANS_EHIP, "plan upload: %s", hipGetErrorString(e));
        return MANS_OK;
    }
    const int b = ps->next;
    ps->next ^= 1;
    e = hipEventSynchronize(ps->copied[b]);          // previous copies from this buffer done
    if (e == hipSuccess && ps->cap[b] < total) {
        // grow geometrically: freeing pinned memory synchronises the device, so it must
        // stay rare on a stream that serves many batches
        const size_t cap = std::max<size_t>({total, 2 * ps->cap[b], (size_t)1 << 20});
        if (ps->pin[b]) e = hipHostFree(ps->pin[b]);
        ps->pin[b] = nullptr;
        ps->cap[b] = 0;
        if (e == hipSuccess) e = hipHostMalloc((void**)&ps->pin[b], cap, hipHostMallocDefault);
        if (e == hipSuccess) ps->cap[b] = cap;
    }
    size_t o = 0;
    for (size_t p = 0; p < parts.size() && e == hipSuccess; ++p) {
        std::memcpy(ps->pin[b] + o, imgs[p].data(), imgs[p].size());
        e = hipMemcpyAsync((char*)workspace + parts[p].off + parts[p].l.o_plans, ps->pin[b] + o,
                           imgs[p].size(), hipMemcpyHostToDevice, st);
        o += (imgs[p].size() + 255u) & ~(size_t)255u;
    }
    if (e == hipSuccess) e = hipEventRecord(ps->copied[b], st);
    if (e != hipSuccess) return set_error(MANS_EHIP, "plan upload: %s", hipGetErrorString(e));
    return MANS_OK;
}

extern "C" int mans_batch_workspace_bytes(const mans_stream_desc* descs, size_t n, uint64_t* bytes) {
    if (!descs || !n || !bytes) return set_error(MANS_EINVAL, "bad batch");
    // enough for the split decode and for the single pipeline (kernel timing)
    std::vector<BatchPart> parts;
    uint64_t b1 = 0, bk = 0;
    int rc = plan_parts(descs, n, 1, parts, &b1);
    if (!rc) rc = plan_parts(descs, n, n_parts(n), parts, &bk);
    if (rc) return rc;
    *bytes = std::max(b1, bk);
    return MANS_OK;
}

extern "C" int mans_decode_batch(const mans_stream_desc* descs, size_t n,
                                 const mans_batch_opts* opts, mans_batch_stats* stats) {
    if (!descs || !n || !opts) return set_error(MANS_EINVAL, "bad batch");
    // per-kernel timing events describe one pipeline, so a timed decode is not split
    const size_t k = (opts->kernel_events || (opts->flags & MANS_BATCH_ONE_PIPELINE)) ? 1 : n_parts(n);
    std::vector<BatchPart> parts;
    uint64_t need = 0;
    int rc = plan_parts(descs, n, k, parts, &need);
    if (!rc) rc = mark_direct(parts, descs, n, opts->flags);
    if (rc) return rc;
    if (((uintptr_t)opts->stream_status) & 3u)
        return set_error(MANS_EINVAL, "stream_status must be 4-byte aligned");
    attach_status(parts, opts);
    if (!opts->workspace || opts->workspace_bytes < need)
        return set_error(MANS_ENOMEM, "batch workspace of %llu bytes < %llu needed",
                         (unsigned long long)opts->workspace_bytes, (unsigned long long)need);
    hipStream_t st = (hipStream_t)opts->stream;
    PartStreams* ps = nullptr;
    rc = part_streams(st, &ps);
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> lk(ps->mu);
        rc = upload_parts(parts, opts->workspace, st, ps);
        if (rc) return rc;
        if (k == 1) {
            const BatchPart& bp = parts[0];
            rc = launch_pipeline(part_args(bp, opts->workspace), bp.ct, bp.it, (uint32_t)bp.chunks,
                                 bp.ml, bp.mk, st, opts->kernel_events);
        } else {
            rc = part_streams_ready(ps);
            if (!rc) rc = launch_parts(parts, opts->workspace, st, *ps);
        }
    }
    if (rc) return rc;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        for (const BatchPart& bp : parts) {
            stats->total_units += bp.units;
            stats->total_symbols += bp.syms;
            stats->n_chunks += (uint32_t)bp.chunks;
            stats->n_items += (uint32_t)bp.cks;
        }
        stats->n_streams = (uint32_t)n;
        stats->status = 0;
    }
    return MANS_OK;
}

// ------------------------------------------------------------ graphs and rules
namespace {
struct GraphHandle {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};
}  // namespace

extern "C" int mans_batch_graph_create(const mans_stream_desc* descs, size_t n,
                                       const mans_batch_opts* opts, void** graph) {
    if (!descs || !n || !opts || !graph) return set_error(MANS_EINVAL, "bad batch");
    *graph = nullptr;
    std::vector<BatchPart> parts;
    uint64_t need = 0;
    int rc = plan_parts(descs, n, n_parts(n), parts, &need);
    if (!rc) rc = mark_direct(parts, descs, n, opts->flags);
    if (rc) return rc;
    if (((uintptr_t)opts->stream_status) & 3u)
        return set_error(MANS_EINVAL, "stream_status must be 4-byte aligned");
    attach_status(parts, opts);
    if (!opts->workspace || opts->workspace_bytes < need)
        return set_error(MANS_ENOMEM, "batch workspace of %llu bytes < %llu needed",
                         (unsigned long long)opts->workspace_bytes, (unsigned long long)need);
    rc = upload_parts(parts, opts->workspace, nullptr, nullptr);
    if (rc) return rc;
    // capture on a private stream; the parts fork onto private streams inside the capture
    hipStream_t cs = nullptr;
    PartStreams ps;
    hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ps.fork, hipEventDisableTiming);
    for (size_t p = 0; p < parts.size() && e == hipSuccess; ++p) {
        e = hipStreamCreateWithFlags(&ps.s[p], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ps.join[p], hipEventDisableTiming);
    }
    // (the staging ring and mutex of ps are unused here)
    GraphHandle* g = new GraphHandle();
    if (e != hipSuccess) {
        rc = set_error(MANS_EHIP, "capture streams: %s", hipGetErrorString(e));
    } else {
        e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            if (parts.size() == 1) {
                const BatchPart& bp = parts[0];
                rc = launch_pipeline(part_args(bp, opts->workspace), bp.ct, bp.it,
                                     (uint32_t)bp.chunks, bp.ml, bp.mk, cs, nullptr);
            } else {
                rc = launch_parts(parts, opts->workspace, cs, ps);
            }
            hipError_t e2 = hipStreamEndCapture(cs, &g->graph);
            if (!rc && e2 != hipSuccess) rc = set_error(MANS_EHIP, "end capture: %s", hipGetErrorString(e2));
        } else {
            rc = set_error(MANS_EHIP, "begin capture: %s", hipGetErrorString(e));
        }
    }
    if (!rc) {
        e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) rc = set_error(MANS_EHIP, "instantiate: %s", hipGetErrorString(e));
    }
    for (size_t p = 0; p < kMaxParts; ++p) {
        if (ps.s[p]) (void)hipStreamDestroy(ps.s[p]);
        if (ps.join[p]) (void)hipEventDestroy(ps.join[p]);
    }
    if (ps.fork) (void)hipEventDestroy(ps.fork);
    if (cs) (void)hipStreamDestroy(cs);
    if (rc) {
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        delete g;
        return rc;
    }
    *graph = g;
    return MANS_OK;
}

extern "C" int mans_graph_launch(void* graph, void* stream) {
    GraphHandle* g = (GraphHandle*)graph;
    if (!g || !g->exec) return set_error(MANS_EINVAL, "graph is NULL");
    hipError_t e = hipGraphLaunch(g->exec, (hipStream_t)stream);
    if (e != hipSuccess) return set_error(MANS_EHIP, "graph launch: %s", hipGetErrorString(e));
    return MANS_OK;
}

extern "C" int mans_graph_destroy(void* graph) {
    GraphHandle* g = (GraphHandle*)graph;
    if (!g) return MANS_OK;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return MANS_OK;
}

extern "C" int mans_set_wait_polls(uint32_t polls) {
    g_wait_polls.store(polls);
    return MANS_OK;
}

uint32_t sync_round_limit() { return mans::g_round_limit.load(); }

extern "C" int mans_set_sync_round_limit(uint32_t rounds) {
    g_round_limit.store(rounds);
    return MANS_OK;
}

extern "C" int mans_set_geometry_rules(const mans_geometry_rule* rules, size_t n) {
    if (n && !rules) return set_error(MANS_EINVAL, "rules is NULL");
    std::vector<mans_geometry_rule> v;
    for (size_t i = 0; i < n; ++i) {
        mans_geometry_rule r = rules[i];
        if (r.scope == MANS_RULE_LINEAGE) {          // lineage spacing: a multiple of 32, or 0
            if (r.chunk_units > (1u << 24))
                return set_error(MANS_EINVAL, "rule %zu: lineage spacing %u above 2^24", i,
                                 r.chunk_units);
            r.chunk_units = r.chunk_units / kLinQ * kLinQ;
            v.push_back(r);
            continue;
        }
        if (r.chunk_units < kMinRuleChunk || r.chunk_units > (1u << 24))
            return set_error(MANS_EINVAL, "rule %zu: chunk_units %u outside [%u, 2^24]", i,
                             r.chunk_units, kMinRuleChunk);
        if (r.scope > (MANS_RULE_SINGLE | MANS_RULE_BATCH))
            return set_error(MANS_EINVAL, "rule %zu: scope %u", i, r.scope);
        uint32_t p = kMinRuleChunk;
        while (p * 2u <= r.chunk_units) p *= 2u;
        r.chunk_units = p;
        v.push_back(r);
    }
    std::lock_guard<std::mutex> lk(g_rules_mu);
    g_rules.swap(v);
    return MANS_OK;
}

extern "C" int mans_sync_point_count(uint64_t n_units, uint32_t chunk_units, uint64_t* count) {
    if (!count) return set_error(MANS_EINVAL, "count is NULL");
    if (chunk_units < kR || chunk_units > (1u << 24) || (chunk_units & (chunk_units - 1u)))
        return set_error(MANS_EINVAL, "chunk_units=%u is not a power of two in [%u, 2^24]",
                         chunk_units, kR);
    *count = (n_units + chunk_units - 1u) / chunk_units;
    return MANS_OK;
}

extern "C" int mans_sync_points_check(const mans_sync_point* ih, uint64_t count, uint32_t L) {
    if (!ih && count) return set_error(MANS_EINVAL, "NULL argument");
    if (L < 2u || L > 16384u || (L & (L - 1u)))
        return set_error(MANS_EINVAL, "L=%u is not a power of two in [2, 16384]", L);
    for (uint64_t c = 0; c < count; ++c) {
        const uint64_t prev = c ? ih[c - 1u].symbols_before : 0u;
        if (ih[c].state >= L || ih[c].bit > (c ? 31u : 32u) || ih[c].symbols_before < prev ||
            (c == 0u && ih[c].symbols_before != 0u))
            return set_error(MANS_ECORRUPT, "sync point %llu: state %u, bit %u, before %llu",
                             (unsigned long long)c, ih[c].state, ih[c].bit,
                             (unsigned long long)ih[c].symbols_before);
    }
    return MANS_OK;
}

extern "C" int mans_stream_range(const mans_stream_desc* f, const mans_sync_point* ih,
                                 const mans_sync_point* idev, uint64_t first, uint64_t last,
                                 mans_stream_desc* sub, uint64_t* skip) {
    if (!f || !ih || !idev || !sub || !skip) return set_error(MANS_EINVAL, "NULL argument");
    uint64_t n = 0;
    int rc = mans_sync_point_count(f->n_units, f->chunk_units, &n);
    if (rc) return rc;
    if (first >= last || last > f->n_out)
        return set_error(MANS_EINVAL, "window [%llu, %llu) is empty or beyond n_out=%llu",
                         (unsigned long long)first, (unsigned long long)last,
                         (unsigned long long)f->n_out);
    if (n == 0 || ih[0].symbols_before != 0u)
        return set_error(MANS_ECORRUPT, "sync points do not start at symbol 0");
    const uint64_t S = f->chunk_units;
    // ca: the last chunk starting at or before `first`; cb: the first chunk starting at or
    // after `last` (n if none).  Records are ascending in symbols_before.
    uint64_t lo = 0, hi = n - 1u;
    while (lo < hi) {
        const uint64_t mid = (lo + hi + 1u) >> 1;
        if (ih[mid].symbols_before <= first) lo = mid; else hi = mid - 1u;
    }
    const uint64_t ca = lo;
    lo = ca + 1u;
    hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (ih[mid].symbols_before >= last) hi = mid; else lo = mid + 1u;
    }
    const uint64_t cb = lo;
    const uint64_t before = ih[ca].symbols_before;
    const uint64_t end = cb < n ? ih[cb].symbols_before : f->n_out;
    // the records used must bracket the window (mans_sync_points_check validates all)
    if (before > first || (ca + 1u < n && ih[ca + 1u].symbols_before <= first) ||
        ih[cb - 1u].symbols_before >= last || end < last || end > f->n_out)
        return set_error(MANS_ECORRUPT, "sync points are not ascending within [0, n_out]");
    *sub = *f;
    if (ca > 0u) {
        if (ih[ca].state >= f->L || ih[ca].bit > 31u)
            return set_error(MANS_ECORRUPT, "sync point %llu out of range",
                             (unsigned long long)ca);
        sub->first_state = ih[ca].state + f->L;
        sub->first_bit = 32u - ih[ca].bit;
    }
    const uint64_t u0 = ca * S;
    // one unit past the last chunk: its last steps may read bits from there
    const uint64_t u1 = std::min<uint64_t>(cb * S + 1u, f->n_units);
    sub->units = f->units + u0;
    sub->n_units = u1 - u0;
    sub->n_out = end - before;
    sub->out = nullptr;
    sub->index_in = idev + ca;
    sub->index_out = nullptr;
    *skip = first - before;
    return MANS_OK;
}

extern "C" int mans_chunk_units(uint64_t n_units, uint64_t n_out, uint32_t L, uint32_t* chunk_units) {
    if (!chunk_units) return set_error(MANS_EINVAL, "chunk_units is NULL");
    uint32_t S, K;
    choose_geometry(n_units, n_out, L, 0, 0, &S, &K);
    *chunk_units = S;
    return MANS_OK;
}
