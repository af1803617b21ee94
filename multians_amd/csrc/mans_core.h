// mans_core.h -- per-lane decode logic shared by the gfx950 kernels and the CPU
// emulation build (tests only).
//
// Method (self-synchronising subsequence decoding, after the reference's
// src/multians_gpu_decoder.cu:16-573 and src/multicore_decoder.cc:92-273, re-planned
// for large sync chunks):
//
//   chunk c     = units [c*S, (c+1)*S) of one stream (S = chunk_units)
//   checkpoint q= unit boundary q*K (K = checkpoint_units, K | S); its value is
//                 the (state, bit offset) of the first decoding step whose start
//                 position p satisfies p >= 32*q*K ("entry" of the boundary).
//   spec pass   : lane c decodes chunk c from a guessed state at bit 0 of the chunk
//                 (chunk 0: the true start), storing every checkpoint of its path
//                 and the number of steps taken before it.
//   walk pass   : lane c walks chunk c again from the exit of chunk c-1's path and
//                 compares each checkpoint with the stored one; equality means the
//                 two paths have merged (decoding is deterministic in (p, state)),
//                 so everything stored from there on is on the walker's path.
//                 Checkpoints before the merge are stored separately (wb*).  A walk
//                 that crosses its chunk without merging continues into the next
//                 chunks (lane_walk_chain) when its own start is likely proven.
//   check       : every chunk whose recorded walk start differs from its
//                 predecessor's exit is listed.
//   fixup       : listed chunks are walked again from their predecessor's exit, in
//                 rounds, until every chunk is consistent; by induction from chunk 0
//                 every exit is then on the true path.
//   write pass  : every checkpoint is now known on the true path together with its
//                 output offset (exclusive scan of per-chunk step counts), so item
//                 q decodes exactly off(q+1)-off(q) steps from checkpoint q and
//                 writes them.  The last item decodes up to n_out, so zero-bit
//                 trailing symbols are produced (the reference loses them when
//                 n_units % S == 0).
//
// All arithmetic is unsigned integer.  A decoding step (multians_gpu_decoder.cu:131-166)
// for a table entry e = {next, symbol, nbits}:
//     state' = (next << nbits) | bits(p, nbits);   p += nbits
// Entries are staged into a derived word:
//     bits 0-4 nbits | bits 8-15 symbol | bits 16-31 (next << nbits) - L
// so that state_index' = (w >> 16) + bits(p, w & 31), state_index = state - L.
#ifndef MANS_CORE_H_
#define MANS_CORE_H_

#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define MANS_HD __host__ __device__ __forceinline__
#else
#define MANS_HD inline
#endif

namespace mans {

constexpr uint32_t kNone = 0xFFFFFFFFu;

// Packed boundary entry: state index (x - L) in bits 0-15, bit offset 0..32 in 16-21.
MANS_HD uint32_t pack_entry(uint32_t xi, uint32_t off) { return xi | (off << 16); }
MANS_HD uint32_t entry_state(uint32_t v) { return v & 0xFFFFu; }
MANS_HD uint32_t entry_off(uint32_t v) { return (v >> 16) & 63u; }

// Derived LDS table word from a reference CUHDCodetableItem word.  Returns 0 and
// sets *bad when the entry cannot be decoded by the fast step: the next state must
// stay inside [L, 2L) for every value of the nbits read, and a zero-bit entry must
// strictly decrease the state (so zero-bit runs are finite).
// A zero-bit entry that does not decrease the state (a table with a symbol owning all
// L states has only such entries) sets *cyc: position-driven decoding could then loop
// forever, so those tables are only decoded when the stream has no coded bits beyond
// its first checkpoint interval (count-driven write pass only).
MANS_HD uint32_t derive_entry(uint32_t e, uint32_t L, uint32_t index, int* bad, int* cyc) {
    uint32_t next = e & 0xFFFFu, sym = (e >> 16) & 0xFFu, nb = e >> 24;
    if (nb > 15u || next == 0u) { *bad = 1; return 0u; }
    uint32_t lo = next << nb, hi = (next + 1u) << nb;
    if (lo < L || hi > 2u * L) { *bad = 1; return 0u; }
    if (nb == 0u && lo - L >= index) *cyc = 1;
    return nb | (sym << 8) | ((lo - L) << 16);
}

MANS_HD uint32_t funnel_shr(uint32_t hi, uint32_t lo, uint32_t s) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_alignbit(hi, lo, s);
#else
    s &= 31u;
    return s ? (lo >> s) | (hi << (32u - s)) : lo;
#endif
}

MANS_HD uint32_t low_bits(uint32_t w, uint32_t word) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_ubfe(w, 0u, word);   // width = word[4:0]
#else
    uint32_t n = word & 31u;
    return n ? (w & (0xFFFFFFFFu >> (32u - n))) : 0u;
#endif
}

// Bit cursor over one stream: lo/hi are units u and u+1, nx is unit u+2 (prefetched),
// o in [0, 32) is the bit offset inside lo.  Units at or beyond n read as zero.
struct Cursor {
    const uint32_t* units;
    uint32_t n;
    uint32_t u;
    uint32_t o;
    uint32_t lo, hi, nx;
};

#if defined(__HIP_DEVICE_COMPILE__)
typedef const uint4 __attribute__((address_space(1)))* gptr_u4;
typedef const uint32_t __attribute__((address_space(1)))* gptr_u32;
#endif

// Unit i of a stream, zero at or beyond n.  Branch-free: the address is clamped to a
// valid unit and the value masked.
MANS_HD uint32_t load_unit(const uint32_t* units, uint32_t n, uint32_t i) {
    const uint32_t ia = i < n ? i : 0u;
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t v = ((gptr_u32)units)[ia];
#else
    const uint32_t v = units[ia];
#endif
    return i < n ? v : 0u;
}

// Position the cursor at absolute bit (32*unit + off), off in [0, 32].
MANS_HD void cursor_init(Cursor& c, const uint32_t* units, uint32_t n, uint32_t unit,
                         uint32_t off) {
    c.units = units;
    c.n = n;
    c.u = unit + (off >> 5);
    c.o = off & 31u;
    c.lo = load_unit(units, n, c.u);
    c.hi = load_unit(units, n, c.u + 1u);
    c.nx = load_unit(units, n, c.u + 2u);
}

// One decoding step on state index xi with derived table tab.  Returns the derived
// word (symbol in bits 8-15).  Advances the cursor; returns crossed=1 when the next
// step starts in the following unit.
template <typename Tab>
MANS_HD uint32_t step(const Tab& tab, Cursor& c, uint32_t& xi, uint32_t& crossed) {
    uint32_t w = tab[xi];
    uint32_t win = funnel_shr(c.hi, c.lo, c.o);
    xi = (w >> 16) + low_bits(win, w);
    c.o += w & 31u;
    crossed = 0u;
    if (c.o >= 32u) {
        c.o -= 32u;
        c.u += 1u;
        c.lo = c.hi;
        c.hi = c.nx;
        c.nx = load_unit(c.units, c.n, c.u + 2u);
        crossed = 1u;
    }
    return w;
}

// Cursor for long sequential walks (spec and walk passes): units arrive in aligned
// blocks of four (one 16-byte load).  Two block registers A and B alternate: the
// cursor reads units from one while the other's load is in flight, and a block is
// reloaded in place once consumed (no register copies of data still in flight, which
// would make the prefetch wait).
struct U4 { uint32_t x, y, z, w; };

// Aligned block of four units starting at b (b % 4 == 0), zeros at or beyond n.
// Branch-free (clamped address, masked values).  Reads whole 16-byte blocks, so the
// buffer must be readable up to n rounded up to a multiple of four units.
MANS_HD U4 load_block(const uint32_t* units, uint32_t n, uint32_t b) {
    const uint32_t ba = b < n ? b : 0u;
    U4 r;
#if defined(__HIP_DEVICE_COMPILE__)
    const uint4 v = ((gptr_u4)units)[ba >> 2];
    r.x = v.x; r.y = v.y; r.z = v.z; r.w = v.w;
#else
    r.x = units[ba]; r.y = ba + 1u < n ? units[ba + 1u] : 0u;
    r.z = ba + 2u < n ? units[ba + 2u] : 0u; r.w = ba + 3u < n ? units[ba + 3u] : 0u;
#endif
    r.x = b < n ? r.x : 0u;
    r.y = b + 1u < n ? r.y : 0u;
    r.z = b + 2u < n ? r.z : 0u;
    r.w = b + 3u < n ? r.w : 0u;
    return r;
}

// The same block with the address clamped into the stream but the values not masked, for
// prefetches: units past the end are only ever read as the unused top of a window (a
// stream's codes end at bit 32 n, and its true path consumes no bit beyond them; any
// other path is discarded whatever it reads), and masking the values would make the
// prefetch wait for its data as soon as it is issued.
MANS_HD U4 load_block_raw(const uint32_t* units, uint32_t n, uint32_t b) {
    const uint32_t ba = b < n ? b : 0u;
    U4 r;
#if defined(__HIP_DEVICE_COMPILE__)
    const uint4 v = ((gptr_u4)units)[ba >> 2];
    r.x = v.x; r.y = v.y; r.z = v.z; r.w = v.w;
#else
    r.x = units[ba]; r.y = ba + 1u < n ? units[ba + 1u] : 0u;
    r.z = ba + 2u < n ? units[ba + 2u] : 0u; r.w = ba + 3u < n ? units[ba + 3u] : 0u;
#endif
    return r;
}

MANS_HD uint32_t pick4(const U4& v, uint32_t k) {
    return k == 0u ? v.x : (k == 1u ? v.y : (k == 2u ? v.z : v.w));
}

struct BCursor {
    const uint32_t* units;
    uint32_t n;
    uint32_t u;       // unit of lo
    uint32_t o;       // bit offset in lo
    uint32_t lo, hi;
    uint32_t k;       // position of unit u+2 in the 8-unit ring A|B (0..7)
    uint32_t nb;      // unit index of the next block to load
    U4 A, B;
};

// Position at absolute bit 32*unit + off (off in [0, 32]).
MANS_HD void bcursor_init(BCursor& c, const uint32_t* units, uint32_t n, uint32_t unit, uint32_t off) {
    c.units = units;
    c.n = n;
    c.u = unit + (off >> 5);
    c.o = off & 31u;
    const uint32_t bb = c.u & ~3u;            // A = block of u, B = the next block
    c.A = load_block(units, n, bb);
    c.B = load_block(units, n, bb + 4u);
    const uint32_t k0 = c.u - bb;             // 0..3
    c.lo = pick4(c.A, k0);
    c.hi = k0 == 3u ? c.B.x : pick4(c.A, k0 + 1u);
    c.k = k0 + 2u;
    c.nb = bb + 8u;
    if (k0 >= 2u) {                           // A is already consumed: refill it now
        c.A = load_block(units, n, c.nb);
        c.nb += 4u;
    }
}

MANS_HD void bcursor_advance(BCursor& c) {
    c.u += 1u;
    c.lo = c.hi;
    const uint32_t k = c.k;
    c.hi = k < 4u ? pick4(c.A, k) : pick4(c.B, k - 4u);
    c.k = (k + 1u) & 7u;
    // a block is reloaded once its last unit has moved into hi
    if (k == 3u) { c.A = load_block(c.units, c.n, c.nb); c.nb += 4u; }
    if (k == 7u) { c.B = load_block(c.units, c.n, c.nb); c.nb += 4u; }
}

template <typename Tab>
MANS_HD uint32_t bstep(const Tab& tab, BCursor& c, uint32_t& xi, uint32_t& crossed) {
    uint32_t w = tab[xi];
    uint32_t win = funnel_shr(c.hi, c.lo, c.o);
    xi = (w >> 16) + low_bits(win, w);
    c.o += w & 31u;
    crossed = 0u;
    if (c.o >= 32u) {
        c.o -= 32u;
        bcursor_advance(c);
        crossed = 1u;
    }
    return w;
}

// ------------------------------------------------------------------ layouts
// Sync point record (layout of the C ABI's mans_sync_point): the true (state, bit
// offset) at a chunk start and the symbols decoded before it.
struct SyncPoint {
    uint32_t state;
    uint32_t bit;
    uint64_t before;
};

// Per-stream parameters (device copy of one mans_stream_desc plus the plan).
struct StreamPlan {
    const uint32_t* units;
    uint8_t* out;
    const uint32_t* table;
    uint64_t n_out;
    uint32_t n_units;
    uint32_t L;
    uint32_t first_state;
    uint32_t first_bit;
    uint32_t S;          // chunk units (multiple of K, power of two)
    uint32_t K;          // checkpoint units (power of two)
    uint32_t log2K;
    uint32_t n_chunks;   // ceil(n_units / S)
    uint32_t n_ck;       // ceil(n_units / K)
    uint32_t chunk_base; // global index of chunk 0
    uint64_t ck_base;    // global index of checkpoint 0
    uint32_t task_chunk; // first spec/walk task
    uint32_t task_item;  // first write task
    uint32_t warm;       // spec warm-up: units decoded before each chunk c >= 1 (multiple
                         // of K), so the spec path is likely on the true one at the chunk
                         // start; 0 = none
    uint32_t pad;
    const SyncPoint* index_in;   // optional: spec of chunk c >= 1 starts at index_in[c]
    SyncPoint* index_out;        // optional: true sync points written after the scan
    SyncPoint* items_out;        // optional: a record per write item (32 units), by k_write
    uint32_t* status_out;        // optional (batch): this stream's status word, zeroed by
                                 // k_init; bit 0 rejected table, bit 1 not converged
};

// Workspace arrays (global indices).
struct Work {
    // per chunk
    uint32_t* exitv;     // spec path entry of boundary (c+1)S
    uint32_t* own_total; // spec path steps in chunk c
    uint32_t* wentry;    // walk start (entry of boundary cS on the proven path)
    uint32_t* wexit;     // walk exit when it did not merge
    uint32_t* wtotal;    // walk steps in chunk when it did not merge
    uint32_t* mq;        // merge checkpoint (stream-local index) or kNone
    uint32_t* wcm;       // walk steps before the merge checkpoint
    uint32_t* fflag;     // 1 = walk into this chunk failed (pending fixup)
    uint64_t* coff;      // exclusive offset of chunk c's first output symbol
    // per checkpoint
    uint32_t* ck;        // spec path entry
    uint32_t* ckc;       // spec path steps since chunk start
    uint32_t* wb;        // walk entry (before merge)
    uint32_t* wbc;       // walk steps since chunk start
    // misc
    uint32_t* status;    // [0] error bits, [3] fixup rounds
    uint32_t* flist;     // fixup lists (global chunk ids), stream s in its chunk range
    uint32_t* flist2;
    uint32_t* ctick;     // task tickets: [0] spec, [1] walk, [2] write
    uint64_t* bsum;      // scan block sums
    uint32_t* fcnt;      // per stream: entries in its fixup list
};

// Storage slot of checkpoint q of a stream.  Chunk c = q / P (P = S/K checkpoints per
// chunk) lies in the 256-chunk range starting at c0 = c & ~255 (one sync workgroup), and
// checkpoint r = q % P of lane l = c - c0 is stored at ck_base + c0*P + r*width + l,
// width = min(256, n_chunks - c0): the lanes of a workgroup store a round's checkpoints
// side by side (full-line writes).  A stream's slots span n_chunks * P entries.
MANS_HD uint64_t ck_slot(const StreamPlan& sp, uint32_t q) {
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t lp = (uint32_t)__builtin_ctz(sp.S) - sp.log2K;
#else
    uint32_t lp = 0;
    while ((sp.K << lp) < sp.S) ++lp;
#endif
    const uint32_t c = q >> lp, r = q & ((1u << lp) - 1u);
    const uint32_t c0 = c & ~255u, l = c & 255u;
    const uint32_t rest = sp.n_chunks - c0;
    const uint32_t width = rest < 256u ? rest : 256u;
    return sp.ck_base + ((uint64_t)c0 << lp) + (uint64_t)r * width + l;
}

// ---------------------------------------------------------------- lane bodies

// Spec pass for chunk c of stream sp (lane body).
template <typename Tab>
MANS_HD void lane_spec(const StreamPlan& sp, const Work& w, const Tab& tab, uint32_t c) {
    const uint32_t L = sp.L;
    const uint32_t u0 = c * sp.S;
    uint32_t u_stop = u0 + sp.S;
    if (u_stop > sp.n_units) u_stop = sp.n_units;
    BCursor cur;
    uint32_t xi, off0;
    if (c == 0u) {
        off0 = 32u - sp.first_bit;
        xi = sp.first_state - L;
    } else if (sp.index_in) {
        // recorded sync point (masked into range: a wrong record only costs walks)
        off0 = sp.index_in[c].bit & 31u;
        xi = sp.index_in[c].state & (L - 1u);
    } else {
        off0 = 0u;
        xi = sp.first_state - L;   // guessed state (the reference guesses the same)
    }
    // warm-up: a chunk c >= 1 starts its guessed decode warm units earlier; the first
    // step starting in unit u0 gives the chunk's first checkpoint
    const uint32_t wu = c == 0u ? 0u : (sp.warm < u0 ? sp.warm : u0);
    bcursor_init(cur, sp.units, sp.n_units, u0 - wu, off0);
    while (cur.u < u0) {
        uint32_t crossed;
        bstep(tab, cur, xi, crossed);
    }
    off0 = cur.o + (cur.u - u0) * 32u;     // = cur.o: a step crosses one unit at most
    const uint32_t gc = sp.chunk_base + c;
    const uint32_t kmask = sp.K - 1u;
    w.ck[ck_slot(sp, u0 >> sp.log2K)] = pack_entry(xi, off0);
    w.ckc[ck_slot(sp, u0 >> sp.log2K)] = 0u;
    if (c == 0u) {
        w.mq[gc] = 0u;
        w.wcm[gc] = 0u;
        w.wentry[gc] = pack_entry(xi, off0);
    }
    uint32_t cnt = 0u;
    while (cur.u < u_stop) {
        uint32_t crossed;
        bstep(tab, cur, xi, crossed);
        ++cnt;
        if (crossed && (cur.u & kmask) == 0u && cur.u < u_stop) {
            uint64_t q = ck_slot(sp, cur.u >> sp.log2K);
            w.ck[q] = pack_entry(xi, cur.o);
            w.ckc[q] = cnt;
        }
    }
    w.exitv[gc] = pack_entry(xi, cur.o);
    w.own_total[gc] = cnt;
}

// Walk into chunk c (c >= 1) from `entry` (packed, offset relative to unit c*S).
// Returns 1 when it merged with the stored path, 0 when it crossed the whole chunk.
template <typename Tab>
MANS_HD uint32_t lane_walk(const StreamPlan& sp, const Work& w, const Tab& tab, uint32_t c,
                           uint32_t entry) {
    const uint32_t u0 = c * sp.S;
    uint32_t u_stop = u0 + sp.S;
    if (u_stop > sp.n_units) u_stop = sp.n_units;
    const uint32_t gc = sp.chunk_base + c;
    const uint32_t kmask = sp.K - 1u;
    w.wentry[gc] = entry;
    uint32_t q0 = u0 >> sp.log2K;
    if (w.ck[ck_slot(sp, q0)] == entry) {
        w.mq[gc] = q0;
        w.wcm[gc] = 0u;
        return 1u;
    }
    w.wb[ck_slot(sp, q0)] = entry;
    w.wbc[ck_slot(sp, q0)] = 0u;
    BCursor cur;
    uint32_t xi = entry_state(entry);
    bcursor_init(cur, sp.units, sp.n_units, u0, entry_off(entry));
    uint32_t cnt = 0u;
    while (cur.u < u_stop) {
        uint32_t crossed;
        bstep(tab, cur, xi, crossed);
        ++cnt;
        if (crossed && (cur.u & kmask) == 0u && cur.u < u_stop) {
            uint32_t q = cur.u >> sp.log2K;
            uint32_t v = pack_entry(xi, cur.o);
            if (w.ck[ck_slot(sp, q)] == v) {
                w.mq[gc] = q;
                w.wcm[gc] = cnt;
                return 1u;
            }
            w.wb[ck_slot(sp, q)] = v;
            w.wbc[ck_slot(sp, q)] = cnt;
        }
    }
    w.mq[gc] = kNone;
    w.wexit[gc] = pack_entry(xi, cur.o);
    w.wtotal[gc] = cnt;
    return 0u;
}

// Walk from `entry` at the start of chunk c (c >= 1) through chunks c .. c_end-1 until it
// merges with the stored spec path, giving every chunk it enters a full record (the
// chunk-crossing form of lane_walk used by the sync kernels).  Returns 1 if merged.
// With stop_after_first != 0 the walk ends at the end of chunk c even if it has not
// merged (the sync kernel's rule for a walk whose predecessor's walk did not merge).
template <typename Tab>
MANS_HD uint32_t lane_walk_chain(const StreamPlan& sp, const Work& w, const Tab& tab, uint32_t c,
                                 uint32_t c_end, uint32_t entry, uint32_t stop_after_first = 0u) {
    if (stop_after_first && c_end > c + 1u) c_end = c + 1u;
    const uint32_t u0 = c * sp.S;
    uint32_t u_stop = c_end * sp.S;
    if (u_stop > sp.n_units) u_stop = sp.n_units;
    uint32_t gc = sp.chunk_base + c;
    const uint32_t kmask = sp.K - 1u, smask = sp.S - 1u;
    w.wentry[gc] = entry;
    uint32_t q0 = u0 >> sp.log2K;
    if (w.ck[ck_slot(sp, q0)] == entry) {
        w.mq[gc] = q0;
        w.wcm[gc] = 0u;
        return 1u;
    }
    w.wb[ck_slot(sp, q0)] = entry;
    w.wbc[ck_slot(sp, q0)] = 0u;
    BCursor cur;
    uint32_t xi = entry_state(entry);
    bcursor_init(cur, sp.units, sp.n_units, u0, entry_off(entry));
    uint32_t cnt = 0u;
    while (cur.u < u_stop) {
        uint32_t crossed;
        bstep(tab, cur, xi, crossed);
        ++cnt;
        if (crossed && (cur.u & kmask) == 0u && cur.u < u_stop) {
            const uint32_t v = pack_entry(xi, cur.o);
            if ((cur.u & smask) == 0u) {          // chunk boundary
                w.mq[gc] = kNone;
                w.wexit[gc] = v;
                w.wtotal[gc] = cnt;
                ++gc;
                w.wentry[gc] = v;
                cnt = 0u;
            }
            const uint32_t q = cur.u >> sp.log2K;
            if (w.ck[ck_slot(sp, q)] == v) {
                w.mq[gc] = q;
                w.wcm[gc] = cnt;
                return 1u;
            }
            w.wb[ck_slot(sp, q)] = v;
            w.wbc[ck_slot(sp, q)] = cnt;
        }
    }
    w.mq[gc] = kNone;
    w.wexit[gc] = pack_entry(xi, cur.o);
    w.wtotal[gc] = cnt;
    return 0u;
}

// 1 if a walk from `entry` at the start of chunk c merges with the stored spec path
// inside chunk c (reads only).
template <typename Tab>
MANS_HD uint32_t lane_merges_in_chunk(const StreamPlan& sp, const Work& w, const Tab& tab,
                                      uint32_t c, uint32_t entry) {
    const uint32_t u0 = c * sp.S;
    uint32_t u_stop = u0 + sp.S;
    if (u_stop > sp.n_units) u_stop = sp.n_units;
    const uint32_t kmask = sp.K - 1u;
    if (w.ck[ck_slot(sp, u0 >> sp.log2K)] == entry) return 1u;
    BCursor cur;
    uint32_t xi = entry_state(entry);
    bcursor_init(cur, sp.units, sp.n_units, u0, entry_off(entry));
    while (cur.u < u_stop) {
        uint32_t crossed;
        bstep(tab, cur, xi, crossed);
        if (crossed && (cur.u & kmask) == 0u && cur.u < u_stop &&
            w.ck[ck_slot(sp, cur.u >> sp.log2K)] == pack_entry(xi, cur.o))
            return 1u;
    }
    return 0u;
}

// Exit of chunk c (entry of boundary (c+1)S) along the chunk's current walk/path.
MANS_HD uint32_t chunk_exit(const StreamPlan& sp, const Work& w, uint32_t c) {
    const uint32_t gc = sp.chunk_base + c;
    return w.mq[gc] == kNone ? w.wexit[gc] : w.exitv[gc];
}

// True number of steps owned by chunk c (valid after walks and fixups).
MANS_HD uint32_t chunk_count(const StreamPlan& sp, const Work& w, uint32_t c) {
    const uint32_t gc = sp.chunk_base + c;
    uint32_t m = w.mq[gc];
    if (m == kNone) return w.wtotal[gc];
    return w.wcm[gc] + (w.own_total[gc] - w.ckc[ck_slot(sp, m)]);
}

// True sync point of chunk c (valid after the fixup rounds and the scan): chunk 0 from
// the stream start, chunk c >= 1 from its recorded walk start (then the exit of chunk
// c-1 on the true path); symbols before it from the scanned chunk offsets.
MANS_HD SyncPoint chunk_sync_point(const StreamPlan& sp, const Work& w, uint32_t c) {
    const uint32_t gc = sp.chunk_base + c;
    SyncPoint p;
    if (c == 0u) {
        p.state = sp.first_state - sp.L;
        p.bit = 32u - sp.first_bit;
    } else {
        const uint32_t e = w.wentry[gc];
        p.state = entry_state(e);
        p.bit = entry_off(e);
    }
    p.before = w.coff[gc] - w.coff[sp.chunk_base];
    return p;
}

// True entry and chunk-relative step count at checkpoint q.
MANS_HD void checkpoint_truth(const StreamPlan& sp, const Work& w, uint32_t q, uint32_t* entry,
                              uint32_t* cum) {
    const uint32_t c = (q << sp.log2K) / sp.S;
    const uint32_t gc = sp.chunk_base + c;
    const uint64_t qq = ck_slot(sp, q);
    uint32_t m = w.mq[gc];
    if (m == kNone || q < m) {
        *entry = w.wb[qq];
        *cum = w.wbc[qq];
    } else {
        *entry = w.ck[qq];
        *cum = w.wcm[gc] + (w.ckc[qq] - w.ckc[ck_slot(sp, m)]);
    }
}

// Output offset of checkpoint q within its stream.  coff holds the exclusive scan of
// chunk counts over all chunks of all streams; the stream's own offsets start at
// coff[chunk_base].
MANS_HD uint64_t checkpoint_offset(const StreamPlan& sp, const Work& w, uint32_t q,
                                   uint32_t* entry) {
    uint32_t cum;
    checkpoint_truth(sp, w, q, entry, &cum);
    const uint32_t c = (q << sp.log2K) / sp.S;
    return w.coff[sp.chunk_base + c] - w.coff[sp.chunk_base] + cum;
}

}  // namespace mans

#endif  // MANS_CORE_H_
