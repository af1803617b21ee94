// mans_lineage.h -- lineage synchronisation of one stream (shared by the kernels and the
// CPU emulation).  No counterpart in the reference: it replaces the sync phases of a
// single-stream decode (src/multians_gpu_decoder.cu:195-466, the speculative decode and
// the block synchronisation) by another way of finding the true decoding path, and feeds
// the count-driven write pass (src/multians_gpu_decoder.cu:492-573) through the trusted
// item-record form of k_write.
//
// Why.  A decode started from a guessed state merges with the true path after a distance
// that is close to exponentially distributed (mean 130-590 units at L = 4096, longest at
// medium and low entropy).  A chunked decode whose chunks each carry one guessed path
// waits, at its end, for the slowest of many such merges: about mean * ln(chunks) units
// of serial decoding.  Here a lineage (a decode from a guessed state) starts every D units
// and runs until it meets the recorded path of any other lineage.  Around any position
// several lineages that have not yet met are alive, so a new one meets one of them much
// sooner than it would meet one particular path: the longest lineage is a few mean merge
// distances, at a cost of 2-5 decodes of the stream in total, spread over many lanes.
//
// Records.  Checkpoint q sits at unit Q*q (Q = 32 units, the write pass's item).  Each
// checkpoint has kLinSlots slots: val = the packed entry (state, bit offset of the first
// step at or after the boundary) of a path that passed it, and inf = where that path is at
// checkpoint q + 1 (a slot there) with the number of steps in between.  A lineage reaching
// checkpoint q searches its slots for its own entry: found, its path from there on is the
// recorded one (decoding is deterministic in (bit, state)), so it links its previous slot
// to the found one and ends; not found, it takes a new slot, links to it and goes on.
// Lineage 0 starts from the stream's true start, recorded in slot 0 of checkpoint 0, so
// the true path is a chain of links from (0, 0).
//
// Resolution.  The checkpoints are grouped into map chunks of kLinChunk.  For every slot
// at a chunk's first checkpoint the map follows the links to the chunk's end (slot there,
// steps).  Paths merge within a few hundred units, so in practice all slots of a chunk
// map to one end slot; then the next chunk's entry slot is known without its own entry.
// Otherwise the entry is found by composing maps back to the nearest such chunk.  The
// steps along the true path per chunk are scanned, and each chunk's trace follows its
// links from its entry slot, writing one mans_sync_point per 32-unit item: the records
// the trusted write pass decodes from.
//
// Anything unexpected (a full slot list, a lineage longer than the cap, a link that was
// never written, maps that do not resolve within kLinLookBack chunks, a rejected table)
// sets the failure word, and the decode runs the chunked sync pipeline instead.
#ifndef MANS_LINEAGE_H_
#define MANS_LINEAGE_H_

#include "mans_core.h"

namespace mans {

constexpr uint32_t kLinQ = 32;            // units per checkpoint (= write item)
constexpr uint32_t kLinSlots = 16;        // slots per checkpoint
constexpr uint32_t kLinChunk = 16;        // checkpoints per map chunk
constexpr uint32_t kLinLookBack = 256;    // map chunks a resolve may compose
constexpr uint32_t kLinEmpty = 0xFFFFFFFFu;   // slot value not yet written
// inf word: next slot (bits 0-4) | steps to the next checkpoint << 5
constexpr uint32_t kLinEnd = 31u;         // next slot: the path reached the last checkpoint
constexpr uint32_t kLinBroken = 30u;      // next slot: the path was not recorded further
constexpr uint32_t kLinMaxSteps = (1u << 27) - 1u;
static_assert(kLinSlots < kLinBroken, "slot numbers below the link markers");

MANS_HD uint32_t lin_link(uint32_t slot, uint32_t steps) { return slot | (steps << 5); }
MANS_HD uint32_t lin_next(uint32_t inf) { return inf & 31u; }
MANS_HD uint32_t lin_steps(uint32_t inf) { return inf >> 5; }

// Control words (status block, words kLinCtl..): [0] lineage ticket, [1] failure (0 = the
// lineage records hold; the chunked pipeline runs only when it is 1), [2] reason bits
// (1 slot list full, 2 lineage cap, 4 unresolved map, 8 link missing, 16 table),
// [3] lineages run, [4] checkpoints decoded by lineages (statistics).
constexpr uint32_t kLinCtl = 64;

struct LinWork {
    uint32_t* val;       // n_q * kLinSlots
    uint32_t* inf;       // n_q * kLinSlots
    uint32_t* cnt;       // n_q: slots taken (may exceed kLinSlots: the rest were refused)
    uint32_t* mslot;     // n_c * kLinSlots: end slot of each start slot's path (or marker)
    uint64_t* msteps;    // n_c * kLinSlots: its steps
    uint32_t* entry;     // n_c: the true path's slot at each chunk's first checkpoint
    uint64_t* count;     // n_c: the true path's steps in each chunk
    SyncPoint* rec;      // n_q: one record per 32-unit item
    uint32_t* ctl;       // control words
    uint32_t n_q;        // checkpoints = items = ceil(n_units / 32)
    uint32_t n_c;        // map chunks = ceil(n_q / kLinChunk)
};

MANS_HD uint32_t lin_slots_used(uint32_t cnt) { return cnt < kLinSlots ? cnt : kLinSlots; }

// Guessed start state (index x - L) of lineage j >= 1: any state will do (a wrong guess
// only makes the lineage longer); a hash spreads the guesses.
MANS_HD uint32_t lin_guess(uint32_t j, uint32_t L) { return ((j * 2654435761u) >> 7) & (L - 1u); }

// Map lane (c, s): the path recorded in slot s of chunk c's first checkpoint, followed to
// the chunk's end (the first checkpoint of chunk c + 1, or the last checkpoint).  Unused
// slots map to kLinEmpty.
MANS_HD void lin_map_lane(const LinWork& lw, uint32_t c, uint32_t s) {
    const uint32_t q0 = c * kLinChunk;
    const uint64_t o = (uint64_t)c * kLinSlots + s;
    if (s >= lin_slots_used(lw.cnt[q0])) {
        lw.mslot[o] = kLinEmpty;
        lw.msteps[o] = 0u;
        return;
    }
    const uint32_t q1 = q0 + kLinChunk < lw.n_q - 1u ? q0 + kLinChunk : lw.n_q - 1u;
    uint32_t slot = s;
    uint64_t steps = 0;
    for (uint32_t q = q0; q < q1; ++q) {
        const uint32_t inf = lw.inf[(uint64_t)q * kLinSlots + slot];
        const uint32_t nx = lin_next(inf);
        if (nx >= kLinSlots) { slot = kLinBroken; break; }     // not recorded further
        steps += lin_steps(inf);
        slot = nx;
    }
    lw.mslot[o] = slot;
    lw.msteps[o] = steps;
}

// The end slot shared by every used start slot of chunk c, or kLinEmpty when they differ
// (or one is broken).
MANS_HD uint32_t lin_common_end(const LinWork& lw, uint32_t c) {
    const uint32_t used = lin_slots_used(lw.cnt[c * kLinChunk]);
    uint32_t e = kLinEmpty;
    for (uint32_t s = 0; s < used; ++s) {
        const uint32_t m = lw.mslot[(uint64_t)c * kLinSlots + s];
        if (m >= kLinSlots) return kLinEmpty;
        if (e == kLinEmpty) e = m;
        else if (m != e) return kLinEmpty;
    }
    return e;
}

// Resolve lane c: the true path's slot at chunk c's first checkpoint and its steps in
// chunk c.  Returns 0, or a failure reason bit.
MANS_HD uint32_t lin_resolve_lane(const LinWork& lw, uint32_t c) {
    uint32_t k = c, slot = 0u;
    // back to the nearest chunk whose start slots all lead to one end slot
    for (;;) {
        if (k == 0u) { slot = 0u; break; }
        const uint32_t e = lin_common_end(lw, k - 1u);
        if (e != kLinEmpty) { slot = e; break; }
        if (c - k >= kLinLookBack) return 4u;
        --k;
    }
    // forward through the maps to chunk c
    for (; k < c; ++k) {
        const uint32_t m = lw.mslot[(uint64_t)k * kLinSlots + slot];
        if (m >= kLinSlots) return 8u;
        slot = m;
    }
    if (slot >= lin_slots_used(lw.cnt[c * kLinChunk])) return 8u;
    if (lw.mslot[(uint64_t)c * kLinSlots + slot] >= kLinSlots) return 8u;
    lw.entry[c] = slot;
    lw.count[c] = lw.msteps[(uint64_t)c * kLinSlots + slot];
    return 0u;
}

// Trace lane c: one record per checkpoint of chunk c, along the true path from its entry
// slot, with output offsets from off (the steps before chunk c).  Returns 0 or 8 (a link
// that is missing).
MANS_HD uint32_t lin_trace_lane(const LinWork& lw, uint32_t c, uint64_t off) {
    const uint32_t q0 = c * kLinChunk;
    const uint32_t q1 = q0 + kLinChunk < lw.n_q ? q0 + kLinChunk : lw.n_q;
    uint32_t slot = lw.entry[c];
    for (uint32_t q = q0; q < q1; ++q) {
        const uint64_t o = (uint64_t)q * kLinSlots + slot;
        const uint32_t v = lw.val[o];
        SyncPoint r;
        r.state = entry_state(v);
        r.bit = entry_off(v);
        r.before = off;
        lw.rec[q] = r;
        if (q + 1u == lw.n_q) break;
        const uint32_t inf = lw.inf[o];
        if (lin_next(inf) >= kLinSlots) return 8u;
        off += lin_steps(inf);
        slot = lin_next(inf);
    }
    return 0u;
}

}  // namespace mans

#endif  // MANS_LINEAGE_H_
