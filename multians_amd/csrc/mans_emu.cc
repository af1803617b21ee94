// mans_emu.cc -- CPU emulation of the decode pipeline (TESTS ONLY, libmans_emu.so).
//
// Runs the per-lane bodies of mans_core.h sequentially, phase by phase, in the same
// order the kernels of mans_kernels.hip run them, so the algorithm (spec, chained
// walks, consistency check, fixup rounds, scan, count-driven write) can be checked
// against the oracle without a GPU.  The product never loads this library.
#include <algorithm>
#include <cstring>
#include <vector>

#include "mans_core.h"
#include "mans_lineage.h"

using namespace mans;

namespace {

struct HostTab {
    const uint32_t* p;
    uint32_t operator[](uint32_t i) const { return p[i]; }
};

uint32_t ilog2(uint32_t v) { uint32_t r = 0; while ((1u << (r + 1)) <= v) ++r; return r; }

}  // namespace

namespace {

int emu_decode(const uint32_t* units, uint64_t n_units, uint8_t* out, uint64_t n_out,
               const uint32_t* table, uint32_t L, uint32_t first_state, uint32_t first_bit,
               uint32_t S, uint32_t K, uint32_t warm, uint64_t* stats, uint32_t* ck_entries_out,
               uint64_t* ck_offsets_out, const SyncPoint* index_in = nullptr,
               SyncPoint* index_out = nullptr, uint32_t round_limit = 0u) {
    if (!n_out) return 0;
    std::vector<uint32_t> tab(L);
    int bad = 0, cyc = 0;
    for (uint32_t i = 0; i < L; ++i) tab[i] = derive_entry(table[i], L, i, &bad, &cyc);
    if (bad || (cyc && (n_units + K - 1) / K > 1)) return -4;
    StreamPlan sp;
    std::memset(&sp, 0, sizeof(sp));
    sp.units = units;
    sp.out = out;
    sp.table = table;
    sp.n_out = n_out;
    sp.n_units = (uint32_t)n_units;
    sp.warm = warm;
    sp.index_in = index_in;
    sp.L = L;
    sp.first_state = first_state;
    sp.first_bit = first_bit;
    sp.S = S;
    sp.K = K;
    sp.log2K = ilog2(K);
    sp.n_chunks = (uint32_t)((n_units + S - 1) / S);
    sp.n_ck = (uint32_t)((n_units + K - 1) / K);
    const uint32_t nc = sp.n_chunks, nq = sp.n_ck;
    std::vector<uint32_t> exitv(nc), own(nc), went(nc), wexit(nc), wtot(nc), mq(nc), wcm(nc),
        ff(nc), fl(nc + 1), fl2(nc + 1), status(64), fcnt(2);
    std::vector<uint64_t> coff(nc);
    const uint64_t ns = (uint64_t)nc * (S / K);     // checkpoint slots (see ck_slot)
    std::vector<uint32_t> ck(ns), ckc(ns), wb(ns), wbc(ns);
    Work w;
    w.exitv = exitv.data(); w.own_total = own.data(); w.wentry = went.data();
    w.wexit = wexit.data(); w.wtotal = wtot.data(); w.mq = mq.data(); w.wcm = wcm.data();
    w.fflag = ff.data(); w.coff = coff.data(); w.ck = ck.data(); w.ckc = ckc.data();
    w.wb = wb.data(); w.wbc = wbc.data(); w.status = status.data(); w.flist = fl.data();
    w.flist2 = fl2.data(); w.ctick = nullptr; w.fcnt = fcnt.data(); w.bsum = nullptr;
    HostTab ht{tab.data()};
    uint64_t st_fail = 0, st_rounds = 0, st_rewalks = 0, st_merges = 0;
    // k_spec (a table with zero-bit cycles has a single checkpoint interval here:
    // only the chunk-0 start is recorded)
    if (cyc) {
        ck[0] = pack_entry(first_state - L, 32u - first_bit);
        ckc[0] = 0; mq[0] = 0; wcm[0] = 0;
    } else {
        for (uint32_t c = 0; c < nc; ++c) lane_spec(sp, w, ht, c);
    }
    // Walks (k_spec's for a stream of one range, k_walk's otherwise): in a range of R =
    // 256 chunks, the walk into chunk c0 + t starts from the spec exit of its predecessor
    // and goes on to the range end until it merges (chunk 0 is not walked).  Lanes run in
    // lockstep rounds on the GPU, so where walks overlap the one that started further
    // back writes last; running the lanes from the back of the range to its front
    // reproduces that.  A walk leaves its first chunk only if the previous
    // lane's walk merged inside its own chunk (the range's first walk always may, and
    // chunk 0 counts as merged).  Ranges are independent: a walk never leaves its range.
    const uint32_t R = 256u;
    std::vector<uint8_t> own_merged(nc, 0);
    own_merged[0] = 1;
    for (uint32_t t = 1; t < nc; ++t) own_merged[t] = (uint8_t)lane_merges_in_chunk(sp, w, ht, t, exitv[t - 1]);
    for (uint32_t c0 = 0; c0 < nc; c0 += R) {
        const uint32_t c_end = c0 + R < nc ? c0 + R : nc;
        for (uint32_t t = c_end; t-- > (c0 > 1u ? c0 : 1u);) {
            const uint32_t stop = (t > c0 && !own_merged[t - 1]) ? 1u : 0u;
            st_merges += lane_walk_chain(sp, w, ht, t, c_end, exitv[t - 1], stop);
        }
        // repair passes (as many as the kernel's): chunks after c0 whose start is not
        // their predecessor's exit, walked again from it (entries captured first; no stop
        // rule; back to front as above)
        for (uint32_t pass = 0; pass < 4u; ++pass) {
            std::vector<uint32_t> again, from;
            for (uint32_t t = c0 + 1u; t < c_end; ++t) {
                const uint32_t e = chunk_exit(sp, w, t - 1u);
                if (went[t] != e) { again.push_back(t); from.push_back(e); }
            }
            if (again.empty()) break;
            for (size_t k = again.size(); k-- > 0;)
                st_merges += lane_walk_chain(sp, w, ht, again[k], c_end, from[k]);
        }
    }
    // k_check: every chunk whose walk start differs from its predecessor's exit
    for (uint32_t c = 0; c < nc; ++c) {
        ff[c] = 0;
        if (c >= 1 && went[c] != chunk_exit(sp, w, c - 1)) fl[status[1]++] = c;
    }
    st_fail = status[1];
    // k_fixup (same phases and the same round bound as the kernel: n_chunks + 1 rounds,
    // or round_limit; a stream still inconsistent then is reported, -5)
    uint32_t* list = fl.data();
    uint32_t* next = fl2.data();
    const uint32_t cap = round_limit && round_limit <= nc ? round_limit : nc + 1u;
    for (uint32_t round = 0;; ++round) {
        uint32_t n = status[1];
        if (!n) break;
        if (round >= cap) return -5;
        ++st_rounds;
        for (uint32_t i = 0; i < n; ++i) {          // A: capture entries, mark
            uint32_t t = list[i];
            next[i] = chunk_exit(sp, w, t - 1);
            ff[t] = 1;
        }
        for (uint32_t i = 0; i < n; ++i) {          // B: walk again
            ++st_rewalks;
            lane_walk(sp, w, ht, list[i], next[i]);
        }
        status[2] = 0;
        for (uint32_t i = 0; i < n; ++i) {          // C: consistency of t and t+1
            uint32_t t = list[i];
            if (w.wentry[t] != chunk_exit(sp, w, t - 1)) next[status[2]++] = t;
            if (t + 1 < nc && !ff[t + 1] && w.wentry[t + 1] != chunk_exit(sp, w, t))
                next[status[2]++] = t + 1;
        }
        for (uint32_t i = 0; i < n; ++i) ff[list[i]] = 0;
        status[1] = status[2];
        std::swap(list, next);
    }
    // k_scan
    uint64_t run = 0;
    for (uint32_t c = 0; c < nc; ++c) {
        coff[c] = run;
        run += chunk_count(sp, w, c);
    }
    // k_index
    if (index_out)
        for (uint32_t c = 0; c < nc; ++c) index_out[c] = chunk_sync_point(sp, w, c);
    // k_write
    for (uint32_t q = 0; q < nq; ++q) {
        uint32_t entry, e2;
        uint64_t off = checkpoint_offset(sp, w, q, &entry);
        uint64_t end = (q + 1 < nq) ? checkpoint_offset(sp, w, q + 1, &e2) : n_out;
        if (ck_entries_out) ck_entries_out[q] = entry;
        if (ck_offsets_out) ck_offsets_out[q] = off;
        if (end > n_out) end = n_out;
        if (off >= end) continue;
        Cursor cur;
        cursor_init(cur, units, (uint32_t)n_units, q << sp.log2K, entry_off(entry));
        uint32_t xi = entry_state(entry);
        for (uint64_t i = off; i < end; ++i) {
            uint32_t crossed;
            uint32_t wd = step(ht, cur, xi, crossed);
            out[i] = (uint8_t)(wd >> 8);
        }
    }
    if (stats) {
        stats[0] = nc; stats[1] = nq; stats[2] = st_fail; stats[3] = st_rounds;
        stats[4] = st_rewalks; stats[5] = st_merges; stats[6] = 0; stats[7] = 0;
    }
    return 0;
}

}  // namespace

extern "C" {

// Returns 0, -4 for a rejected table.  stats (optional, 8 u64): n_chunks, n_ck,
// chunks listed by the consistency check, fixup rounds, total re-walks, merges, 0, 0.
int mans_emu_decode(const uint32_t* units, uint64_t n_units, uint8_t* out, uint64_t n_out,
                    const uint32_t* table, uint32_t L, uint32_t first_state, uint32_t first_bit,
                    uint32_t S, uint32_t K, uint64_t* stats, uint32_t* ck_entries_out,
                    uint64_t* ck_offsets_out) {
    return emu_decode(units, n_units, out, n_out, table, L, first_state, first_bit, S, K, 0u,
                      stats, ck_entries_out, ck_offsets_out);
}

// The same with a spec warm-up of `warm` units (a multiple of K) before each chunk.
int mans_emu_decode_warm(const uint32_t* units, uint64_t n_units, uint8_t* out, uint64_t n_out,
                         const uint32_t* table, uint32_t L, uint32_t first_state,
                         uint32_t first_bit, uint32_t S, uint32_t K, uint32_t warm,
                         uint64_t* stats) {
    return emu_decode(units, n_units, out, n_out, table, L, first_state, first_bit, S, K, warm,
                      stats, nullptr, nullptr);
}

// The same with at most round_limit fixup rounds (0 = n_chunks + 1): -5 when the stream
// is still inconsistent after them (the kernel's status bit 1).
int mans_emu_decode_limited(const uint32_t* units, uint64_t n_units, uint8_t* out,
                            uint64_t n_out, const uint32_t* table, uint32_t L,
                            uint32_t first_state, uint32_t first_bit, uint32_t S, uint32_t K,
                            uint32_t round_limit, uint64_t* stats) {
    return emu_decode(units, n_units, out, n_out, table, L, first_state, first_bit, S, K, 0u,
                      stats, nullptr, nullptr, nullptr, nullptr, round_limit);
}

// Zero-bit run folding of the sync kernels' tables (fold_zero_bit_runs in
// mans_kernels.hip), checked on the CPU: the in-place pointer jumping the kernel runs
// (here sequentially, in thread-stride order) must give the words that following each
// zero-bit run one step at a time gives, and decoding with the folded words (one lookup =
// 1 + z steps at one bit position) must meet every unit boundary in the same state, at
// the same offset, after the same number of steps as decoding one step at a time.
// Returns 0, or a negative code naming the first difference; stats (4 u64): zero-bit
// states, longest run, lookups of the folded decode, steps.
int mans_emu_fold_check(const uint32_t* units, uint64_t n_units, const uint32_t* table,
                        uint32_t L, uint32_t first_state, uint32_t first_bit, uint64_t n_out,
                        uint64_t* stats) {
    const uint32_t kMax = 255u;
    std::vector<uint32_t> w(L), P(L), ref(L), fold(L);
    int bad = 0, cyc = 0;
    for (uint32_t i = 0; i < L; ++i) w[i] = derive_entry(table[i], L, i, &bad, &cyc) & ~0xFF00u;
    if (bad || cyc) return -4;
    // pointer jumping (next | z << 16), thread-stride order, in place, until no change
    uint64_t zb = 0;
    for (uint32_t i = 0; i < L; ++i) {
        const bool z = (w[i] & 31u) == 0u;
        P[i] = z ? ((w[i] >> 16) | (1u << 16)) : i;
        zb += z;
    }
    for (bool changed = true; changed;) {
        changed = false;
        for (uint32_t t = 0; t < 256u; ++t)
            for (uint32_t i = t; i < L; i += 256u) {
                const uint32_t p = P[i], z = p >> 16;
                if (!z) continue;
                const uint32_t q = P[p & 0xFFFFu], zq = q >> 16;
                if (!zq || z + zq > kMax) continue;
                P[i] = (q & 0xFFFFu) | ((z + zq) << 16);
                changed = true;
            }
    }
    uint64_t longest = 0;
    for (uint32_t i = 0; i < L; ++i) {
        const uint32_t z = P[i] >> 16;
        fold[i] = z ? (w[P[i] & 0xFFFFu] | (z << 8)) : w[i];     // symbol bytes are 0
        // reference: follow the run one step at a time (at most kMax steps)
        uint32_t j = i, n = 0;
        while ((w[j] & 31u) == 0u && n < kMax) { j = w[j] >> 16; ++n; }
        if (n == kMax && (w[j] & 31u) == 0u) {
            // capped: pointer jumping stops at some run length <= kMax, which is valid
            // too; check that its word describes a real position on this run
            uint32_t k = i, m = 0;
            while (m < z && (w[k] & 31u) == 0u) { k = w[k] >> 16; ++m; }
            if (m != z || (fold[i] & ~0xFF00u) != w[k]) return -10;
        } else {
            ref[i] = n ? (w[j] | (n << 8)) : w[i];
            if (fold[i] != ref[i]) return -11;
        }
        longest = std::max<uint64_t>(longest, z);
    }
    // decode both ways and compare at every unit boundary
    auto bit = [&](uint64_t q) -> uint32_t {
        const uint64_t u = q >> 5;
        return u < n_units ? (units[u] >> (q & 31u)) & 1u : 0u;
    };
    auto bits = [&](uint64_t q, uint32_t k) {
        uint32_t v = 0;
        for (uint32_t i = 0; i < k; ++i) v |= bit(q + i) << i;
        return v;
    };
    std::vector<uint64_t> e1, c1, e2, c2;
    {   // plain
        uint64_t p = 32u - first_bit, cnt = 0, next_u = 0;
        uint32_t xi = first_state - L;
        while (cnt < n_out) {
            while (next_u <= n_units && 32ull * next_u <= p) {
                e1.push_back(xi | ((p - 32ull * next_u) << 16));
                c1.push_back(cnt);
                ++next_u;
            }
            const uint32_t d = w[xi];
            xi = (d >> 16) + bits(p, d & 31u);
            p += d & 31u;
            ++cnt;
        }
    }
    uint64_t lookups = 0;
    {   // folded
        uint64_t p = 32u - first_bit, cnt = 0, next_u = 0;
        uint32_t xi = first_state - L;
        while (cnt < n_out) {
            while (next_u <= n_units && 32ull * next_u <= p) {
                e2.push_back(xi | ((p - 32ull * next_u) << 16));
                c2.push_back(cnt);
                ++next_u;
            }
            const uint32_t d = fold[xi];
            xi = (d >> 16) + bits(p, d & 31u);
            p += d & 31u;
            cnt += 1u + ((d >> 8) & 0xFFu);
            ++lookups;
        }
    }
    if (e1 != e2) return -12;
    if (c1 != c2) return -13;
    if (stats) { stats[0] = zb; stats[1] = longest; stats[2] = lookups; stats[3] = n_out; }
    return 0;
}

// The same from recorded sync points (index_in, n_chunks records, or NULL) and/or
// writing the true ones (index_out, or NULL); no warm-up.
int mans_emu_decode_index(const uint32_t* units, uint64_t n_units, uint8_t* out, uint64_t n_out,
                          const uint32_t* table, uint32_t L, uint32_t first_state,
                          uint32_t first_bit, uint32_t S, uint32_t K, const void* index_in,
                          void* index_out, uint64_t* stats) {
    return emu_decode(units, n_units, out, n_out, table, L, first_state, first_bit, S, K, 0u,
                      stats, nullptr, nullptr, (const SyncPoint*)index_in, (SyncPoint*)index_out);
}

}  // extern "C"

// ------------------------------------------------ lineage synchronisation
// The lineage path of a single-stream decode (mans_lineage.h), run on the CPU: lineages in
// the kernel's ticket order (the last start first), each decoded one step at a time,
// searching and taking slots as k_lin does; then the map, resolve, scan and trace lanes
// (shared code) and the trusted write pass.  Returns 0 with the output written, 1 when the
// lineage records do not hold (the decode would run the chunked pipeline; reason bits in
// stats[2]), -4 for a rejected table.  D: units between lineage starts (a multiple of 32);
// slot_limit (<= kLinSlots, 0 = all) and cap_ck (0 = none: checkpoints one lineage may
// pass) force the failure paths.  stats (8 u64): lineages, checkpoints decoded by
// lineages, reason bits, map chunks, most slots taken at a checkpoint, items, 0, 0.
static int emu_lineage(const uint32_t* units, uint64_t n_units, uint8_t* out, uint64_t n_out,
                const uint32_t* table, uint32_t L, uint32_t first_state, uint32_t first_bit,
                uint32_t D, uint32_t slot_limit, uint32_t cap_ck, uint64_t* stats,
                SyncPoint* rec_out) {
    if (!n_out) return 0;
    if (D < kLinQ || D % kLinQ) return -2;
    std::vector<uint32_t> tab(L);
    int bad = 0, cyc = 0;
    for (uint32_t i = 0; i < L; ++i) tab[i] = derive_entry(table[i], L, i, &bad, &cyc);
    if (bad) return -4;
    const uint32_t nu = (uint32_t)n_units;
    const uint32_t n_q = (nu + kLinQ - 1u) / kLinQ;
    const uint32_t n_c = (n_q + kLinChunk - 1u) / kLinChunk;
    const uint32_t SL = slot_limit && slot_limit < kLinSlots ? slot_limit : kLinSlots;
    std::vector<uint32_t> val((size_t)n_q * kLinSlots, kLinEmpty), inf((size_t)n_q * kLinSlots, 0u),
        cnt(n_q, 0u), mslot((size_t)n_c * kLinSlots), entry(n_c), ctl(8, 0u);
    std::vector<uint64_t> msteps((size_t)n_c * kLinSlots), count(n_c), coff(n_c);
    std::vector<SyncPoint> rec(n_q);
    LinWork lw;
    lw.val = val.data(); lw.inf = inf.data(); lw.cnt = cnt.data(); lw.mslot = mslot.data();
    lw.msteps = msteps.data(); lw.entry = entry.data(); lw.count = count.data();
    lw.rec = rec.data(); lw.ctl = ctl.data(); lw.n_q = n_q; lw.n_c = n_c;
    HostTab ht{tab.data()};
    uint32_t reason = cyc && n_q > 1u ? 16u : 0u;
    uint64_t st_lin = 0, st_ck = 0, most = 1;
    // the true start: slot 0 of checkpoint 0
    val[0] = pack_entry(first_state - L, 32u - first_bit);
    cnt[0] = 1u;
    const uint32_t n_lin = (nu + D - 1u) / D;
    for (uint32_t t = 0; t < n_lin && !reason; ++t) {
        const uint32_t j = n_lin - 1u - t;            // ticket order: the last start first
        const uint32_t u0 = j * D;
        uint32_t q = u0 / kLinQ, slot = kLinEmpty, xi;
        Cursor cur;
        if (j == 0u) {
            slot = 0u;
            xi = first_state - L;
            cursor_init(cur, units, nu, 0u, 32u - first_bit);
        } else {
            if (q + 1u >= n_q) continue;              // no checkpoint after its start
            xi = lin_guess(j, L);
            cursor_init(cur, units, nu, u0, 0u);
        }
        ++st_lin;
        uint32_t passed = 0;
        while (q + 1u < n_q) {
            // decode to checkpoint q + 1: every step that starts before its unit
            const uint32_t b = (q + 1u) * kLinQ;
            uint32_t steps = 0;
            while (cur.u < b) {
                uint32_t crossed;
                step(ht, cur, xi, crossed);
                ++steps;
            }
            ++q;
            ++st_ck;
            const uint32_t v = pack_entry(xi, cur.o);
            const uint64_t base = (uint64_t)q * kLinSlots;
            uint32_t found = kLinEmpty;
            for (uint32_t s = 0; s < lin_slots_used(cnt[q]) && s < SL; ++s)
                if (val[base + s] == v) { found = s; break; }
            if (found != kLinEmpty) {
                if (slot != kLinEmpty) inf[(uint64_t)(q - 1u) * kLinSlots + slot] = lin_link(found, steps);
                break;                                  // merged
            }
            const uint32_t ns = cnt[q]++;
            if (ns >= SL) {
                if (slot != kLinEmpty)
                    inf[(uint64_t)(q - 1u) * kLinSlots + slot] = lin_link(kLinBroken, 0u);
                reason |= 1u;
                break;
            }
            val[base + ns] = v;
            most = std::max<uint64_t>(most, ns + 1u);
            if (slot != kLinEmpty) inf[(uint64_t)(q - 1u) * kLinSlots + slot] = lin_link(ns, steps);
            slot = ns;
            if (q + 1u == n_q) {                        // the last checkpoint
                inf[base + ns] = lin_link(kLinEnd, 0u);
                break;
            }
            if (cap_ck && ++passed >= cap_ck) {
                inf[base + ns] = lin_link(kLinBroken, 0u);
                reason |= 2u;
                break;
            }
        }
        // lineage 0 in a stream of one checkpoint: nothing to link
        if (j == 0u && n_q == 1u) inf[0] = lin_link(kLinEnd, 0u);
    }
    if (!reason) {
        for (uint32_t c = 0; c < n_c; ++c)
            for (uint32_t s = 0; s < kLinSlots; ++s) lin_map_lane(lw, c, s);
        for (uint32_t c = 0; c < n_c && !reason; ++c) reason |= lin_resolve_lane(lw, c);
    }
    if (!reason) {
        uint64_t run = 0;                               // k_scan over the chunk counts
        for (uint32_t c = 0; c < n_c; ++c) { coff[c] = run; run += count[c]; }
        for (uint32_t c = 0; c < n_c && !reason; ++c) reason |= lin_trace_lane(lw, c, coff[c]);
    }
    if (stats) {
        stats[0] = st_lin; stats[1] = st_ck; stats[2] = reason; stats[3] = n_c;
        stats[4] = most; stats[5] = n_q; stats[6] = 0; stats[7] = 0;
    }
    if (reason) return 1;
    if (rec_out) std::copy(rec.begin(), rec.end(), rec_out);
    // the trusted write pass (k_write with item records)
    for (uint32_t q = 0; q < n_q; ++q) {
        uint64_t off = rec[q].before, end = q + 1u < n_q ? rec[q + 1u].before : n_out;
        if (end > n_out) end = n_out;
        if (off >= end) continue;
        Cursor cur;
        cursor_init(cur, units, nu, q * kLinQ, q == 0u ? 32u - first_bit : rec[q].bit);
        uint32_t xi = q == 0u ? first_state - L : rec[q].state;
        for (uint64_t i = off; i < end; ++i) {
            uint32_t crossed;
            out[i] = (uint8_t)(step(ht, cur, xi, crossed) >> 8);
        }
    }
    return 0;
}

extern "C" int mans_emu_lineage_decode(const uint32_t* units, uint64_t n_units, uint8_t* out,
                                       uint64_t n_out, const uint32_t* table, uint32_t L,
                                       uint32_t first_state, uint32_t first_bit, uint32_t D,
                                       uint32_t slot_limit, uint32_t cap_ck, uint64_t* stats,
                                       void* rec_out) {
    return emu_lineage(units, n_units, out, n_out, table, L, first_state, first_bit, D,
                       slot_limit, cap_ck, stats, (SyncPoint*)rec_out);
}

// ------------------------------------------------ two-state interleaved streams
// The lanes of k_il_decode run one after another through the same per-lane code
// (mans_il_lane.h), with a host array standing in for LDS: tables at dword kT0 (slot 0)
// and after it (slot 1), the ring after both.
#include "mans_il_lane.h"

namespace {

struct HostMem {
    uint32_t* lds;
    uint32_t ld(uint32_t a) const { return lds[a >> 2]; }
    void st(uint32_t a, uint32_t v) const { lds[a >> 2] = v; }
    void out8(uint8_t* p, uint32_t v) const { *p = (uint8_t)v; }
    void out32(uint8_t* p, uint32_t v) const { std::memcpy(p, &v, 4); }
    void out128(uint8_t* p, uint32_t a, uint32_t b, uint32_t c, uint32_t d) const {
        const uint32_t v[4] = {a, b, c, d};
        std::memcpy(p, v, 16);
    }
};

}  // namespace

extern "C" int mans_emu_il_decode(const uint32_t* units, uint64_t n_units, const uint32_t* tab0,
                                  const uint32_t* tab1, uint32_t L, const uint32_t* first_states,
                                  uint32_t first_bit, uint64_t n, uint32_t R,
                                  const mans_il_record* rec, uint8_t* out, uint32_t* lane_status) {
    constexpr uint32_t kT0 = 36;      // a nonzero table base, as in LDS after static words
    const uint32_t tw = (L + 3u) & ~3u;
    std::vector<uint32_t> lds(kT0 + 2u * tw + kIlRingStride + 4u, 0xDEADBEEFu);
    const bool two = tab0 != tab1;
    const uint32_t t1 = two ? kT0 + tw : kT0;
    int bad = 0, cyc = 0;
    for (uint32_t k = 0; k < (two ? 2u : 1u); ++k) {
        const uint32_t tb = k ? t1 : kT0;
        const uint32_t* t = k ? tab1 : tab0;
        for (uint32_t i = 0; i < L; ++i) lds[tb + i] = derive_entry(t[i], L, i, &bad, &cyc) + (tb << 16);
    }
    if (bad) return -4;
    IlArgs a;
    std::memset(&a, 0, sizeof(a));
    a.units = units;
    a.tab[0] = tab0;
    a.tab[1] = tab1;
    a.two_tables = two;
    a.rec = rec;
    a.out = out;
    a.n = n;
    a.n_units = (uint32_t)n_units;
    a.L = L;
    a.first_state[0] = first_states[0];
    a.first_state[1] = first_states[1];
    a.first_bit = first_bit;
    a.R = R;
    a.M = (uint32_t)((n_units + R - 1) / R);
    HostMem m{lds.data()};
    uint32_t all = 0;
    for (uint32_t q = 0; q <= a.M; ++q) {
        const uint32_t f = il_lane(a, q, m, kT0, t1, 4u * (kT0 + (two ? 2u : 1u) * tw));
        if (lane_status) lane_status[q] = f;
        all |= f;
    }
    return (int)all;
}
