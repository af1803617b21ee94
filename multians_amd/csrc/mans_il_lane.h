// mans_il_lane.h -- one lane of the two-state interleaved decoder (device and host).
//
// The per-lane body of k_il_decode (mans_interleaved.hip), written over a memory policy
// so that the CPU emulation (mans_emu.cc, tests only) runs the same code: `Mem` gives
// the LDS words by byte address (ld / st) and the output stores (out8 / out32 / out128).
#ifndef MANS_IL_LANE_H_
#define MANS_IL_LANE_H_

#include "../../include/mans_c_api.h"
#include "mans_core.h"

namespace mans {

// The ring holds two blocks of kIlBlk units; a lane checks for a block change at most
// kIlCheckBits bits apart, so a window (two units from bit p) never reaches past the
// block after the one that held p at the last check.
constexpr uint32_t kIlBlk = 16;                  // units per ring block
constexpr uint32_t kIlRing = 2u * kIlBlk;        // ring slots
constexpr uint32_t kIlRingStride = kIlRing + 1u; // + guard word (odd: fewer bank conflicts)
constexpr uint32_t kIlCheckBits = 32u * (kIlBlk - 1u);   // 480: 16 pairs of <= 30 bits
static_assert(16u * 30u <= kIlCheckBits, "a check every 16 pairs");

struct IlArgs {
    const uint32_t* units;
    const uint32_t* tab[2];
    const mans_il_record* rec;
    uint8_t* out;
    uint32_t* status;
    uint64_t n;                 // symbols
    uint32_t n_units;
    uint32_t L;
    uint32_t two_tables;        // tab[1] differs from tab[0]
    uint32_t first_state[2];
    uint32_t first_bit;
    uint32_t R;
    uint32_t M;                 // records
};

MANS_HD uint32_t il_bfe(uint32_t v, uint32_t off, uint32_t width) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_ubfe(v, off, width);   // offset and width: low five bits
#else
    off &= 31u;
    width &= 31u;
    return (uint32_t)(((uint64_t)v >> off) & ((1ull << width) - 1ull));
#endif
}

// symbols (byte 1) of two table words: wa's in byte 0, wb's in byte 1
MANS_HD uint32_t il_sym2(uint32_t wb, uint32_t wa) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_perm(wb, wa, 0x0c0c0501u);
#else
    return ((wa >> 8) & 0xFFu) | (((wb >> 8) & 0xFFu) << 8);
#endif
}

// low halves of two pairs: a's in bytes 0-1, b's in bytes 2-3
MANS_HD uint32_t il_join(uint32_t b, uint32_t a) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_perm(b, a, 0x05040100u);
#else
    return (a & 0xFFFFu) | ((b & 0xFFFFu) << 16);
#endif
}

// Lane state.  States are held as LDS dword indices of their table words (table base +
// x - L), with the word itself prefetched: `a` is the slot of the next step.
struct IlLane {
    uint32_t ia, wa, ib, wb;
    uint32_t p;          // bit offset from unit ub
    uint32_t cb;         // ring block holding bit p (p >> 8 at the last check)
    uint32_t ring;       // LDS byte address of the lane's ring
    uint32_t ubs;        // ring slot of unit ub (0 or 8)
    uint32_t ub;         // first unit staged (a multiple of 8)
    uint32_t lo, hi;     // window units at p
    U4 f0, f1, f2, f3;   // the block after the ring's two (in flight)
};

template <class Mem>
MANS_HD void il_window(IlLane& s, const Mem& m) {
    const uint32_t slot = (s.ubs + (s.p >> 5)) & (kIlRing - 1u);
    const uint32_t a = s.ring + 4u * slot;
    s.lo = m.ld(a);
    s.hi = m.ld(a + 4u);
}

template <class Mem>
MANS_HD void il_put4(const Mem& m, uint32_t a, const U4& x) {
    m.st(a, x.x); m.st(a + 4u, x.y); m.st(a + 8u, x.z); m.st(a + 12u, x.w);
}

template <class Mem>
MANS_HD void il_put_block(const IlLane& s, const Mem& m, uint32_t slot0, const U4& x0,
                          const U4& x1, const U4& x2, const U4& x3) {
    const uint32_t a = s.ring + 4u * slot0;
    il_put4(m, a, x0);
    il_put4(m, a + 16u, x1);
    il_put4(m, a + 32u, x2);
    il_put4(m, a + 48u, x3);
    if (slot0 == 0u) m.st(s.ring + 4u * kIlRing, x0.x);      // guard repeats slot 0
}

// Aligned block of four units at b, the address clamped into the stream but the values
// not masked: units past the end are only ever read as the unused top of a window (a
// consistent stream's codes end at bit 32 n_units), and a mask would make the load's
// consumer wait for it at once.
MANS_HD U4 il_load4(const uint32_t* units, uint32_t n, uint32_t b) {
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

MANS_HD void il_fetch(IlLane& s, const uint32_t* units, uint32_t n_units, uint32_t u) {
    s.f0 = il_load4(units, n_units, u);
    s.f1 = il_load4(units, n_units, u + 4u);
    s.f2 = il_load4(units, n_units, u + 8u);
    s.f3 = il_load4(units, n_units, u + 12u);
}

// Block check (at most kIlCheckBits after the previous one): when p entered the next
// block, the block it left takes the one after the next (held in f0-f3, loaded a block
// earlier), and those registers take the following block's load.  The checks sit where
// the lane's last output store is a batch of steps old, so waiting for the block's load
// does not also wait for a store just issued.
template <class Mem>
MANS_HD void il_check(IlLane& s, const Mem& m, const uint32_t* units, uint32_t n_units) {
    if ((s.p >> 9) != s.cb) {
        const uint32_t slot0 = (s.ubs + kIlBlk * s.cb) & (kIlRing - 1u);
        il_put_block(s, m, slot0, s.f0, s.f1, s.f2, s.f3);
        s.cb += 1u;
        il_fetch(s, units, n_units, s.ub + kIlBlk * (s.cb + 2u));
    }
}

// One step of slot a; the slots then swap roles.  Returns the symbol.
template <class Mem>
MANS_HD uint32_t il_step1(IlLane& s, const Mem& m, const uint32_t* units, uint32_t n_units) {
    const uint32_t win = funnel_shr(s.hi, s.lo, s.p);
    const uint32_t i = (s.wa >> 16) + low_bits(win, s.wa);
    const uint32_t sym = (s.wa >> 8) & 0xFFu;
    s.p += s.wa & 31u;
    s.ia = s.ib;
    s.wa = s.wb;
    s.ib = i;
    s.wb = m.ld(i << 2);
    il_window(s, m);
    return sym;
}

// Steps of slot a then slot b from one window; the roles are unchanged after them.
// Returns the two symbols in bytes 0 and 1.  Slot b's code starts nb(a) bits into the
// window (the bit counts sit in the low five bits of both words).
template <class Mem>
MANS_HD uint32_t il_step2(IlLane& s, const Mem& m, const uint32_t* units, uint32_t n_units) {
    const uint32_t win = funnel_shr(s.hi, s.lo, s.p);
    const uint32_t i1 = (s.wa >> 16) + low_bits(win, s.wa);
    const uint32_t i2 = (s.wb >> 16) + il_bfe(win, s.wa, s.wb);
    const uint32_t two = il_sym2(s.wb, s.wa);
    s.p += (s.wa & 31u) + (s.wb & 31u);
    s.ia = i1;
    s.ib = i2;
    s.wa = m.ld(i1 << 2);
    s.wb = m.ld(i2 << 2);
    il_window(s, m);
    return two;
}

// 16 symbols (8 pairs) as four dwords.
template <class Mem>
MANS_HD void il_quad(IlLane& s, Mem& m, const uint32_t* units, uint32_t n_units, uint32_t* v) {
    for (int k = 0; k < 4; ++k) {
        const uint32_t x = il_step2(s, m, units, n_units);
        const uint32_t y = il_step2(s, m, units, n_units);
        v[k] = il_join(y, x);
    }
}

// Decodes cnt steps into dst: bytes up to 4-byte alignment, dwords up to 16-byte
// alignment, 16-byte stores up to 64-byte alignment, then whole 64-byte halves of lines
// held in registers and stored back to back (a lane's half line reaches the L2 at once
// instead of over the time its 64 symbols take, which would leave it partly written while
// thousands of lanes each hold one open line), then 16-byte stores, dwords and bytes.
// Block checks, at most 240 bits apart: after the head (at most 3 single steps and
// 3 dwords, 225 bits), after every 16 symbols (the last one of a line just before its
// stores), and after the tail (225 bits; the next check may be a head later, 450 bits).
template <class Mem>
MANS_HD void il_emit(IlLane& s, Mem& m, const uint32_t* units, uint32_t n_units, uint8_t* dst,
                     uint32_t cnt) {
    uint32_t i = 0;
    const uint32_t head = (uint32_t)((4u - ((uintptr_t)dst & 3u)) & 3u);
    const uint32_t h = head < cnt ? head : cnt;
    for (; i < h; ++i) m.out8(dst + i, il_step1(s, m, units, n_units));
    uint8_t* dw = dst + i;
    uint32_t nd = (cnt - i) >> 2;
    const uint32_t wa = (uint32_t)(((16u - ((uintptr_t)dw & 15u)) & 15u) >> 2);
    const uint32_t ha = wa < nd ? wa : nd;
    for (uint32_t d = 0; d < ha; ++d) {
        const uint32_t x = il_step2(s, m, units, n_units);
        const uint32_t y = il_step2(s, m, units, n_units);
        m.out32(dw + 4u * d, il_join(y, x));
    }
    il_check(s, m, units, n_units);
    dw += 4u * ha;
    nd -= ha;
    uint32_t nq = nd >> 2;                          // 16-byte blocks
    const uint32_t wl = (uint32_t)(((64u - ((uintptr_t)dw & 63u)) & 63u) >> 4);
    const uint32_t hl = wl < nq ? wl : nq;
    for (uint32_t d = 0; d < hl; ++d) {
        uint32_t v[4];
        il_quad(s, m, units, n_units, v);
        il_check(s, m, units, n_units);
        m.out128(dw, v[0], v[1], v[2], v[3]);
        dw += 16;
    }
    nq -= hl;
    for (uint32_t l = nq >> 2; l; --l) {
        uint32_t v[16];
        for (int k = 0; k < 4; ++k) {
            il_quad(s, m, units, n_units, v + 4 * k);
            il_check(s, m, units, n_units);
        }
        for (int k = 0; k < 4; ++k) m.out128(dw + 16 * k, v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
        dw += 64;
    }
    for (uint32_t d = 0; d < (nq & 3u); ++d) {
        uint32_t v[4];
        il_quad(s, m, units, n_units, v);
        il_check(s, m, units, n_units);
        m.out128(dw, v[0], v[1], v[2], v[3]);
        dw += 16;
    }
    for (uint32_t d = 0; d < (nd & 3u); ++d) {
        const uint32_t x = il_step2(s, m, units, n_units);
        const uint32_t y = il_step2(s, m, units, n_units);
        m.out32(dw + 4u * d, il_join(y, x));
    }
    i += (ha + nd) << 2;
    for (; i < cnt; ++i) m.out8(dst + i, il_step1(s, m, units, n_units));
    il_check(s, m, units, n_units);
}

// Interval q's start (q = 0: the stream's start) or end (q = M: the stream's end) as
// (state of slot 0, state of slot 1, absolute bit, steps before).  Returns false for a
// record whose fields are out of range.
MANS_HD bool il_point(const IlArgs& a, uint32_t q, bool end, uint32_t* x0, uint32_t* x1,
                      uint64_t* pos, uint64_t* before) {
    if (!end && q == 0u) {
        *x0 = a.first_state[0] - a.L;
        *x1 = a.first_state[1] - a.L;
        *pos = 32u - a.first_bit;
        *before = 0u;
        return true;
    }
    if (end && q == a.M) {
        *x0 = 0u;
        *x1 = 0u;
        *pos = 32ull * a.n_units;
        *before = a.n;
        return true;
    }
    const uint32_t m = end ? a.M - 1u - q : a.M - q;
    const mans_il_record r = a.rec[m];
    *x0 = r.state[0];
    *x1 = r.state[1];
    *pos = 32ull * ((uint64_t)a.n_units - (uint64_t)a.R * m) + r.bit;
    *before = r.before;
    return r.state[0] < a.L && r.state[1] < a.L && r.bit < 32u * a.R && r.before <= a.n;
}

// Lane q (0 <= q <= M): decodes interval q with its ring at LDS byte address `ring`, the
// staged tables at dwords t0 (slot 0) and t1 (slot 1).  Returns status bits: 2 when the
// interval's records are out of range or its decode does not end on the next record.
template <class Mem>
MANS_HD uint32_t il_lane(const IlArgs& a, uint32_t q, Mem& m, uint32_t t0, uint32_t t1,
                         uint32_t ring) {
    uint32_t x0, x1, e0, e1;
    uint64_t P0, Pe, j0, j1;
    const bool ok = il_point(a, q, false, &x0, &x1, &P0, &j0) &&
                    il_point(a, q, true, &e0, &e1, &Pe, &j1) && j0 <= j1 && P0 <= Pe;
    if (!ok) return 2u;
    IlLane s;
    s.ring = ring;
    s.ub = (uint32_t)(P0 >> 5) & ~(kIlBlk - 1u);
    s.ubs = s.ub & (kIlRing - 1u);
    s.p = (uint32_t)(P0 - 32ull * s.ub);
    s.cb = 0u;
    {
        const U4 b0 = il_load4(a.units, a.n_units, s.ub);
        const U4 b1 = il_load4(a.units, a.n_units, s.ub + 4u);
        const U4 b2 = il_load4(a.units, a.n_units, s.ub + 8u);
        const U4 b3 = il_load4(a.units, a.n_units, s.ub + 12u);
        il_put_block(s, m, s.ubs, b0, b1, b2, b3);
        il_fetch(s, a.units, a.n_units, s.ub + kIlBlk);
        il_put_block(s, m, s.ubs ^ kIlBlk, s.f0, s.f1, s.f2, s.f3);
        il_fetch(s, a.units, a.n_units, s.ub + 2u * kIlBlk);
    }
    il_window(s, m);
    // the slot of step j0 is (n - 1 - j0) % 2 (no step when j0 = n)
    const uint32_t sa = (uint32_t)((a.n - 1u - j0) & 1u);
    s.ia = (sa ? t1 : t0) + (sa ? x1 : x0);
    s.ib = (sa ? t0 : t1) + (sa ? x0 : x1);
    s.wa = m.ld(s.ia << 2);
    s.wb = m.ld(s.ib << 2);
    uint64_t left = j1 - j0;
    uint8_t* dst = a.out + j0;
    while (left) {
        const uint32_t c = left > (1u << 30) ? (1u << 30) : (uint32_t)left;
        il_emit(s, m, a.units, a.n_units, dst, c);
        dst += c;
        left -= c;
    }
    // after the steps, slot a is the slot of the next step
    const uint32_t sn = (uint32_t)((sa + (j1 - j0)) & 1u);
    const uint32_t fa = s.ia - (sn ? t1 : t0), fb = s.ib - (sn ? t0 : t1);
    const uint32_t g0 = sn ? fb : fa, g1 = sn ? fa : fb;
    const uint64_t pend = 32ull * s.ub + s.p;
    return (g0 != e0 || g1 != e1 || pend != Pe) ? 2u : 0u;
}

}  // namespace mans

#endif
