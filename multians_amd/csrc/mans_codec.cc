// mans_codec.cc -- host-side producers of the decoder's inputs (C ABI).
//
// These are the reference's codec-construction layer, rebuilt for throughput
// (thousands of tables and gigabytes of symbols for the batched workloads):
//   mans_build_tables   ANSTableGenerator::generate_table / generate_encoder_table /
//                       get_decoder_table (src/ans_table_generator.cc:164-288)
//   mans_encode         ANSEncoder::encode (src/ans_encoder.cc:13-97) + the unit
//                       reversal of CUHDInputBuffer (src/multians_input_buffer.cc:29)
//   mans_histogram      symbol counts of a buffer (the counting step of
//                       ANSTableGenerator::generate_distribution_from_buffer)
//   mans_normalize_counts  counts from a histogram (absent symbols get 0, present >= 1)
//   mans_gen_symbols    s = min(255, floor(X)), X ~ Exp(lambda), counter-based
#include <algorithm>
#include <cmath>
#include <cstring>
#include <queue>
#include <thread>
#include <vector>

#include "mans_internal.h"

namespace {

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// s(u) = #{k in 1..255 : u < thr[k]} for u uniform on [0, 2^53); s is non-increasing
// in u.  Bucketing u by its top 12 bits gives, per bucket, the range [s(end), s(start)];
// most buckets hold one symbol, the rest are finished by a short binary search.
struct SymTable {
    uint64_t thr[256];
    uint8_t hi[4096], lo[4096];
};

inline uint32_t sym_of(const uint64_t* thr, uint64_t u, uint32_t lo, uint32_t hi) {
    // largest k in [lo, hi] with u < thr[k] (u < thr[lo] holds)
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (u < thr[mid]) lo = mid; else hi = mid - 1;
    }
    return lo;
}

void gen_range(uint64_t seed, const SymTable* t, uint8_t* out, uint64_t b, uint64_t e) {
    for (uint64_t i = b; i < e; ++i) {
        const uint64_t u = mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ull) >> 11;
        const uint32_t k = (uint32_t)(u >> 41);
        const uint32_t h = t->hi[k], l = t->lo[k];
        out[i] = (uint8_t)(h == l ? h : sym_of(t->thr, u, l, h));
    }
}

struct Cand {
    double p;
    uint32_t s;
    bool operator>(const Cand& o) const { return p > o.p || (p == o.p && s > o.s); }
};

}  // namespace

extern "C" {

int mans_gen_symbols(uint64_t seed, double lambda, uint8_t* out, uint64_t n, int threads) {
    if (!out && n) return set_error(MANS_EINVAL, "out is NULL");
    if (!(lambda > 0.0)) return set_error(MANS_EINVAL, "lambda must be > 0");
    SymTable t;
    for (int k = 0; k < 256; ++k) t.thr[k] = (uint64_t)std::floor(std::ldexp(std::exp(-lambda * k), 53));
    for (uint32_t k = 0; k < 4096; ++k) {
        t.hi[k] = (uint8_t)sym_of(t.thr, (uint64_t)k << 41, 0, 255);
        t.lo[k] = (uint8_t)sym_of(t.thr, (((uint64_t)k + 1) << 41) - 1, 0, 255);
    }
    if (threads <= 1 || n < (1u << 20)) {
        gen_range(seed, &t, out, 0, n);
        return MANS_OK;
    }
    std::vector<std::thread> th;
    uint64_t per = (n + threads - 1) / threads;
    for (int i = 0; i < threads; ++i) {
        uint64_t b = std::min<uint64_t>(n, per * i), e = std::min<uint64_t>(n, b + per);
        th.emplace_back(gen_range, seed, &t, out, b, e);
    }
    for (auto& x : th) x.join();
    return MANS_OK;
}

int mans_histogram(const uint8_t* in, uint64_t n, uint64_t* hist, int threads) {
    if ((!in && n) || !hist) return set_error(MANS_EINVAL, "NULL argument");
    if (threads < 1) threads = 1;
    if (n < (1u << 22)) threads = 1;
    std::vector<std::vector<uint64_t>> part(threads, std::vector<uint64_t>(256, 0));
    auto run = [&](int t) {
        const uint64_t per = (n + threads - 1) / threads;
        const uint64_t b = std::min<uint64_t>(n, per * t), e = std::min<uint64_t>(n, b + per);
        uint64_t* h = part[t].data();
        for (uint64_t i = b; i < e; ++i) ++h[in[i]];
    };
    std::vector<std::thread> th;
    for (int t = 1; t < threads; ++t) th.emplace_back(run, t);
    run(0);
    for (auto& x : th) x.join();
    for (int s = 0; s < 256; ++s) {
        uint64_t v = 0;
        for (int t = 0; t < threads; ++t) v += part[t][s];
        hist[s] = v;
    }
    return MANS_OK;
}

int mans_normalize_counts(const uint64_t* hist, uint32_t nsym, uint32_t L, uint32_t* counts) {
    if (!hist || !counts || !nsym) return set_error(MANS_EINVAL, "bad histogram");
    uint64_t total = 0;
    uint32_t present = 0;
    for (uint32_t s = 0; s < nsym; ++s) { total += hist[s]; present += hist[s] != 0; }
    if (!total || present > L) return set_error(MANS_EINVAL, "histogram cannot be normalised to L=%u", L);
    int64_t sum = 0;
    for (uint32_t s = 0; s < nsym; ++s) {
        uint64_t c = hist[s] ? (uint64_t)(((unsigned __int128)hist[s] * L) / total) : 0;
        counts[s] = hist[s] ? (c ? (uint32_t)c : 1u) : 0u;
        sum += counts[s];
    }
    while (sum != (int64_t)L) {
        uint32_t best = nsym;
        for (uint32_t s = 0; s < nsym; ++s) {
            if (!counts[s] || (sum > (int64_t)L && counts[s] <= 1)) continue;
            if (best == nsym || counts[s] > counts[best]) best = s;
        }
        if (best == nsym) return set_error(MANS_EINVAL, "normalisation failed");
        if (sum < (int64_t)L) { counts[best] += (uint32_t)((int64_t)L - sum); sum = L; }
        else { --counts[best]; --sum; }
    }
    return MANS_OK;
}

int mans_build_tables(const uint32_t* counts, uint32_t nsym, uint32_t L, uint32_t* enc,
                      uint32_t* dec) {
    if (!counts || !nsym || nsym > 256 || L < 2 || L > 32768)
        return set_error(MANS_EINVAL, "bad table parameters");
    uint64_t sum = 0;
    for (uint32_t s = 0; s < nsym; ++s) sum += counts[s];
    if (sum != L) return set_error(MANS_EINVAL, "counts sum to %llu, not L=%u", (unsigned long long)sum, L);
    // Spread: state L+i goes to the symbol with the smallest key p_s (lowest symbol
    // on ties); p_s starts at 0.5/P_s and grows by 1/P_s, P_s = count/L in double.
    std::priority_queue<Cand, std::vector<Cand>, std::greater<Cand>> pq;
    std::vector<double> P(nsym);
    std::vector<uint32_t> X(nsym);
    for (uint32_t s = 0; s < nsym; ++s) {
        X[s] = counts[s];
        if (!counts[s]) continue;
        P[s] = (double)counts[s] / (double)L;
        pq.push({0.5 / P[s], s});
    }
    std::vector<uint32_t> st_sym(L), st_x(L);
    for (uint32_t i = 0; i < L; ++i) {
        Cand c = pq.top();
        pq.pop();
        st_sym[i] = c.s;
        st_x[i] = X[c.s]++;
        pq.push({c.p + 1.0 / P[c.s], c.s});
    }
    if (enc) std::memset(enc, 0, sizeof(uint32_t) * 3ull * nsym * L);
    for (uint32_t i = 0; i < L; ++i) {
        const uint32_t s = st_sym[i], xv = st_x[i];
        uint32_t first_len = 0;
        bool have_len = false;
        for (uint32_t k = 0; ((uint64_t)xv << k) < 2ull * L; ++k) {
            uint64_t lo = (uint64_t)xv << k, hi = lo + (1ull << k);
            lo = std::max<uint64_t>(lo, L);
            hi = std::min<uint64_t>(hi, 2ull * L);
            if (lo >= hi) continue;
            if (!have_len && k > 0) { first_len = k; have_len = true; }
            if (enc) {
                for (uint64_t x = lo; x < hi; ++x) {
                    uint32_t* e = enc + 3ull * ((uint64_t)s * L + (x - L));
                    e[0] = L + i;
                    e[1] = (uint32_t)(x & ((1ull << k) - 1));
                    e[2] = k;
                }
            }
        }
        // reference decoder entry: next = X, nbits = first nonzero length met in
        // ascending source-state order (src/ans_table_generator.cc:274-281)
        if (dec) dec[i] = (xv & 0xFFFFu) | (s << 16) | (first_len << 24);
    }
    return MANS_OK;
}

uint64_t mans_max_encoded_units(uint64_t n) { return n / 2 + 2; }

}  // extern "C"

namespace {

// Compact form of an encoder table (the full one is 3 x 256 x L words, 12 MiB at
// L = 4096, and its rows are read at random): for a symbol s owning f_s states, a step
// from state x emits nb = m_s - (x < thr_s) bits (the low bits of x, the code) and moves to
// ys[base_s + (x >> nb) - f_s].  Built from the full table and checked against every one
// of its entries; a table of another shape keeps the full-table path.
struct CompactEnc {
    std::vector<uint32_t> f, base, m, thr, ys;
};

bool compact_encoder(const uint32_t* enc, uint32_t L, CompactEnc& c) {
    c.f.assign(256, 0);
    c.base.assign(256, 0);
    c.m.assign(256, 0);
    c.thr.assign(256, 0);
    c.ys.assign(L, 0);
    uint32_t used = 0;
    for (uint32_t s = 0; s < 256; ++s) {
        const uint32_t* row = enc + 3ull * s * L;
        if (!row[0]) {                       // symbol without states: its row is unset
            for (uint32_t i = 1; i < L; ++i)
                if (row[3ull * i]) return false;
            continue;
        }
        uint32_t mx = 0, thr = 2u * L;
        for (uint32_t i = 0; i < L; ++i) mx = std::max(mx, row[3ull * i + 2]);
        for (uint32_t i = 0; i < L; ++i)
            if (row[3ull * i + 2] == mx) { thr = L + i; break; }
        // f: the number of distinct next states (k = x >> nb covers [f, 2f))
        const uint32_t f = (thr >> mx);
        if (f == 0 || used + f > L) return false;
        c.f[s] = f;
        c.m[s] = mx;
        c.thr[s] = thr;
        c.base[s] = used;
        for (uint32_t i = 0; i < L; ++i) {
            const uint32_t x = L + i;
            const uint32_t* e = row + 3ull * i;
            const uint32_t nb = mx - (x < thr ? 1u : 0u);
            const uint32_t k = x >> nb;
            if (!e[0] || e[2] != nb || e[1] != (x & ((1u << nb) - 1u)) || k < f || k >= 2u * f)
                return false;
            uint32_t& y = c.ys[used + k - f];
            if (y && y != e[0]) return false;
            y = e[0];
        }
        used += f;
    }
    return true;
}

}  // namespace

extern "C" {

int mans_encode(const uint8_t* in, uint64_t n, const uint32_t* enc, uint32_t L, uint32_t* units,
                uint64_t* n_units, uint32_t* first_bit, uint32_t* first_state) {
    if ((!in && n) || !enc || !units || !n_units || !first_bit || !first_state)
        return set_error(MANS_EINVAL, "NULL argument");
    // codewords appended MSB-first to 32-bit units, state starting at L; the last
    // unit keeps `first_bit` valid bits left-aligned (zeros below)
    uint64_t acc = 0;
    uint32_t fill = 0, x = L;
    uint64_t nu = 0;
    CompactEnc c;
    if (n >= 4096u && compact_encoder(enc, L, c)) {
        for (uint64_t i = 0; i < n; ++i) {
            const uint32_t s = in[i];
            if (!c.f[s]) return set_error(MANS_EINVAL, "symbol %u has no code", s);
            const uint32_t nb = c.m[s] - (x < c.thr[s] ? 1u : 0u);
            acc = (acc << nb) | (x & ((1u << nb) - 1u));
            fill += nb;
            x = c.ys[c.base[s] + (x >> nb) - c.f[s]];
            if (fill >= 32) {
                fill -= 32;
                units[nu++] = (uint32_t)(acc >> fill);
                acc &= (1ull << fill) - 1;
            }
        }
    } else {
        for (uint64_t i = 0; i < n; ++i) {
            const uint32_t* e = enc + 3ull * ((uint64_t)in[i] * L + (x - L));
            if (!e[0]) return set_error(MANS_EINVAL, "symbol %u has no code", in[i]);
            acc = (acc << e[2]) | e[1];
            fill += e[2];
            x = e[0];
            if (fill >= 32) {
                fill -= 32;
                units[nu++] = (uint32_t)(acc >> fill);
                acc &= (1ull << fill) - 1;
            }
        }
    }
    units[nu++] = fill ? (uint32_t)(acc << (32 - fill)) : 0u;
    std::reverse(units, units + nu);
    *n_units = nu;
    *first_bit = fill;
    *first_state = x;
    return MANS_OK;
}

int mans_il_record_count(uint64_t n_units, uint32_t record_units, uint64_t* count) {
    if (!count) return set_error(MANS_EINVAL, "count is NULL");
    if (record_units < MANS_IL_MIN_RECORD_UNITS || record_units > MANS_IL_MAX_RECORD_UNITS)
        return set_error(MANS_EINVAL, "record_units=%u outside [%u, %u]", record_units,
                         MANS_IL_MIN_RECORD_UNITS, MANS_IL_MAX_RECORD_UNITS);
    *count = (n_units + record_units - 1) / record_units;
    return MANS_OK;
}

}  // extern "C"

namespace {

// One slot's encoder: the compact form when the table has its shape, else the full table.
struct SlotEnc {
    const uint32_t* full;
    uint32_t L;
    CompactEnc c;
    bool compact;
    // codeword of s from state x: returns its length (0 = s has no code), sets *code, *x
    inline int step(uint32_t s, uint32_t& x, uint32_t& code) const {
        if (compact) {
            if (!c.f[s]) return -1;
            const uint32_t nb = c.m[s] - (x < c.thr[s] ? 1u : 0u);
            code = x & ((1u << nb) - 1u);
            x = c.ys[c.base[s] + (x >> nb) - c.f[s]];
            return (int)nb;
        }
        const uint32_t* e = full + 3ull * ((uint64_t)s * L + (x - L));
        if (!e[0]) return -1;
        code = e[1];
        x = e[0];
        return (int)e[2];
    }
};

}  // namespace

extern "C" int mans_encode_interleaved(const uint8_t* in, uint64_t n, const uint32_t* enc0,
                                       const uint32_t* enc1, uint32_t L, uint32_t record_units,
                                       uint32_t* units, uint64_t* n_units, uint32_t* first_bit,
                                       uint32_t* first_states, mans_il_record* records,
                                       uint64_t records_cap, uint64_t* n_records) {
    if ((!in && n) || !enc0 || !enc1 || !units || !n_units || !first_bit || !first_states ||
        !records || !n_records)
        return set_error(MANS_EINVAL, "NULL argument");
    if (L < 2u || L > 16384u || (L & (L - 1u)))
        return set_error(MANS_EINVAL, "L=%u is not a power of two in [2, 16384]", L);
    uint64_t rc_max = 0;
    int rc = mans_il_record_count(0, record_units, &rc_max);
    if (rc) return rc;
    SlotEnc se[2];
    const uint32_t* enc[2] = {enc0, enc1};
    for (int k = 0; k < 2; ++k) {
        se[k].full = enc[k];
        se[k].L = L;
        se[k].compact = n >= 4096u && compact_encoder(enc[k], L, se[k].c);
    }
    // Forward pass.  E counts the bits emitted so far; record m belongs to forward bit
    // F_m = 32 R m: it is taken from the state before the first symbol whose codeword
    // would end beyond F_m (the last symbol ending at or before F_m is the first the
    // decoder meets at or after the mirrored boundary), so a record is written the moment
    // a codeword straddles or starts at its F_m.
    const uint64_t span = 32ull * record_units;
    uint64_t acc = 0, E = 0, F = 0, m = 0;
    uint32_t fill = 0, x[2] = {L, L};
    uint64_t nu = 0;
    auto put = [&](uint64_t before, uint32_t bit) -> bool {
        if (m >= records_cap) return false;
        records[m].state[0] = (uint16_t)(x[0] - L);
        records[m].state[1] = (uint16_t)(x[1] - L);
        records[m].bit = bit;
        records[m].before = before;
        ++m;
        F += span;
        return true;
    };
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t k = (uint32_t)(i & 1u);
        uint32_t xs = x[k], code = 0;
        const int nb = se[k].step(in[i], xs, code);
        if (nb < 0) return set_error(MANS_EINVAL, "symbol %u has no code in slot %u", in[i], k);
        if (E + (uint64_t)nb > F && !put(n - i, (uint32_t)(F - E)))
            return set_error(MANS_EINVAL, "records_cap %llu too small", (unsigned long long)records_cap);
        x[k] = xs;
        acc = (acc << nb) | code;
        fill += (uint32_t)nb;
        E += (uint64_t)nb;
        if (fill >= 32) {
            fill -= 32;
            units[nu++] = (uint32_t)(acc >> fill);
            acc &= (1ull << fill) - 1;
        }
    }
    units[nu++] = fill ? (uint32_t)(acc << (32 - fill)) : 0u;
    // boundaries at or after the last codeword's end: the decoder's first step
    const uint64_t M = (nu + record_units - 1) / record_units;
    while (m < M) {
        if (!put(0, (uint32_t)(F - E)))
            return set_error(MANS_EINVAL, "records_cap %llu too small", (unsigned long long)records_cap);
    }
    if (m != M) return set_error(MANS_EINVAL, "record count %llu != %llu", (unsigned long long)m,
                                 (unsigned long long)M);
    std::reverse(units, units + nu);
    *n_units = nu;
    *first_bit = fill;
    first_states[0] = x[0];
    first_states[1] = x[1];
    *n_records = m;
    return MANS_OK;
}

extern "C" int mans_make_stream(uint64_t seed, double lambda, uint64_t n, uint32_t L,
                                uint8_t* symbols_out, uint32_t* units_out, uint64_t units_cap,
                                uint64_t* n_units, uint32_t* first_bit, uint32_t* first_state,
                                uint32_t* dec_out) {
    if (!units_out || !n_units || !first_bit || !first_state || !dec_out)
        return set_error(MANS_EINVAL, "NULL argument");
    if (units_cap < mans_max_encoded_units(n)) return set_error(MANS_EINVAL, "units_cap too small");
    std::vector<uint8_t> tmp;
    uint8_t* sym = symbols_out;
    if (!sym) { tmp.resize(n); sym = tmp.data(); }
    // large streams are generated on several threads (callers that make many streams at
    // once run one per thread, and their streams are small)
    const unsigned hc = std::thread::hardware_concurrency();
    const int th = n >= (1ull << 26) ? (int)std::min(8u, hc ? hc : 1u) : 1;
    int rc = mans_gen_symbols(seed, lambda, sym, n, th);
    if (rc) return rc;
    uint64_t hist[256] = {0};
    for (uint64_t i = 0; i < n; ++i) ++hist[sym[i]];
    uint32_t counts[256];
    rc = mans_normalize_counts(hist, 256, L, counts);
    if (rc) return rc;
    std::vector<uint32_t> enc(3ull * 256 * L);
    rc = mans_build_tables(counts, 256, L, enc.data(), dec_out);
    if (rc) return rc;
    return mans_encode(sym, n, enc.data(), L, units_out, n_units, first_bit, first_state);
}
