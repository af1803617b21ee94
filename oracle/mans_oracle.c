/*
 * mans_oracle.c -- CPU oracle for tANS subsequence decoding.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP decoder in multians_amd/.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * Each function names the reference lines it restates (paths relative to the
 * reference repository root).  See mans_oracle.h for how each one is pinned.
 */
#define _GNU_SOURCE
#include "mans_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ symbols */

static inline uint64_t mo_mix64(uint64_t z) {
    /* splitmix64 finaliser (Steele, Lea, Flood 2014) */
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void mo_gen_symbols(uint64_t seed, double lambda, uint8_t* out, size_t n) {
    /* X ~ Exp(lambda): P(X >= k) = exp(-lambda k).  With u uniform on [0, 2^53),
     * floor(X) >= k  <=>  u < thr[k] = floor(2^53 exp(-lambda k)).
     * s = min(255, floor(X)) = #{k in 1..255 : u < thr[k]} (thr is decreasing). */
    uint64_t thr[256];
    for (int k = 0; k < 256; ++k) thr[k] = (uint64_t)floor(ldexp(exp(-lambda * (double)k), 53));
    for (size_t i = 0; i < n; ++i) {
        uint64_t u = mo_mix64(seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull) >> 11;
        int lo = 0, hi = 255; /* largest k with u < thr[k] (thr[0] = 2^53 > u) */
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (u < thr[mid]) lo = mid; else hi = mid - 1;
        }
        out[i] = (uint8_t)lo;
    }
}

/* ------------------------------------------------------------------ counts */

int mo_normalize(const uint64_t* hist, int nsym, uint32_t L, uint32_t* counts) {
    uint64_t total = 0;
    int present = 0;
    for (int s = 0; s < nsym; ++s) { total += hist[s]; present += hist[s] != 0; }
    if (total == 0 || (uint32_t)present > L) return -1;
    int64_t sum = 0;
    for (int s = 0; s < nsym; ++s) {
        if (!hist[s]) { counts[s] = 0; continue; }
        uint64_t c = (uint64_t)(((__uint128_t)hist[s] * L) / total);
        counts[s] = c ? (uint32_t)c : 1u;
        sum += counts[s];
    }
    /* surplus/deficit goes to (comes from) the largest count first */
    while (sum != (int64_t)L) {
        int best = -1;
        for (int s = 0; s < nsym; ++s) {
            if (!counts[s]) continue;
            if (sum > (int64_t)L && counts[s] <= 1) continue;
            if (best < 0 || counts[s] > counts[best]) best = s;
        }
        if (best < 0) return -1;
        if (sum < (int64_t)L) { counts[best] += (uint32_t)((int64_t)L - sum); sum = L; }
        else { counts[best] -= 1; sum -= 1; }
    }
    return 0;
}

/* ------------------------------------------------------------------ tables */

int mo_build_tables(const uint32_t* counts, int nsym, uint32_t L, mo_enc_entry* enc,
                    uint32_t* dec) {
    /* Spread: ans_table_generator.cc:164-225 (priority queue over p = 0.5/P_s,
     * first minimum wins, p += 1/P_s; state i gets (s, X_s++), X_s starts at L_s). */
    double* P = (double*)malloc(sizeof(double) * nsym);
    double* q = (double*)malloc(sizeof(double) * nsym);
    uint32_t* X = (uint32_t*)malloc(sizeof(uint32_t) * nsym);
    uint8_t* ps = (uint8_t*)malloc(L);
    uint32_t* px = (uint32_t*)malloc(sizeof(uint32_t) * L);
    uint8_t* nb = (uint8_t*)calloc(L, 1);
    uint64_t sum = 0;
    for (int s = 0; s < nsym; ++s) {
        sum += counts[s];
        P[s] = (double)counts[s] / (double)L;    /* prob_ans = count / N (:152-153) */
        q[s] = 0.5 / P[s];                        /* 0.5f / P_s (:184); inf if absent */
        X[s] = counts[s];
    }
    int rc = 0;
    if (sum != L) { rc = -1; goto done; }
    for (uint32_t i = 0; i < L; ++i) {
        int best = 0;
        for (int s = 1; s < nsym; ++s) if (q[s] < q[best]) best = s;
        q[best] = q[best] + (1.0 / P[best]);
        ps[i] = (uint8_t)best;
        px[i] = X[best]++;
    }
    /* Encoder table: ans_table_generator.cc:234-253.  State i = L+ii assigned to
     * (s, Xv) serves every x in [L, 2L) with x >> k == Xv, emitting k bits x & (2^k-1). */
    memset(enc, 0, sizeof(mo_enc_entry) * (size_t)nsym * L);
    for (uint32_t ii = 0; ii < L; ++ii) {
        uint32_t s = ps[ii], Xv = px[ii];
        for (uint32_t k = 0; ((uint64_t)Xv << k) < 2ull * L; ++k) {
            uint64_t lo = (uint64_t)Xv << k, hi = lo + (1ull << k);
            if (lo < L) lo = L;
            if (hi > 2ull * L) hi = 2ull * L;
            for (uint64_t x = lo; x < hi; ++x) {
                mo_enc_entry* e = &enc[(size_t)s * L + (x - L)];
                e->next_state = L + ii;
                e->code = (uint32_t)(x & ((1ull << k) - 1));
                e->len = k;
            }
        }
    }
    /* Decoder table: ans_table_generator.cc:261-288, including its choice of
     * min_num_bits (first nonzero length met in (symbol, state) order).  Rows of
     * absent symbols are skipped (the reference would index state 0 - L there). */
    memset(dec, 0, sizeof(uint32_t) * L);
    for (int s = 0; s < nsym; ++s) {
        for (uint32_t j = 0; j < L; ++j) {
            const mo_enc_entry* e = &enc[(size_t)s * L + j];
            if (!e->next_state) continue;
            uint32_t prev = e->next_state - L, len = e->len;
            uint32_t nextv = ((L + j) >> len) & 0xFFFFu;
            uint32_t mb = (nb[prev] == 0 && len > 0) ? len : nb[prev];
            nb[prev] = (uint8_t)mb;
            dec[prev] = nextv | ((uint32_t)s << 16) | (mb << 24);
        }
    }
    for (uint32_t i = 0; i < L; ++i) if ((dec[i] & 0xFFFFu) == 0) { rc = -2; break; }
done:
    free(P); free(q); free(X); free(ps); free(px); free(nb);
    return rc;
}

/* ------------------------------------------------------------------ encoder */

size_t mo_max_units(size_t n, uint32_t L) {
    (void)L;
    return n / 2 + 2; /* codewords are <= 16 bits */
}

size_t mo_encode(const uint8_t* in, size_t n, const mo_enc_entry* enc, uint32_t L,
                 uint32_t* out_units, uint32_t* first_bit, uint32_t* first_state) {
    /* ans_encoder.cc:13-72: state starts at index 0 (x = L); each codeword is
     * appended MSB-first to 32-bit units; the final unit holds `first_bit` valid
     * bits left-aligned.  The unit count is floor(total_bits/32) + 1 and
     * first_bit = total_bits mod 32 (the reference also emits a final unit with no
     * valid bits when the total is a multiple of 32).  Units are then reversed
     * into decoder order (multians_input_buffer.cc:29). */
    uint64_t acc = 0;
    uint32_t nacc = 0, x = L;
    size_t nu = 0;
    for (size_t i = 0; i < n; ++i) {
        const mo_enc_entry* e = &enc[(size_t)in[i] * L + (x - L)];
        if (!e->next_state) return 0;
        acc = (acc << e->len) | e->code;
        nacc += e->len;
        x = e->next_state;
        if (nacc >= 32) {
            nacc -= 32;
            out_units[nu++] = (uint32_t)(acc >> nacc);
            acc &= (nacc ? ((1ull << nacc) - 1) : 0);
        }
    }
    out_units[nu++] = nacc ? (uint32_t)(acc << (32 - nacc)) : 0u;
    *first_bit = nacc;
    *first_state = x;
    for (size_t a = 0, b = nu - 1; a < b; ++a, --b) {
        uint32_t t = out_units[a]; out_units[a] = out_units[b]; out_units[b] = t;
    }
    return nu;
}

/* ------------------------------------------------------------------ decoding */

static inline uint32_t mo_unit(const uint32_t* u, size_t n, uint64_t i) {
    return i < n ? u[i] : 0u;
}

/* k bits at absolute position p, LSB-first (k <= 32). */
static inline uint32_t mo_bits(const uint32_t* u, size_t n, uint64_t p, uint32_t k) {
    if (!k) return 0;
    uint64_t w = ((uint64_t)mo_unit(u, n, (p >> 5) + 1) << 32) | mo_unit(u, n, p >> 5);
    w >>= (p & 31);
    return (uint32_t)(w & ((1ull << k) - 1));
}

/* One decoding step, multians_gpu_decoder.cu:131-166: lookup, take nbits bits
 * below next<<nbits, then extend one bit at a time while the state is below L. */
static inline int mo_step(const uint32_t* units, size_t n_units, const uint32_t* table,
                          uint32_t L, uint32_t* x, uint64_t* p, uint8_t* sym) {
    if (*x < L || *x >= 2 * L) return -1;
    uint32_t e = table[*x - L];
    uint32_t nbits = e >> 24;
    uint32_t nx = ((e & 0xFFFFu) << nbits) | mo_bits(units, n_units, *p, nbits);
    uint64_t q = *p + nbits;
    int guard = 0;
    while (nx < L) {
        if (++guard > 32) return -1;
        nx = (nx << 1) | mo_bits(units, n_units, q, 1);
        ++q;
    }
    *sym = (uint8_t)(e >> 16);
    *x = nx;
    *p = q;
    return 0;
}

int mo_decode_seq(const uint32_t* units, size_t n_units, const uint32_t* table, uint32_t L,
                  uint32_t first_state, uint32_t first_bit, uint8_t* out, size_t n_out,
                  uint32_t S, uint32_t* entries, uint32_t* counts, uint32_t* end_state,
                  uint64_t* end_bit) {
    uint64_t p = 32u - first_bit;
    uint32_t x = first_state;
    size_t n_sub = S ? (n_units + S - 1) / S : 0;
    size_t next_entry = 0;
    uint64_t sub_bits = 32ull * S;
    if (S) {
        for (size_t j = 0; j < n_sub; ++j) { entries[j] = 0xFFFFFFFFu; counts[j] = 0; }
    }
    for (size_t k = 0; k < n_out; ++k) {
        if (S) {
            while (next_entry < n_sub && p >= sub_bits * next_entry) {
                entries[next_entry] = (x - L) | (uint32_t)((p - sub_bits * next_entry) << 16);
                ++next_entry;
            }
            size_t j = (size_t)(p / sub_bits);
            if (j >= n_sub) j = n_sub - 1;
            counts[j] += 1;
        }
        uint8_t s;
        if (mo_step(units, n_units, table, L, &x, &p, &s)) return -1;
        out[k] = s;
    }
    if (end_state) *end_state = x;
    if (end_bit) *end_bit = p;
    return 0;
}

/* Last step owned by each subsequence of S units (the step the reference's phase 1
 * records per subsequence: multians_gpu_decoder.cu:160-168 keeps last_state, last_word_unit
 * and last_word_bit of the last codeword that starts in the subsequence), over the first
 * n_out steps: its state before the step and its absolute start bit.  A step belongs to
 * the subsequence holding the unit it starts in; steps at or beyond unit n_sub*S belong to
 * the last one.  Subsequences owning no step keep last_state 0. */
int mo_subseq_last(const uint32_t* units, size_t n_units, const uint32_t* table, uint32_t L,
                   uint32_t first_state, uint32_t first_bit, size_t n_out, uint32_t S,
                   uint32_t* last_state, uint64_t* last_bit) {
    uint64_t p = 32u - first_bit;
    uint32_t x = first_state;
    size_t n_sub = (n_units + S - 1) / S;
    for (size_t j = 0; j < n_sub; ++j) { last_state[j] = 0; last_bit[j] = 0; }
    for (size_t k = 0; k < n_out; ++k) {
        size_t j = (size_t)(p / (32ull * S));
        if (j >= n_sub) j = n_sub - 1;
        last_state[j] = x;
        last_bit[j] = p;
        uint8_t s;
        if (mo_step(units, n_units, table, L, &x, &p, &s)) return -1;
    }
    return 0;
}


/* ------------------------------------------------- interleaved-state streams */
/* No reference counterpart: the reference names "multiple tables / multiple states" as
 * future work (README.md:11).  Format restated here (own design): N states share one bit
 * stream; input symbol i is encoded by state slot k = i mod N with that slot's encoder
 * table, in input order, each code appended MSB-first exactly as the single-state encoder
 * (ans_encoder.cc:13-72) appends it; every slot starts at x = L; after the last symbol the
 * slots' states are the decoder's first states; units are reversed into decoder order.
 * Decoding (decode order j = 0 .. n-1 yields input symbol n-1-j) runs the single-state
 * step (multians_gpu_decoder.cu:127-169) on slot (n-1-j) mod N with that slot's decoder
 * table.  With N = 1 both reduce to mo_encode / mo_decode_seq. */
size_t mo_encode_interleaved(const uint8_t* in, size_t n, const mo_enc_entry* const* enc,
                             uint32_t L, uint32_t N, uint32_t* out_units, uint32_t* first_bit,
                             uint32_t* first_states) {
    uint64_t acc = 0;
    uint32_t nacc = 0, x[16];
    size_t nu = 0;
    if (N == 0 || N > 16) return 0;
    for (uint32_t k = 0; k < N; ++k) x[k] = L;
    for (size_t i = 0; i < n; ++i) {
        const uint32_t k = (uint32_t)(i % N);
        const mo_enc_entry* e = &enc[k][(size_t)in[i] * L + (x[k] - L)];
        if (!e->next_state) return 0;
        acc = (acc << e->len) | e->code;
        nacc += e->len;
        x[k] = e->next_state;
        if (nacc >= 32) {
            nacc -= 32;
            out_units[nu++] = (uint32_t)(acc >> nacc);
            acc &= (nacc ? ((1ull << nacc) - 1) : 0);
        }
    }
    out_units[nu++] = nacc ? (uint32_t)(acc << (32 - nacc)) : 0u;
    *first_bit = nacc;
    for (uint32_t k = 0; k < N; ++k) first_states[k] = x[k];
    for (size_t a = 0, b = nu - 1; a < b; ++a, --b) {
        uint32_t t = out_units[a]; out_units[a] = out_units[b]; out_units[b] = t;
    }
    return nu;
}

int mo_decode_interleaved(const uint32_t* units, size_t n_units, const uint32_t* const* tables,
                          uint32_t L, uint32_t N, const uint32_t* first_states,
                          uint32_t first_bit, uint8_t* out, size_t n_out, size_t n_total,
                          uint32_t R, mo_il_record* records, uint32_t* end_states,
                          uint64_t* end_bit) {
    uint64_t p = 32u - first_bit;
    uint32_t x[16];
    if (N == 0 || N > 16) return -1;
    if (R && (N > 2 || n_out != n_total)) return -1;
    for (uint32_t k = 0; k < N; ++k) x[k] = first_states[k];
    /* records are crossed in decreasing m (increasing T_m) */
    int64_t m = R ? (int64_t)((n_units + R - 1) / R) - 1 : -1;
    for (size_t j = 0; j <= n_out; ++j) {
        while (m >= 0 && p >= 32ull * (n_units - (uint64_t)R * (uint64_t)m)) {
            for (uint32_t k = 0; k < 2; ++k)
                records[m].state[k] = (uint16_t)(k < N ? x[k] - L : 0u);
            records[m].bit = (uint32_t)(p - 32ull * (n_units - (uint64_t)R * (uint64_t)m));
            records[m].before = j;
            --m;
        }
        if (j == n_out) break;
        const uint32_t k = (uint32_t)((n_total - 1 - j) % N);
        uint8_t s;
        if (mo_step(units, n_units, tables[k], L, &x[k], &p, &s)) return -1;
        out[j] = s;
    }
    if (m >= 0) return -1;      /* a record past the stream's end: not a valid stream */
    if (end_states) for (uint32_t k = 0; k < N; ++k) end_states[k] = x[k];
    if (end_bit) *end_bit = p;
    return 0;
}

/* ------------------------------------------------------- multicore decoder */
/* Restatement of MulticoreDecoder (multicore_decoder.cc:17-333): T threads, each
 * owning an interval of subsequences (:275-303); a speculative pass from the
 * first state at bit 0 (:125-128), repeated overflow rounds from the previous
 * thread's last sync point until a thread re-meets its own stored sync point
 * (:136-143, :223-236), a per-thread prefix of symbol counts (:305-333), and a
 * write pass (:79-89) clipped at the output size.  Steps starting at or beyond
 * unit ceil(n/S)*S are never decoded, so bytes they would produce keep their old
 * contents (the reference behaves the same way).  The 32-bit window registers
 * follow the reference exactly, including its first window for at == 0, built
 * with 32-bit shifts by 32 that x86-64 performs modulo 32 (so the speculative
 * paths, and where they meet, are the reference's).  One difference: each round
 * reads its start point from a snapshot taken before the round (the reference
 * reads it while the previous thread may be rewriting it, a data race). */

typedef struct { uint32_t state, bit, unit, num_symbols; } mo_mc_sync;

typedef struct {
    const uint32_t* in;
    size_t n_units;
    const uint32_t* table;
    uint32_t L, first_state, first_bit;
    uint8_t* out;
    size_t n_out, S, T;
    mo_mc_sync* sync;
    mo_mc_sync* start_snap; /* per thread: sync[begin-1] at round start */
    int* synced;
    size_t* out_positions;
} mo_mc_ctx;

typedef struct {
    mo_mc_ctx* c;
    size_t tid, begin, end, sub;
    int overflow, write;
    int err;
} mo_mc_job;

/* 32-bit shifts as x86-64 executes them for a 32-bit operand: the count is taken
 * modulo 32 (the reference shifts by 32 when at == 0 or at == 32, :167-172). */
static inline uint32_t mo_shl32(uint32_t v, uint32_t n) { return v << (n & 31u); }
static inline uint32_t mo_shr32(uint32_t v, uint32_t n) { return v >> (n & 31u); }

static void* mo_mc_phase(void* arg) {
    mo_mc_job* jb = (mo_mc_job*)arg;
    mo_mc_ctx* c = jb->c;
    const size_t tid = jb->tid, S = c->S;
    if (jb->overflow && c->synced[tid]) return NULL;
    const uint32_t L = c->L;
    const uint32_t* in = c->in;
    uint32_t state = c->first_state;
    uint32_t at = (tid == 0) ? 32u - c->first_bit : 0u;
    size_t in_pos = jb->begin, out_pos = 0, out_size = 0, cur_sub = jb->sub;
    uint32_t cur_unit = 0;
    if (jb->overflow || (jb->write && tid > 0)) {
        mo_mc_sync sp = c->start_snap[tid];
        state = sp.state;
        at = sp.bit;
        in_pos = in_pos - S + sp.unit;
        cur_unit = sp.unit;
    }
    if (jb->write) {
        out_pos = c->out_positions[tid];
        out_size = (tid < c->T - 1) ? c->out_positions[tid + 1] : c->n_out;
    }
    /* window/next registers exactly as :157-172 (including the modulo-32 shifts) */
    uint32_t window = in[in_pos], next = in[in_pos + 1];
    uint32_t copy_next = mo_shl32(next, 32u - at);
    next = mo_shr32(next, at);
    window = mo_shr32(window, at);
    window += copy_next;
    uint32_t last_state = 0, last_bit = 0, num_symbols = 0;
    int reset = jb->write && tid == 0;
    while (in_pos < jb->end) {
        while (at < 32) {
            last_state = state;
            if (state < L || state >= 2 * L) { jb->err = -1; return NULL; }
            uint32_t e = c->table[state - L];
            uint32_t taken = e >> 24;
            ++num_symbols;
            uint32_t low = taken ? (window & (0xFFFFFFFFu >> (32u - taken))) : 0u;
            state = ((e & 0xFFFFu) << taken) + low;
            int guard = 0;
            while (state < L) {                          /* :193-197 */
                if (++guard > 32) { jb->err = -1; return NULL; }
                uint32_t sh = mo_shr32(window, taken);
                ++taken;
                state = (state << 1) + (sh & 1u);
            }
            if (jb->write && reset && out_pos < out_size) c->out[out_pos++] = (uint8_t)(e >> 16);
            copy_next = taken > 0 ? mo_shl32(next, 32u - taken) : 0u;
            last_bit = at;
            next = mo_shr32(next, taken);
            window = mo_shr32(window, taken);
            at += taken;
            window += copy_next;
        }
        ++in_pos;
        ++cur_unit;
        if (cur_unit == S) {
            if (jb->overflow && reset) {
                mo_mc_sync sp = c->sync[cur_sub];
                if (sp.state == last_state && sp.bit == last_bit && sp.unit == cur_unit - 1) {
                    c->sync[cur_sub].num_symbols = num_symbols;
                    c->synced[tid] = 1;
                    return NULL;
                }
            }
            if (!jb->overflow || reset) {
                if (!jb->write) {
                    mo_mc_sync v = {last_state, last_bit, cur_unit - 1, num_symbols};
                    c->sync[cur_sub] = v;
                }
                ++cur_sub;
            }
            if (jb->overflow && in_pos > c->n_units) c->synced[tid] = 1;
            reset = 1;
            cur_unit = 0;
            num_symbols = 0;
        }
        window = in[in_pos];                              /* :256-271 */
        next = in[in_pos + 1];
        if (at == 32) {
            at = 0;
        } else {
            at -= 32;
            window = mo_shr32(window, at);
            next = mo_shr32(next, at);
            window += mo_shl32(in[in_pos + 1], 32u - at);
        }
    }
    return NULL;
}

static int mo_mc_run(mo_mc_ctx* c, mo_mc_job* jobs, size_t first, int overflow, int write) {
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * c->T);
    for (size_t i = first; i < c->T; ++i) {
        jobs[i].overflow = overflow;
        jobs[i].write = write;
        if (i > 0) c->start_snap[i] = c->sync[jobs[i].sub - 1];
    }
    for (size_t i = first; i < c->T; ++i) pthread_create(&th[i], NULL, mo_mc_phase, &jobs[i]);
    int err = 0;
    for (size_t i = first; i < c->T; ++i) {
        pthread_join(th[i], NULL);
        if (jobs[i].err) err = -1;
    }
    free(th);
    return err;
}

int mo_decode_multicore(size_t S, size_t T, const uint32_t* units, size_t n_units,
                        const uint32_t* table, uint32_t L, uint32_t first_state,
                        uint32_t first_bit, uint8_t* out, size_t n_out) {
    if (!S || !T) return -1;
    size_t nsub = (n_units + S - 1) / S;
    if (nsub < T) return -1; /* the reference asserts this (:284) */
    mo_mc_ctx c;
    c.in = units; c.n_units = n_units; c.table = table; c.L = L;
    c.first_state = first_state; c.first_bit = first_bit; c.out = out; c.n_out = n_out;
    c.S = S; c.T = T;
    c.sync = (mo_mc_sync*)calloc(nsub, sizeof(mo_mc_sync));
    c.start_snap = (mo_mc_sync*)calloc(T, sizeof(mo_mc_sync));
    c.synced = (int*)calloc(T, sizeof(int));
    c.out_positions = (size_t*)calloc(T, sizeof(size_t));
    mo_mc_job* jobs = (mo_mc_job*)calloc(T, sizeof(mo_mc_job));
    /* intervals, :275-303 */
    size_t spt = nsub / T, rem = nsub % T, at = 0;
    for (size_t i = 0; i < T; ++i) {
        jobs[i].c = &c; jobs[i].tid = i;
        jobs[i].begin = at; jobs[i].end = at + spt * S; jobs[i].sub = spt * i;
        at += spt * S;
    }
    jobs[T - 1].end += rem * S;
    int err = mo_mc_run(&c, jobs, 0, 0, 0);
    int all = 0;
    size_t rounds = 0;
    while (!err && !all) {
        err = mo_mc_run(&c, jobs, 1, 1, 0);
        all = 1;
        for (size_t i = 1; i < T; ++i) if (!c.synced[i]) all = 0;
        if (++rounds > T + 2) { err = -1; break; }
    }
    if (!err) {
        /* prefix_sum, :305-333 (32-bit running sum as in the reference) */
        uint32_t sum = 0;
        size_t pos = 0, nxt = spt;
        c.out_positions[0] = 0;
        for (size_t i = 1; i < T; ++i) {
            for (; pos < nxt; ++pos) sum += c.sync[pos].num_symbols;
            c.out_positions[i] = sum;
            pos = nxt;
            nxt += spt;
        }
        err = mo_mc_run(&c, jobs, 0, 0, 1);
    }
    free(c.sync); free(c.start_snap); free(c.synced); free(c.out_positions); free(jobs);
    return err;
}
