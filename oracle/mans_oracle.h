/*
 * mans_oracle.h -- CPU oracle for tANS subsequence decoding (TEST INFRASTRUCTURE ONLY).
 *
 * This library is the checker.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product (multians_amd/) never does.
 *
 * It restates, in plain C, the reference's algorithm for the decode path:
 *   - tANS spread / encoder table / decoder table    src/ans_table_generator.cc:164-288
 *   - single-state encoder + stream format            src/ans_encoder.cc:13-97,
 *                                                     src/multians_input_buffer.cc:15-30
 *   - one decoding step (table lookup, bit refill)    src/multians_gpu_decoder.cu:127-169
 *   - MulticoreDecoder (the CPU decoder)              src/multicore_decoder.cc:17-333
 * plus a canonical sequential decoder that also reports, per subsequence of S
 * units, the true entry point (bit, state) and the number of owned steps.
 *
 * How each function is pinned (see tests/test_oracle.py):
 *   mo_build_tables      vs the compiled reference ANSTableGenerator (oracle/_ref)
 *                        on the demo distributions; invariants (every state filled,
 *                        decode(encode(x)) == x).
 *   mo_encode            vs the compiled reference ANSEncoder (units, first_bit,
 *                        first_state; don't-care tail bits masked); integrity
 *                        invariant (decoder ends at state L, bit 32*n_units).
 *   mo_decode_seq        round trip against the original symbols; brute-force
 *                        re-derivation of entries/counts on tiny streams.
 *   mo_subseq_last       brute-force re-derivation on tiny streams; one step from each
 *                        record lands on the next subsequence's entry (mo_decode_seq).
 *   mo_encode_interleaved / mo_decode_interleaved (own format, no reference code)
 *                        N = 1 equals mo_encode / mo_decode_seq unit for unit; N = 2, 3
 *                        against an independent pure-Python encoder and decoder on tiny
 *                        streams (one and two tables); round trip; end invariant (every
 *                        slot back at L, bit 32 n_units); records by brute force.
 *   mo_decode_multicore  vs the compiled reference MulticoreDecoder (same S, T),
 *                        including the streams where the reference drops trailing
 *                        zero-bit symbols.
 *   mo_gen_symbols       closed form: empirical frequencies vs exp(-lambda k).
 */
#ifndef MANS_ORACLE_H_
#define MANS_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Encoder table row entry: enc[s * L + (x - L)] for symbol s, state x in [L, 2L). */
typedef struct {
    uint32_t next_state; /* state after encoding s from x, in [L, 2L); 0 = unset */
    uint32_t code;       /* low `len` bits of x, emitted MSB-first              */
    uint32_t len;        /* number of emitted bits                               */
} mo_enc_entry;

/* symbols s = min(255, floor(X)), X ~ Exp(lambda), from a counter-based PRNG.
 * Element i depends only on (seed, i). */
void mo_gen_symbols(uint64_t seed, double lambda, uint8_t* out, size_t n);

/* Histogram -> counts summing to L (absent -> 0, present -> >= 1). */
int mo_normalize(const uint64_t* hist, int nsym, uint32_t L, uint32_t* counts);

/* Reference distribution of the demo (ans_table_generator.cc:18-59) needs libstdc++
 * <random>; it is taken from oracle/_ref, not restated. */

/* Spread + encoder table + decoder table (reference layout u32: next | sym<<16 | nbits<<24).
 * counts[nsym] sum to L; zero counts allowed (those rows stay unset). */
int mo_build_tables(const uint32_t* counts, int nsym, uint32_t L,
                    mo_enc_entry* enc /* nsym*L */, uint32_t* dec /* L */);

/* Encode in[0..n) -> units in DECODER order (reversed), returns n_units.
 * out must hold mo_max_units(n, L) units.  Don't-care tail bits are zero. */
size_t mo_max_units(size_t n, uint32_t L);
size_t mo_encode(const uint8_t* in, size_t n, const mo_enc_entry* enc, uint32_t L,
                 uint32_t* out_units, uint32_t* first_bit, uint32_t* first_state);

/* Canonical sequential decoder.
 *   units[n_units] (+ reads beyond n_units treated as zero), table[L]
 *   writes out[n_out] symbols (decode order).
 *   If S > 0: entries[n_sub], counts[n_sub] with n_sub = ceil(n_units / S):
 *     entries[j] = (x - L) | (p - 32*S*j) << 16 : first step with bit position p >= 32*S*j
 *     counts[j]  = number of the first n_out steps owned by subsequence j
 *                  (the last subsequence also owns every later step).
 *   end_state/end_bit: state and absolute bit position after n_out steps.
 *   Returns 0, or -1 if the walk leaves [L, 2L) (invalid table). */
int mo_decode_seq(const uint32_t* units, size_t n_units, const uint32_t* table, uint32_t L,
                  uint32_t first_state, uint32_t first_bit, uint8_t* out, size_t n_out,
                  uint32_t S, uint32_t* entries, uint32_t* counts,
                  uint32_t* end_state, uint64_t* end_bit);

/* Last step owned by each of the ceil(n_units / S) subsequences within the first n_out
 * steps: last_state[j] = state before it, last_bit[j] = its absolute start bit (the step
 * whose position the reference's phase 1 records per subsequence,
 * multians_gpu_decoder.cu:160-168).  Steps beyond the last subsequence's units count as
 * its own.  Returns 0 or -1. */
int mo_subseq_last(const uint32_t* units, size_t n_units, const uint32_t* table, uint32_t L,
                   uint32_t first_state, uint32_t first_bit, size_t n_out, uint32_t S,
                   uint32_t* last_state, uint64_t* last_bit);

/* Interleaved-state streams (own format; the reference names multiple states / tables as
 * future work, README.md:11).  N states (slots) share one bit stream: input symbol i is
 * encoded by slot i mod N with encoder table enc[i mod N]; every slot starts at x = L;
 * first_states[k] = slot k's final state; units in decoder order.  Decode step j
 * (yielding input symbol n_total-1-j) runs the single-state step on slot
 * (n_total-1-j) mod N with decoder table tables[that slot].  N = 1 is the single-state
 * format (mo_encode / mo_decode_seq).
 * Records (R > 0, N <= 2, n_out = n_total), anchored at the stream's end so that an
 * encoder can emit them in its one forward pass: record m (m = 0 .. ceil(n_units/R) - 1)
 * is at unit boundary T_m = n_units - R*m (bit 32*T_m of the decoder's stream, forward
 * bit 32*R*m of the encoder's).  It holds, for the first step j (0 <= j <= n_total; j =
 * n_total is the end) whose start bit p_j >= 32*T_m: the slots' states before it (x - L),
 * bit = p_j - 32*T_m, and before = j.  Record 0 is the end (usually states L, bit 0,
 * before n_total; earlier when the last decoded symbols take zero bits). */
typedef struct {
    uint16_t state[2];
    uint32_t bit;
    uint64_t before;
} mo_il_record;
size_t mo_encode_interleaved(const uint8_t* in, size_t n, const mo_enc_entry* const* enc,
                             uint32_t L, uint32_t N, uint32_t* out_units, uint32_t* first_bit,
                             uint32_t* first_states);
int mo_decode_interleaved(const uint32_t* units, size_t n_units, const uint32_t* const* tables,
                          uint32_t L, uint32_t N, const uint32_t* first_states,
                          uint32_t first_bit, uint8_t* out, size_t n_out, size_t n_total,
                          uint32_t R, mo_il_record* records, uint32_t* end_states,
                          uint64_t* end_bit);

/* Restatement of MulticoreDecoder::decode (multicore_decoder.cc:17-90) with pthreads.
 * units must be padded by >= S+2 zero units after n_units (the reference pads 4 and
 * reads up to ceil(n/S)*S+1).  out is NOT cleared: bytes the method never writes
 * keep their previous contents, as in the reference.  Returns 0 or -1. */
int mo_decode_multicore(size_t S, size_t T, const uint32_t* units, size_t n_units,
                        const uint32_t* table, uint32_t L, uint32_t first_state,
                        uint32_t first_bit, uint8_t* out, size_t n_out);

#ifdef __cplusplus
}
#endif

#endif
