// ref_driver.cc -- thin extern "C" driver over the reference's own CPU classes.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile together with the
// reference's CPU sources where they lie (never copied), into oracle/_ref/.
// It lets tests pin the oracle restatement against the real reference code:
//   ANSTableGenerator (src/ans_table_generator.cc), ANSEncoder (src/ans_encoder.cc),
//   CUHDInputBuffer / CUHDOutputBuffer / CUHDCodetable (src/multians_*.cc),
//   MulticoreDecoder (src/multicore_decoder.cc).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include "multians.h"

namespace {

struct Tables {
    std::shared_ptr<ANSEncoderTable> enc;
    std::shared_ptr<CUHDCodetable> dec;
};

Tables tables_from(Distribution& d, size_t nsym, size_t L) {
    auto t = ANSTableGenerator::generate_table(d.prob, d.dist, nullptr, nsym, L);
    Tables r;
    r.enc = ANSTableGenerator::generate_encoder_table(t);
    r.dec = ANSTableGenerator::get_decoder_table(r.enc);
    return r;
}

Distribution dist_from_counts(const uint32_t* counts, size_t nsym, size_t L) {
    auto prob = std::make_shared<std::vector<double>>(nsym);
    auto dist = std::make_shared<std::vector<size_t>>(nsym);
    for (size_t i = 0; i < nsym; ++i) {
        (*dist)[i] = counts[i];
        (*prob)[i] = (double)counts[i] / L;   // as generate_distribution returns it
    }
    return {prob, dist, nullptr};
}

void export_tables(const Tables& t, size_t nsym, size_t L, uint32_t* dec_out,
                   uint32_t* enc_next, uint32_t* enc_code, uint32_t* enc_len) {
    std::memcpy(dec_out, t.dec->get(), L * sizeof(uint32_t));
    if (enc_next) {
        for (size_t s = 0; s < nsym; ++s)
            for (size_t j = 0; j < L; ++j) {
                const auto& e = t.enc->table.at(s).at(j);
                enc_next[s * L + j] = e.next_state;
                enc_code[s * L + j] = e.code_sequence;
                enc_len[s * L + j] = e.code_length;
            }
    }
}

}  // namespace

extern "C" {

// Demo distribution + tables exactly as src/main.cc:55-73 builds them.
int ref_demo_tables(double lambda, uint32_t nsym, uint32_t L, uint32_t seed,
                    uint32_t* counts_out, uint32_t* dec_out, uint32_t* enc_next,
                    uint32_t* enc_code, uint32_t* enc_len) {
    float lf = (float)lambda;
    auto d = ANSTableGenerator::generate_distribution(
        seed, nsym, L, [&](double x) { return lf * exp(-lf * x); });
    for (size_t i = 0; i < nsym; ++i) counts_out[i] = (uint32_t)d.dist->at(i);
    Tables t = tables_from(d, nsym, L);
    export_tables(t, nsym, L, dec_out, enc_next, enc_code, enc_len);
    return 0;
}

// Tables from explicit counts (every count >= 1; the reference cannot take zeros).
int ref_tables_from_counts(const uint32_t* counts, uint32_t nsym, uint32_t L,
                           uint32_t* dec_out, uint32_t* enc_next, uint32_t* enc_code,
                           uint32_t* enc_len) {
    for (uint32_t i = 0; i < nsym; ++i) if (!counts[i]) return -1;
    auto d = dist_from_counts(counts, nsym, L);
    Tables t = tables_from(d, nsym, L);
    export_tables(t, nsym, L, dec_out, enc_next, enc_code, enc_len);
    return 0;
}

// src/ans_table_generator.cc:133-162 (symbols drawn from the spread counts).
int ref_gen_test_data(const uint32_t* counts, uint32_t nsym, size_t size, uint32_t L,
                      uint32_t seed, uint8_t* out) {
    auto dist = std::make_shared<std::vector<size_t>>(counts, counts + nsym);
    auto v = ANSTableGenerator::generate_test_data(dist, size, L, seed);
    std::memcpy(out, v->data(), size);
    return 0;
}

// ANSEncoder::encode with tables from counts; units returned in DECODER order
// (the CUHDInputBuffer order).  `in` must be readable at in[n] (the reference
// encoder reads one symbol past the end, src/ans_encoder.cc:32-33).
long ref_encode(const uint32_t* counts, uint32_t nsym, uint32_t L, const uint8_t* in,
                size_t n, uint32_t* out_units, size_t max_units, uint32_t* first_bit,
                uint32_t* first_state) {
    for (uint32_t i = 0; i < nsym; ++i) if (!counts[i]) return -1;
    auto d = dist_from_counts(counts, nsym, L);
    Tables t = tables_from(d, nsym, L);
    std::vector<uint8_t> buf(in, in + n + 1);
    auto ib = ANSEncoder::encode(buf.data(), n, t.enc);
    size_t nu = ib->get_compressed_size();
    if (nu > max_units) return -2;
    std::memcpy(out_units, ib->get_compressed_data(), nu * sizeof(uint32_t));
    *first_bit = (uint32_t)ib->get_first_bit();
    *first_state = (uint32_t)ib->get_first_state();
    return (long)nu;
}

// MulticoreDecoder::decode (src/multicore_decoder.cc:17-90) on a stream given in
// decoder order.  out[] is a fresh zero-filled CUHDOutputBuffer, copied back.
// Returns the decode wall time in microseconds (steady_clock), or < 0.
double ref_multicore(size_t S, size_t T, const uint32_t* units, size_t n_units,
                     const uint32_t* table, uint32_t L, uint32_t first_state,
                     uint32_t first_bit, uint8_t* out, size_t n_out) {
    size_t nsub = (n_units + S - 1) / S;
    if (nsub < T) return -1.0;
    std::vector<uint32_t> enc_order(units, units + n_units);
    std::reverse(enc_order.begin(), enc_order.end());
    auto in = std::make_shared<CUHDInputBuffer>(enc_order.data(), n_units, first_bit,
                                                first_state);
    auto tab = std::make_shared<CUHDCodetable>(L);
    std::memcpy(tab->get(), table, L * sizeof(uint32_t));
    auto ob = std::make_shared<CUHDOutputBuffer>(n_out);
    auto t0 = std::chrono::steady_clock::now();
    MulticoreDecoder::decode(S, T, n_units, ob, in, tab);
    auto t1 = std::chrono::steady_clock::now();
    std::memcpy(out, ob->get_decompressed_data().get(), n_out);
    return std::chrono::duration<double, std::micro>(t1 - t0).count();
}

}  // extern "C"
