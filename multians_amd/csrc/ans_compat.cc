// ans_compat.cc -- ANSTableGenerator / ANSEncoder of the drop-in C++ layer
// (include/ans_table_generator.h, include/ans_encoder.h).
//
// The demo distribution and test data follow the reference's arithmetic and random
// draws step for step (src/ans_table_generator.cc:18-59, :133-162: std::mt19937,
// float accumulation of the probabilities, two draws per accepted sample), so code
// shaped like src/main.cc produces the same symbols and tables.  The spread, the
// encoder table and the decoder table follow :164-288 (with absent symbols skipped).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <queue>
#include <random>
#include <unordered_map>

#include "ans_encoder.h"
#include "ans_table_generator.h"

namespace {

// first smallest key wins, as std::min_element over the queue does
struct SpreadKey {
    double p;
    std::size_t idx;
    bool operator>(const SpreadKey& o) const { return p > o.p || (p == o.p && idx > o.idx); }
};

}  // namespace

Distribution ANSTableGenerator::generate_distribution(size_t seed, size_t n, size_t N,
                                                      std::function<double(double)> fun) {
    std::mt19937 engine(seed);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    std::vector<double> p(n);
    for (size_t i = 0; i < n; ++i) {
        double v = fun((double)i);
        p[i] = v < 0.001 ? v + 0.001 : v;
    }
    float fsum = 0.0f;                       // accumulated in float
    for (double v : p) fsum = (float)(fsum + v);
    std::vector<size_t> cnt(n);
    int isum = 0;                            // accumulated in int
    for (size_t i = 0; i < n; ++i) {
        p[i] = p[i] / fsum;
        cnt[i] = (size_t)std::ceil(p[i] * N);
        isum = (int)(isum + cnt[i]);
    }
    size_t surplus = (size_t)isum - N;
    while (surplus > 0) {
        size_t idx = (size_t)std::floor(uni(engine) * n);
        uni(engine);
        while (cnt[idx] == 1) {
            idx = (size_t)std::floor(uni(engine) * n);
            uni(engine);
        }
        --cnt[idx];
        --surplus;
    }
    auto prob = std::make_shared<std::vector<double>>(n);
    for (size_t i = 0; i < n; ++i) (*prob)[i] = (double)cnt[i] / N;
    return {prob, std::make_shared<std::vector<size_t>>(cnt), nullptr};
}

Distribution ANSTableGenerator::generate_distribution_from_buffer(size_t seed, size_t N,
                                                                  std::uint8_t* in, size_t size) {
    std::vector<std::uint64_t> freq(256, 0);
    for (size_t i = 0; i < size; ++i) ++freq[in[i]];
    std::vector<size_t> f;
    std::vector<SYMBOL_TYPE> sym;
    for (size_t s = 0; s < 256; ++s)
        if (freq[s]) { f.push_back(freq[s]); sym.push_back((SYMBOL_TYPE)s); }
    const size_t m = f.size();
    std::vector<double> pa(m);
    float fs = 0.0f;
    for (size_t v : f) fs = (float)(fs + (double)v);
    for (size_t i = 0; i < m; ++i) {
        double x = (double)f[i] / fs;
        pa[i] = x < 0.001 ? x + 0.001 : x;
    }
    fs = 0.0f;
    for (double v : pa) fs = (float)(fs + v);
    int isum = 0;
    for (size_t i = 0; i < m; ++i) {
        f[i] = (size_t)std::ceil(pa[i] / fs * N);
        isum = (int)(isum + f[i]);
    }
    std::mt19937 engine(seed);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    size_t surplus = (size_t)isum - N;
    while (surplus > 0) {
        size_t idx = (size_t)std::floor(uni(engine) * m);
        uni(engine);
        while (f[idx] == 1) {
            idx = (size_t)std::floor(uni(engine) * m);
            uni(engine);
        }
        --f[idx];
        --surplus;
    }
    auto prob = std::make_shared<std::vector<double>>(m);
    for (size_t i = 0; i < m; ++i) (*prob)[i] = (double)f[i] / N;
    return {prob, std::make_shared<std::vector<size_t>>(f),
            std::make_shared<std::vector<SYMBOL_TYPE>>(sym)};
}

std::shared_ptr<std::vector<SYMBOL_TYPE>> ANSTableGenerator::generate_test_data(
    std::shared_ptr<std::vector<size_t>> distr, size_t size, size_t num_states, size_t seed) {
    std::vector<SYMBOL_TYPE> spread(num_states, 0);
    size_t k = 0;
    for (size_t s = 0; s < distr->size(); ++s)
        for (size_t j = 0; j < distr->at(s) && k < num_states; ++j) spread[k++] = (SYMBOL_TYPE)s;
    auto out = std::make_shared<std::vector<SYMBOL_TYPE>>(size + 8, 0);
    std::mt19937 engine(seed);
    std::uniform_int_distribution<size_t> pick(0, num_states - 1);
    for (size_t i = 0; i < size; ++i) {
        (*out)[i] = spread[pick(engine)];
        pick(engine);
    }
    return out;
}

std::shared_ptr<std::vector<std::vector<Encoder_Table_Entry>>> ANSTableGenerator::generate_table(
    std::shared_ptr<std::vector<double>> P_s, std::shared_ptr<std::vector<size_t>> L_s,
    std::shared_ptr<std::vector<SYMBOL_TYPE>> symbols, size_t num_symbols, size_t num_states) {
    const size_t L = num_states;
    const size_t rows = symbols ? 256 : num_symbols;
    auto table = std::make_shared<std::vector<std::vector<Encoder_Table_Entry>>>(
        rows, std::vector<Encoder_Table_Entry>(L, Encoder_Table_Entry{0, 0, 0, 0}));
    std::priority_queue<SpreadKey, std::vector<SpreadKey>, std::greater<SpreadKey>> pq;
    std::vector<std::uint32_t> X(num_symbols);
    for (size_t i = 0; i < num_symbols; ++i) {
        X[i] = (std::uint32_t)L_s->at(i);
        pq.push({0.5 / P_s->at(i), i});
    }
    for (size_t st = L; st < 2 * L; ++st) {
        SpreadKey k = pq.top();
        pq.pop();
        const size_t i = k.idx;
        const SYMBOL_TYPE s = symbols ? symbols->at(i) : (SYMBOL_TYPE)i;
        const std::uint32_t xv = X[i]++;
        pq.push({k.p + 1.0 / P_s->at(i), i});
        // every x in [L, 2L) with x >> shift == xv encodes s into state st
        for (std::uint32_t shift = 0; ((std::uint64_t)xv << shift) < 2 * L; ++shift) {
            std::uint64_t lo = std::max<std::uint64_t>((std::uint64_t)xv << shift, L);
            std::uint64_t hi = std::min<std::uint64_t>(((std::uint64_t)xv + 1) << shift, 2 * L);
            for (std::uint64_t x = lo; x < hi; ++x)
                table->at(s)[x - L] = {(std::uint32_t)st, (std::uint32_t)(x - (((std::uint64_t)xv) << shift)),
                                       shift, s};
        }
    }
    return table;
}

std::shared_ptr<ANSEncoderTable> ANSTableGenerator::generate_encoder_table(
    std::shared_ptr<std::vector<std::vector<Encoder_Table_Entry>>> tab) {
    auto e = std::make_shared<ANSEncoderTable>();
    e->number_of_states = tab->at(0).size();
    e->max_number_of_symbols = 256;
    e->number_of_symbols = tab->size();
    e->table.resize(tab->size());
    for (size_t s = 0; s < tab->size(); ++s) {
        e->table[s].resize(e->number_of_states);
        for (size_t j = 0; j < e->number_of_states; ++j) {
            const Encoder_Table_Entry& t = tab->at(s)[j];
            e->table[s][j] = {t.x, t.rem, t.shift, t.symbol};
        }
    }
    return e;
}

std::shared_ptr<CUHDCodetable> ANSTableGenerator::get_decoder_table(
    std::shared_ptr<ANSEncoderTable> enc) {
    const size_t L = enc->number_of_states;
    auto t = std::make_shared<CUHDCodetable>(L);
    CUHDCodetableItem* d = t->get();
    for (size_t s = 0; s < enc->number_of_symbols; ++s) {
        for (size_t j = 0; j < L; ++j) {
            const auto& e = enc->table[s][j];
            if (e.next_state < L) continue;          // row of an absent symbol
            const size_t prev = e.next_state - L;
            const std::uint32_t len = e.code_length;
            CUHDCodetableItem it;
            it.next_state = (std::uint16_t)((L + j) >> len);
            it.symbol = e.symbol;
            it.min_num_bits = (d[prev].min_num_bits == 0 && len > 0) ? (std::uint8_t)len
                                                                      : d[prev].min_num_bits;
            d[prev] = it;
        }
    }
    return t;
}

size_t ANSTableGenerator::get_max_compressed_size(std::shared_ptr<ANSEncoderTable> enc,
                                                  size_t input_size) {
    std::uint32_t mx = 0;
    for (const auto& row : enc->table)
        for (const auto& e : row) mx = std::max(mx, e.code_length);
    return (mx * input_size) / 32 + 2;
}

std::shared_ptr<CUHDInputBuffer> ANSEncoder::encode(SYMBOL_TYPE* in, size_t size_in,
                                                    std::shared_ptr<ANSEncoderTable> enc) {
    const size_t L = enc->number_of_states;
    std::vector<UNIT_TYPE> units;
    units.reserve(ANSTableGenerator::get_max_compressed_size(enc, size_in));
    std::uint64_t acc = 0;
    std::uint32_t fill = 0, x = (std::uint32_t)L;
    for (size_t i = 0; i < size_in; ++i) {
        const auto& e = enc->table.at(in[i])[x - L];
        acc = (acc << e.code_length) | e.code_sequence;
        fill += e.code_length;
        x = e.next_state;
        if (fill >= 32) {
            fill -= 32;
            units.push_back((UNIT_TYPE)(acc >> fill));
            acc &= (1ull << fill) - 1;
        }
    }
    units.push_back(fill ? (UNIT_TYPE)(acc << (32 - fill)) : 0u);
    return std::make_shared<CUHDInputBuffer>(units.data(), units.size(), fill, x);
}
