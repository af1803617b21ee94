// demo_main.cc -- the reference demo's flow (src/main.cc:34-207) on the drop-in classes.
//
//   mans_demo <device> <size in MB> [--no-gpu] [--checksums] [--dump-sync DIR]
//
// For lambda = 0.1, 0.26, ..., 2.34: build the exponential distribution and test data,
// the tables and the encoded stream with ANSTableGenerator/ANSEncoder, decode on the GPU
// with CUHDGPUDecoder (S = 4, 128 threads, as the demo passes them), copy back, reverse
// and compare with the input.  --checksums prints FNV-1a hashes of the counts, decoder
// table, data and units per lambda so the host codec can be compared with the
// reference's; --no-gpu skips the decode; --dump-sync DIR writes, for the first two
// lambdas, the stream and the decoder memory's reference sync state after the decode
// (get_sync_info, get_output_sizes, retrieve_sync_data + get_sequence_synced_host) to
// DIR/sync_<k>.bin, for a test to check against the oracle.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "multians.h"

#define NUM_SYMBOLS 256
#define NUM_STATES 1024
#define SEED 5
#define SUBSEQUENCE_SIZE 4
#define THREADS_PER_BLOCK 128

static std::uint64_t fnv(const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    std::uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < n; ++i) { h ^= b[i]; h *= 1099511628211ull; }
    return h;
}

template <typename T>
static void put(std::FILE* f, const T* p, size_t n) { std::fwrite(p, sizeof(T), n, f); }

static int run(long input_size, bool gpu, bool checksums, const char* dump) {
    int failures = 0;
    int k = 0;
    std::cout << "lambda | compressed size (bytes) | time [gpu decode] (us)" << std::endl;
    for (float lambda = 0.1f; lambda < 2.5f; lambda += 0.16) {
        auto dist = ANSTableGenerator::generate_distribution(
            SEED, NUM_SYMBOLS, NUM_STATES, [&](double x) { return lambda * exp(-lambda * x); });
        auto data = ANSTableGenerator::generate_test_data(dist.dist, input_size, NUM_STATES, SEED);
        auto table = ANSTableGenerator::generate_table(dist.prob, dist.dist, nullptr, NUM_SYMBOLS,
                                                       NUM_STATES);
        auto encoder_table = ANSTableGenerator::generate_encoder_table(table);
        auto decoder_table = ANSTableGenerator::get_decoder_table(encoder_table);
        auto input_buffer = ANSEncoder::encode(data->data(), input_size, encoder_table);
        auto output_buffer = std::make_shared<CUHDOutputBuffer>(input_size);
        if (checksums) {
            std::vector<std::uint32_t> counts(dist.dist->begin(), dist.dist->end());
            std::cout << "CHK " << std::setprecision(9) << lambda << " counts="
                      << fnv(counts.data(), counts.size() * 4)
                      << " dec=" << fnv(decoder_table->get(), NUM_STATES * 4)
                      << " data=" << fnv(data->data(), input_size)
                      << " units=" << fnv(input_buffer->get_compressed_data(),
                                          input_buffer->get_compressed_size() * 4)
                      << " nunits=" << input_buffer->get_compressed_size()
                      << " first_bit=" << input_buffer->get_first_bit()
                      << " first_state=" << input_buffer->get_first_state() << std::endl;
        }
        size_t us = 0;
        if (gpu) {
            auto gpu_in = std::make_shared<cuhd::CUHDGPUInputBuffer>(input_buffer);
            auto gpu_table = std::make_shared<cuhd::CUHDGPUCodetable>(decoder_table);
            auto gpu_out = std::make_shared<cuhd::CUHDGPUOutputBuffer>(output_buffer);
            auto mem = std::make_shared<cuhd::CUHDGPUDecoderMemory>(
                input_buffer->get_compressed_size(), SUBSEQUENCE_SIZE, THREADS_PER_BLOCK);
            gpu_table->allocate();
            gpu_in->allocate();
            gpu_out->allocate();
            mem->allocate();
            gpu_table->cpy_host_to_device();
            gpu_in->cpy_host_to_device();
            std::vector<std::pair<std::string, size_t>> timings;
            TIMER_START(timings, "gpu")
            cuhd::CUHDGPUDecoder::decode(gpu_in, input_buffer->get_compressed_size(), gpu_out,
                                         output_buffer->get_uncompressed_size(), gpu_table, mem,
                                         input_buffer->get_first_state(),
                                         input_buffer->get_first_bit(),
                                         decoder_table->get_num_entries(), 11, SUBSEQUENCE_SIZE,
                                         THREADS_PER_BLOCK);
            TIMER_STOP
            us = timings.at(0).second;
            gpu_out->cpy_device_to_host();
            output_buffer->reverse();
            if (!cuhd::CUHDUtil::equals(data->data(), output_buffer->get_decompressed_data().get(),
                                        input_size)) {
                std::cout << "mismatch" << std::endl;
                ++failures;
            }
            if (cuhd::CUHDGPUDecoder::decode_status(mem) != 0) {
                std::cout << "status error" << std::endl;
                ++failures;
            }
            if (dump && k < 2) {
                const size_t nu = input_buffer->get_compressed_size();
                const size_t nsub = SDIV(nu, SUBSEQUENCE_SIZE);
                const size_t nblk = SDIV(nsub, THREADS_PER_BLOCK);
                std::vector<std::uint32_t> sync(4 * nsub), sizes(nsub);
                MANS_CHECK(mans_memcpy_d2h(sync.data(), mem->get_sync_info(), 16 * nsub, nullptr));
                MANS_CHECK(mans_memcpy_d2h(sizes.data(), mem->get_output_sizes(), 4 * nsub, nullptr));
                mem->retrieve_sync_data();
                const std::string path = std::string(dump) + "/sync_" + std::to_string(k) + ".bin";
                std::FILE* f = std::fopen(path.c_str(), "wb");
                if (!f) { std::cout << "error: cannot write " << path << std::endl; return 1; }
                const std::uint64_t hdr[9] = {nu, input_buffer->get_first_bit(),
                                              input_buffer->get_first_state(), NUM_STATES,
                                              (std::uint64_t)input_size, SUBSEQUENCE_SIZE,
                                              THREADS_PER_BLOCK, nsub, nblk};
                put(f, hdr, 9);
                put(f, input_buffer->get_compressed_data(), nu);
                put(f, reinterpret_cast<const std::uint32_t*>(decoder_table->get()), NUM_STATES);
                put(f, sync.data(), 4 * nsub);
                put(f, sizes.data(), nsub);
                put(f, mem->get_sequence_synced_host(), nblk);
                std::fclose(f);
            }
        }
        std::cout << std::left << std::setw(7) << lambda << std::setw(12)
                  << input_buffer->get_compressed_size() * sizeof(UNIT_TYPE) << us << std::endl;
        ++k;
    }
    return failures;
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::cout << "USAGE: " << argv[0] << " <device> <size in MB> [--no-gpu] [--checksums]"
                  << std::endl;
        return 1;
    }
    int dev = std::atoi(argv[1]);
    long mb = std::atol(argv[2]);
    bool gpu = true, checksums = false;
    const char* dump = nullptr;
    for (int i = 3; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--no-gpu")) gpu = false;
        if (!std::strcmp(argv[i], "--checksums")) checksums = true;
        if (!std::strcmp(argv[i], "--dump-sync") && i + 1 < argc) dump = argv[++i];
    }
    long size = mb > 0 ? mb * 1024 * 1024 : 100000;
    if (gpu) MANS_CHECK(mans_set_device(dev));
    int f = run(size, gpu, checksums, dump);
    return f ? 2 : 0;
}
