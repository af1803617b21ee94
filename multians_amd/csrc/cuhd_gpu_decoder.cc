// cuhd_gpu_decoder.cc -- cuhd::CUHDGPUDecoder::decode over the C ABI (mans_decode).
// Replaces src/multians_gpu_decoder.cu:575-687 of the reference.
#include "cuhd_gpu_decoder.h"

#include <cstring>

#include "mans_c_api.h"

namespace cuhd {

void CUHDGPUDecoder::decode(std::shared_ptr<CUHDGPUInputBuffer> input, size_t input_size,
                            std::shared_ptr<CUHDGPUOutputBuffer> output, size_t output_size,
                            std::shared_ptr<CUHDGPUCodetable> table,
                            std::shared_ptr<CUHDGPUDecoderMemory> aux, STATE_TYPE initial_state,
                            std::uint32_t initial_bit, std::uint32_t number_of_states,
                            size_t max_codeword_length, size_t preferred_subsequence_size,
                            size_t threads_per_block) {
    (void)max_codeword_length;
    // the subsequence size and block size only shape the reference's sync state arrays
    // (CUHDGPUDecoderMemory's accessors), which must have been sized for them, as in the
    // reference; this decoder picks its own sync-chunk lengths
    if (preferred_subsequence_size != aux->subsequence_size() ||
        threads_per_block != aux->num_threads()) {
        std::printf("MANS error: decoder memory built for subsequence size %zu and %zu threads, "
                    "decode called with %zu and %zu\n", aux->subsequence_size(),
                    aux->num_threads(), preferred_subsequence_size, threads_per_block);
        std::exit(1);
    }
    if (input_size > aux->num_units()) {
        std::printf("MANS error: decoder memory sized for %zu units, stream has %zu\n",
                    aux->num_units(), input_size);
        std::exit(1);
    }
    mans_decode_args a;
    std::memset(&a, 0, sizeof(a));
    a.units = input->get();
    a.n_units = input_size;
    a.out = output->get();
    a.n_out = output_size;
    a.table = reinterpret_cast<const std::uint32_t*>(table->get());
    a.L = number_of_states;
    a.first_state = initial_state;
    a.first_bit = initial_bit;
    a.workspace = aux->workspace();
    a.workspace_bytes = aux->workspace_bytes();
    a.stream = nullptr;
    MANS_CHECK(mans_decode(&a));
    aux->note_decode(&a);
}

int CUHDGPUDecoder::decode_status(std::shared_ptr<CUHDGPUDecoderMemory> aux) {
    int st = 0;
    MANS_CHECK(mans_decode_status(aux->workspace(), nullptr, &st));
    return st;
}

}  // namespace cuhd
