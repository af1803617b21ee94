// mans_interleaved.hip -- decoder of two-state interleaved streams (C ABI).
//
// The reference decodes one tANS state per stream and names multiple states and
// multiple tables as future work (README.md:11).  This format is the library's own: two
// states (slots) share one bit stream, input symbol i is coded by slot i % 2 with that
// slot's table, and the encoder (mans_encode_interleaved) writes records that cut the
// stream into intervals of record_units units, anchored at the stream's end (the format
// and the records are stated in include/mans_c_api.h; the oracle restates both,
// mo_encode_interleaved / mo_decode_interleaved).
//
// Decoding needs no synchronisation: lane q of the grid decodes interval q, from the
// record at its start (both states, the bit position, the output offset) to the record
// at its end, and checks that it lands exactly there.  Within a lane the two slots form
// two independent lookup chains: a pair of steps reads one 32-bit window, takes slot a's
// code at the window's start and slot b's right after it (both bit counts are in the
// table words already held), and issues both next table lookups together, so each LDS
// lookup latency is paid once per two symbols.
//
// Units reach the lane through a ring of two 16-unit blocks in LDS (plus a guard word
// repeating slot 0, so a window never wraps); every 16 symbols the lane checks whether
// its bit position entered a new block, and if so writes the block after it from
// registers (loaded one block earlier) into the slots just left and loads the next one.  Symbols are packed four to a dword and
// stored 16 bytes at a time, as in the single-stream write pass.
#include <hip/hip_runtime.h>

#include "mans_il_lane.h"
#include "mans_internal.h"

namespace {

using namespace mans;

constexpr uint32_t kIWG = 256;
constexpr size_t kLdsBytes = 160u * 1024u;   // LDS of one CU (gfx950)

typedef const uint32_t __attribute__((address_space(3)))* lds_cptr;
typedef uint32_t __attribute__((address_space(3)))* lds_ptr;
typedef uint8_t __attribute__((address_space(1)))* gptr_u8;
typedef uint32_t __attribute__((address_space(1)))* gptr_w;
typedef uint32_t v4u __attribute__((ext_vector_type(4)));
typedef v4u __attribute__((address_space(1)))* gptr_q;

__device__ __forceinline__ uint32_t lds_addr(const void* p) {
    return (uint32_t)(size_t)(const __attribute__((address_space(3))) void*)p;
}

// LDS by byte address; output through global (not flat) stores, which would also count
// on the LDS counter and force LDS waits at every step
struct DevMem {
    __device__ __forceinline__ uint32_t ld(uint32_t a) const { return *(lds_cptr)(size_t)a; }
    __device__ __forceinline__ void st(uint32_t a, uint32_t v) const { *(lds_ptr)(size_t)a = v; }
    __device__ __forceinline__ void out8(uint8_t* p, uint32_t v) const { *(gptr_u8)p = (uint8_t)v; }
    __device__ __forceinline__ void out32(uint8_t* p, uint32_t v) const { *(gptr_w)p = v; }
    __device__ __forceinline__ void out128(uint8_t* p, uint32_t a, uint32_t b, uint32_t c,
                                           uint32_t d) const {
        v4u v;
        v.x = a; v.y = b; v.z = c; v.w = d;
        *(gptr_q)p = v;
    }
};

__global__ void k_il_init(uint32_t* status) {
    if (threadIdx.x == 0) status[0] = 0u;
}

__global__ __launch_bounds__(kIWG) void k_il_decode(IlArgs a) {
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t s_bad;
    if (threadIdx.x == 0) s_bad = 0u;
    // tables: slot 0 at dword 0, slot 1 after it (or the same words); each word carries
    // its own table's base, so a state stays in its slot's table
    const uint32_t tw = (a.L + 3u) & ~3u;
    const uint32_t t0 = lds_addr(lds) >> 2;
    const uint32_t t1 = a.two_tables ? t0 + tw : t0;
    int bad = 0, cyc = 0;
    for (uint32_t k = 0; k < 1u + a.two_tables; ++k) {
        const uint32_t tb = k ? t1 : t0;
        for (uint32_t i = threadIdx.x; i < a.L; i += blockDim.x) {
            const uint32_t e = load_unit(a.tab[k], a.L, i);
            lds[tb - t0 + i] = derive_entry(e, a.L, i, &bad, &cyc) + (tb << 16);
        }
    }
    if (bad) atomicOr(&s_bad, 1u);
    __syncthreads();
    if (s_bad) {
        if (threadIdx.x == 0) atomicOr(&a.status[0], 1u);
        return;                              // a rejected table decodes nothing
    }
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q > a.M) return;
    DevMem m;
    const uint32_t ring = (1u + a.two_tables) * tw + kIlRingStride * threadIdx.x;
    const uint32_t f = il_lane(a, q, m, t0, t1, lds_addr(lds + ring));
    if (f) atomicOr(&a.status[0], f);
}

}  // namespace

extern "C" int mans_decode_interleaved(const mans_il_args* d) {
    if (!d) return set_error(MANS_EINVAL, "args is NULL");
    const uint32_t L = d->L;
    if (L < 2u || L > 16384u || (L & (L - 1u)))
        return set_error(MANS_EINVAL, "L=%u is not a power of two in [2, 16384]", L);
    if (!d->tables[0] || !d->tables[1]) return set_error(MANS_EINVAL, "table is NULL");
    if (!d->workspace) return set_error(MANS_EINVAL, "workspace is NULL");
    if (d->record_units < MANS_IL_MIN_RECORD_UNITS || d->record_units > MANS_IL_MAX_RECORD_UNITS)
        return set_error(MANS_EINVAL, "record_units=%u outside [%u, %u]", d->record_units,
                         MANS_IL_MIN_RECORD_UNITS, MANS_IL_MAX_RECORD_UNITS);
    if (d->n_units == 0 || d->n_units >= 0xFFFFFF00ull)
        return set_error(MANS_EINVAL, "n_units=%llu out of range", (unsigned long long)d->n_units);
    if (!d->units || !d->records || (d->n_symbols && !d->out))
        return set_error(MANS_EINVAL, "units/records/out is NULL");
    if (((uintptr_t)d->units & 15u) != 0u)
        return set_error(MANS_EINVAL, "units must be 16-byte aligned (device allocations are)");
    if (d->first_bit > 31u) return set_error(MANS_EINVAL, "first_bit=%u > 31", d->first_bit);
    for (int k = 0; k < 2; ++k)
        if (d->first_states[k] < L || d->first_states[k] >= 2u * L)
            return set_error(MANS_ECORRUPT, "first state %u outside [L, 2L)", d->first_states[k]);
    uint64_t M = 0;
    int rc = mans_il_record_count(d->n_units, d->record_units, &M);
    if (rc) return rc;
    IlArgs a;
    a.units = d->units;
    a.tab[0] = d->tables[0];
    a.tab[1] = d->tables[1];
    a.two_tables = d->tables[0] != d->tables[1];
    a.rec = d->records;
    a.out = d->out;
    a.status = (uint32_t*)d->workspace;
    a.n = d->n_symbols;
    a.n_units = (uint32_t)d->n_units;
    a.L = L;
    a.first_state[0] = d->first_states[0];
    a.first_state[1] = d->first_states[1];
    a.first_bit = d->first_bit;
    a.R = d->record_units;
    a.M = (uint32_t)M;
    // lanes per workgroup: 256 unless the tables and rings would not fit one CU's LDS
    // (two 16384-state tables)
    const size_t tab_bytes = 4u * (1u + a.two_tables) * ((L + 3u) & ~3u);
    uint32_t lanes = kIWG;
    while (lanes > 64u && tab_bytes + 4u * kIlRingStride * lanes > kLdsBytes - 256u) lanes -= 64u;
    const size_t lds_bytes = tab_bytes + 4u * kIlRingStride * lanes;
    hipStream_t st = (hipStream_t)d->stream;
    hipLaunchKernelGGL(k_il_init, dim3(1), dim3(64), 0, st, a.status);
    const uint32_t nwg = (uint32_t)((M + lanes) / lanes);
    hipLaunchKernelGGL(k_il_decode, dim3(nwg), dim3(lanes), lds_bytes, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(MANS_EHIP, "interleaved decode launch: %s", hipGetErrorString(e));
    if (d->flags & MANS_F_SYNC_ERRORS) {
        int status = 0;
        rc = mans_decode_status(d->workspace, d->stream, &status);
        if (rc) return rc;
        if (status) return set_error(status, "device reported status %d", status);
    }
    return MANS_OK;
}
