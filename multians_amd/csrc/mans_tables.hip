// mans_tables.hip -- decoder tables built on the device, from the symbols already there:
// histogram -> counts normalised to L -> spread -> decoder table in the reference's
// CUHDCodetableItem layout.  The bytes are those of the host codec (mans_histogram,
// mans_normalize_counts, mans_build_tables in mans_codec.cc), which restates
// ANSTableGenerator (src/ans_table_generator.cc:61-95 counting, :164-288 spread and
// decoder table); the reference itself only builds tables on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "mans_c_api.h"
#include "mans_internal.h"

namespace {

// Per job in the workspace: hist[256] (u64), counts[256] (u32), status (u32).
struct JobWork {
    uint64_t hist[256];
    uint32_t counts[256];
    uint32_t status;
    uint32_t pad[63];
};
static_assert(sizeof(JobWork) % 256 == 0, "job work alignment");

constexpr uint32_t kHistWG = 256;
constexpr uint64_t kHistSlice = 1ull << 22;     // symbols per histogram workgroup (at most)

// Table-build status codes (JobWork::status).
constexpr uint32_t kTabEmpty = 1;        // no symbols
constexpr uint32_t kTabTooMany = 2;      // more distinct symbols than states
constexpr uint32_t kTabFailed = 3;       // normalisation found no symbol to adjust

__global__ void k_tab_clear(JobWork* jw, uint32_t n_jobs) {
    const uint32_t j = blockIdx.x;
    if (j >= n_jobs) return;
    jw[j].hist[threadIdx.x] = 0u;
    jw[j].counts[threadIdx.x] = 0u;
    if (threadIdx.x == 0) jw[j].status = 0u;
}

// Histogram of job blockIdx.y's symbols, slice blockIdx.x (kHistSlice symbols): sixteen
// LDS histograms, one per lane residue mod 16, so that on skewed data (one symbol holding
// most of the mass) at most four lanes of a wave add to the same counter at once; folded
// into the job's u64 histogram.
constexpr uint32_t kHistCopies = 16;
__global__ __launch_bounds__(kHistWG) void k_tab_hist(const mans_table_job* jobs, JobWork* jw) {
    __shared__ uint32_t h[kHistCopies][256];
    const mans_table_job jb = jobs[blockIdx.y];
    const uint64_t b = (uint64_t)blockIdx.x * kHistSlice;
    if (b >= jb.n) return;                       // uniform
    const uint64_t e = min(jb.n, b + kHistSlice);
    for (uint32_t i = threadIdx.x; i < kHistCopies * 256u; i += kHistWG) (&h[0][0])[i] = 0u;
    __syncthreads();
    uint32_t* mine = h[threadIdx.x % kHistCopies];
    const uint8_t* s = (const uint8_t*)jb.symbols;
    // 16-byte blocks of the middle (16-byte aligned addresses), single bytes at the edges
    const uint64_t mis = (uint64_t)((uintptr_t)(s + b) & 15u);
    const uint64_t a0 = min(e, b + ((16u - mis) & 15u));
    const uint64_t a1 = max(a0, a0 + ((e - a0) & ~15ull));
    for (uint64_t i = b + threadIdx.x; i < a0; i += kHistWG) atomicAdd(&mine[s[i]], 1u);
    const uint4* q = (const uint4*)(s + a0);
    const uint64_t nq = (a1 - a0) >> 4;
    for (uint64_t i = threadIdx.x; i < nq; i += kHistWG) {
        const uint4 v = q[i];
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int t = 0; t < 4; ++t) atomicAdd(&mine[(w[k] >> (8 * t)) & 255u], 1u);
    }
    for (uint64_t i = a1 + threadIdx.x; i < e; i += kHistWG) atomicAdd(&mine[s[i]], 1u);
    __syncthreads();
    uint32_t c = 0u;
#pragma unroll
    for (uint32_t k = 0; k < kHistCopies; ++k) c += h[k][threadIdx.x];
    if (c) atomicAdd((unsigned long long*)&jw[blockIdx.y].hist[threadIdx.x], (unsigned long long)c);
}

// (key, symbol) order of the spread: smaller key first, lower symbol on ties.
__device__ __forceinline__ bool before(double pa, uint32_t sa, double pb, uint32_t sb) {
    return pa < pb || (pa == pb && sa < sb);
}

// One wave per job: normalise the histogram to L, then spread and write the decoder
// table.  Lane l owns symbols l, l+64, l+128, l+192.
__global__ __launch_bounds__(64) void k_tab_build(const mans_table_job* jobs, JobWork* jw) {
    const mans_table_job jb = jobs[blockIdx.x];
    JobWork& w = jw[blockIdx.x];
    const uint32_t lane = threadIdx.x, L = jb.L;
    uint64_t hs[4];
    uint64_t total = 0;
    uint32_t present = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        hs[k] = w.hist[lane + 64u * k];
        total += hs[k];
        present += hs[k] != 0u;
    }
    for (int d = 32; d >= 1; d >>= 1) {
        total += __shfl_xor(total, d, 64);
        present += __shfl_xor(present, d, 64);
    }
    if (!total || present > L) {
        if (lane == 0) w.status = total ? kTabTooMany : kTabEmpty;
        return;
    }
    // counts: floor(h * L / total), at least 1 for a present symbol (h * L < 2^64: the
    // host allows n < 2^50 symbols)
    uint32_t c[4];
    int64_t sum = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint64_t v = hs[k] ? hs[k] * L / total : 0u;
        c[k] = hs[k] ? (v ? (uint32_t)v : 1u) : 0u;
        sum += c[k];
    }
    for (int d = 32; d >= 1; d >>= 1) sum += __shfl_xor(sum, d, 64);
    // adjust the largest count (lowest symbol on ties) until the counts sum to L
    while (sum != (int64_t)L) {
        uint32_t bc = 0u, bs = 256u;              // best count, its symbol (256 = none)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t s = lane + 64u * k;
            const bool ok = c[k] && !(sum > (int64_t)L && c[k] <= 1u);
            if (ok && (bs == 256u || c[k] > bc)) { bc = c[k]; bs = s; }
        }
        for (int d = 32; d >= 1; d >>= 1) {
            const uint32_t oc = __shfl_xor(bc, d, 64), os = __shfl_xor(bs, d, 64);
            if (os != 256u && (bs == 256u || oc > bc || (oc == bc && os < bs))) { bc = oc; bs = os; }
        }
        if (bs == 256u) {
            if (lane == 0) w.status = kTabFailed;
            return;
        }
        if ((bs & 63u) == lane) {
            const int k = (int)(bs >> 6);
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
                if (kk == k) c[kk] = sum < (int64_t)L ? c[kk] + (uint32_t)((int64_t)L - sum) : c[kk] - 1u;
        }
        sum = sum < (int64_t)L ? (int64_t)L : sum - 1;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        w.counts[lane + 64u * k] = c[k];
        if (jb.counts) ((uint32_t*)jb.counts)[lane + 64u * k] = c[k];
    }
    // spread: state L+i goes to the symbol with the smallest key; a symbol's key starts
    // at 0.5/P and grows by 1/P after each state it takes, P = count/L (double)
    double p[4], ip[4];
    uint32_t x[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        x[k] = c[k];
        if (c[k]) {
            const double P = (double)c[k] / (double)L;
            ip[k] = 1.0 / P;
            p[k] = 0.5 / P;
        } else {
            ip[k] = 0.0;
            p[k] = __builtin_huge_val();
        }
    }
    uint32_t* dec = (uint32_t*)jb.table;
    for (uint32_t i = 0; i < L; ++i) {
        double bp = p[0];
        uint32_t bs = lane;
#pragma unroll
        for (int k = 1; k < 4; ++k)
            if (before(p[k], lane + 64u * k, bp, bs)) { bp = p[k]; bs = lane + 64u * k; }
        for (int d = 32; d >= 1; d >>= 1) {
            const double op = __shfl_xor(bp, d, 64);
            const uint32_t os = __shfl_xor(bs, d, 64);
            if (before(op, os, bp, bs)) { bp = op; bs = os; }
        }
        if ((bs & 63u) == lane) {
            const int k = (int)(bs >> 6);
            uint32_t xv = 0u;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
                if (kk == k) { xv = x[kk]++; p[kk] = p[kk] + ip[kk]; }
            // nbits: the first nonzero shift k whose source states [xv<<k, (xv+1)<<k)
            // meet [L, 2L), as the reference decoder table records it
            uint32_t first_len = 0u;
            for (uint32_t sh = 1; ((uint64_t)xv << sh) < 2ull * L; ++sh) {
                const uint64_t lo = max((uint64_t)xv << sh, (uint64_t)L);
                const uint64_t hi = min(((uint64_t)xv << sh) + (1ull << sh), 2ull * L);
                if (lo < hi) { first_len = sh; break; }
            }
            dec[i] = (xv & 0xFFFFu) | (bs << 16) | (first_len << 24);
        }
    }
}

}  // namespace

static uint64_t jobs_offset(size_t n_jobs) { return (uint64_t)n_jobs * sizeof(JobWork); }

extern "C" int mans_table_workspace_bytes(size_t n_jobs, uint64_t* bytes) {
    if (!bytes) return set_error(MANS_EINVAL, "bytes is NULL");
    n_jobs = std::max<size_t>(n_jobs, 1);
    *bytes = jobs_offset(n_jobs) + (((uint64_t)n_jobs * sizeof(mans_table_job) + 255u) & ~255ull);
    return MANS_OK;
}

extern "C" int mans_build_tables_device(const mans_table_job* jobs, size_t n_jobs, void* workspace,
                                        uint64_t workspace_bytes, void* stream) {
    if (!jobs || !n_jobs) return set_error(MANS_EINVAL, "no jobs");
    if (n_jobs > 65535u) return set_error(MANS_EINVAL, "at most 65535 jobs per call");
    uint64_t need = 0;
    (void)mans_table_workspace_bytes(n_jobs, &need);
    if (!workspace || workspace_bytes < need)
        return set_error(MANS_ENOMEM, "table workspace of %llu bytes < %llu needed",
                         (unsigned long long)workspace_bytes, (unsigned long long)need);
    uint64_t max_n = 0;
    for (size_t j = 0; j < n_jobs; ++j) {
        const mans_table_job& jb = jobs[j];
        if (!jb.symbols || !jb.table) return set_error(MANS_EINVAL, "job %zu: NULL buffer", j);
        if (jb.L < 2u || jb.L > 16384u || (jb.L & (jb.L - 1u)))
            return set_error(MANS_EINVAL, "job %zu: L=%u is not a power of two in [2, 16384]", j, jb.L);
        if (jb.n == 0 || jb.n >= (1ull << 50))
            return set_error(MANS_EINVAL, "job %zu: n=%llu outside [1, 2^50)", j,
                             (unsigned long long)jb.n);
        max_n = std::max(max_n, jb.n);
    }
    const uint64_t slices = (max_n + kHistSlice - 1) / kHistSlice;
    if (slices > 0x7FFFFFFFull) return set_error(MANS_EINVAL, "input too large");
    hipStream_t st = (hipStream_t)stream;
    JobWork* jw = (JobWork*)workspace;
    mans_table_job* djobs = (mans_table_job*)((char*)workspace + jobs_offset(n_jobs));
    // the job array goes into the workspace (the host waits for the copy, so the caller's
    // array may go away on return)
    hipError_t e = hipMemcpyAsync(djobs, jobs, n_jobs * sizeof(mans_table_job), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_tab_clear, dim3((uint32_t)n_jobs), dim3(256), 0, st, jw, (uint32_t)n_jobs);
        hipLaunchKernelGGL(k_tab_hist, dim3((uint32_t)slices, (uint32_t)n_jobs), dim3(kHistWG), 0, st,
                           djobs, jw);
        hipLaunchKernelGGL(k_tab_build, dim3((uint32_t)n_jobs), dim3(64), 0, st, djobs, jw);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return set_error(MANS_EHIP, "table build: %s", hipGetErrorString(e));
    return MANS_OK;
}

extern "C" int mans_table_status(const void* workspace, size_t n_jobs, void* stream, uint32_t* status) {
    if (!workspace || !status) return set_error(MANS_EINVAL, "NULL argument");
    hipStream_t st = (hipStream_t)stream;
    const JobWork* jw = (const JobWork*)workspace;
    for (size_t j = 0; j < n_jobs; ++j) {
        hipError_t e = hipMemcpyAsync(&status[j], &jw[j].status, 4, hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) return set_error(MANS_EHIP, "status: %s", hipGetErrorString(e));
    }
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return set_error(MANS_EHIP, "status: %s", hipGetErrorString(e));
    return MANS_OK;
}

// ------------------------------------------------------------ symbols on the device
// mans_gen_symbols_device: the same symbols as mans_gen_symbols (mans_codec.cc), element
// i from (seed, i) alone: u = splitmix64(seed + (i + 1) * golden) >> 11, uniform on
// [0, 2^53), s = #{k in 1..255 : u < thr[k]} with thr[k] = floor(2^53 exp(-lambda k))
// computed on the host (the same libm call as the host generator, so equal bytes).  The
// generator of the demo data (src/main.cc:55-73 draws from std::exponential_distribution);
// used to make large test and benchmark streams without a host round trip.
namespace {

struct GenThr {
    uint64_t thr[256];
};

__device__ __forceinline__ uint64_t dev_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// largest k in [lo, hi] with u < thr[k] (u < thr[lo] holds)
__device__ __forceinline__ uint32_t dev_sym_of(const uint64_t* thr, uint64_t u, uint32_t lo,
                                               uint32_t hi) {
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1u) >> 1;
        if (u < thr[mid]) lo = mid; else hi = mid - 1u;
    }
    return lo;
}

constexpr uint32_t kGenWG = 256;

// Grid-stride over 16-symbol blocks; per workgroup the thresholds and, per 2^41-wide
// bucket of u (its top 12 bits), the symbol range [s(bucket end), s(bucket start)]
// sit in LDS, so most symbols take one lookup.
__global__ __launch_bounds__(kGenWG) void k_gen_symbols(GenThr t, uint64_t seed, uint8_t* out,
                                                        uint64_t n) {
    __shared__ uint64_t thr[256];
    __shared__ uint8_t hi[4096], lo[4096];
    thr[threadIdx.x] = t.thr[threadIdx.x];
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < 4096u; k += kGenWG) {
        hi[k] = (uint8_t)dev_sym_of(thr, (uint64_t)k << 41, 0u, 255u);
        lo[k] = (uint8_t)dev_sym_of(thr, (((uint64_t)k + 1u) << 41) - 1u, 0u, 255u);
    }
    __syncthreads();
    const uint64_t nb = (n + 15u) / 16u;
    for (uint64_t b = (uint64_t)blockIdx.x * kGenWG + threadIdx.x; b < nb;
         b += (uint64_t)gridDim.x * kGenWG) {
        uint32_t w[4];
        for (uint32_t q = 0; q < 4u; ++q) {
            uint32_t v = 0u;
            for (uint32_t r = 0; r < 4u; ++r) {
                const uint64_t i = b * 16u + q * 4u + r;
                const uint64_t u = dev_mix64(seed + (i + 1u) * 0x9E3779B97F4A7C15ull) >> 11;
                const uint32_t k = (uint32_t)(u >> 41);
                const uint32_t h = hi[k], l = lo[k];
                const uint32_t s = h == l ? h : dev_sym_of(thr, u, l, h);
                v |= s << (8u * r);
            }
            w[q] = v;
        }
        const uint64_t i0 = b * 16u;
        uint8_t* dst = out + i0;
        if (i0 + 16u <= n && (((uintptr_t)dst) & 15u) == 0u) {
            *reinterpret_cast<uint4*>(dst) = make_uint4(w[0], w[1], w[2], w[3]);
        } else {
            for (uint32_t j = 0; j < 16u && i0 + j < n; ++j)
                dst[j] = (uint8_t)(w[j >> 2] >> (8u * (j & 3u)));
        }
    }
}

}  // namespace

extern "C" int mans_gen_symbols_device(uint64_t seed, double lambda, uint8_t* out, uint64_t n,
                                       void* stream) {
    if (!out && n) return set_error(MANS_EINVAL, "out is NULL");
    if (!(lambda > 0.0)) return set_error(MANS_EINVAL, "lambda must be > 0");
    if (!n) return MANS_OK;
    GenThr t;
    for (int k = 0; k < 256; ++k)
        t.thr[k] = (uint64_t)std::floor(std::ldexp(std::exp(-lambda * k), 53));
    const uint64_t nb = (n + 15u) / 16u;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((nb + kGenWG - 1u) / kGenWG, 4096u);
    hipLaunchKernelGGL(k_gen_symbols, dim3(grid), dim3(kGenWG), 0, (hipStream_t)stream, t, seed,
                       out, n);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(MANS_EHIP, "gen launch: %s", hipGetErrorString(e));
    return MANS_OK;
}
