// mans_kernels.hip -- gfx950 kernels of the tANS subsequence decoder and the
// decode entry points of the C ABI (mans_decode, mans_decode_batch).
//
// Replaces cuhd::CUHDGPUDecoder::decode and its kernels
// (src/multians_gpu_decoder.cu:195-687 of the reference):
//   phase1_decode_subseq           -> k_spec   (speculative decode, checkpoints)
//   phase2_synchronise_blocks+loop -> k_walk + k_fixup (merge detection on device,
//                                     no host round trip)
//   phase3 copies + thrust scan    -> k_scan   (per-stream exclusive scan, u64)
//   phase4_decode_write_output     -> k_write  (count-driven decode + store)
// The per-lane logic lives in mans_core.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "mans_core.h"
#include "mans_internal.h"

using namespace mans;

namespace {

constexpr int kWG = 256;           // lanes per workgroup (4 waves)
constexpr int kScanWG = 1024;

struct LdsTab {
    const uint32_t* p;
    __device__ __forceinline__ uint32_t operator[](uint32_t i) const { return p[i]; }
};

struct KArgs {
    StreamPlan one;                 // used when plans == nullptr
    const StreamPlan* plans;        // batch
    uint32_t n_streams;
    uint32_t n_tasks;
    Work w;
};

__device__ __forceinline__ const StreamPlan* plan_of(const KArgs& a, uint32_t s) {
    return a.plans ? &a.plans[s] : &a.one;
}

// Stream owning task t (tasks of each stream are contiguous, ascending in s).
template <bool kItems>
__device__ uint32_t stream_of_task(const KArgs& a, uint32_t t) {
    if (!a.plans) return 0u;
    uint32_t lo = 0, hi = a.n_streams - 1u;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1u) >> 1;
        uint32_t b = kItems ? a.plans[mid].task_item : a.plans[mid].task_chunk;
        if (b <= t) lo = mid; else hi = mid - 1u;
    }
    return lo;
}

__device__ uint32_t stream_of_chunk(const KArgs& a, uint32_t gc) {
    if (!a.plans) return 0u;
    uint32_t lo = 0, hi = a.n_streams - 1u;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1u) >> 1;
        if (a.plans[mid].chunk_base <= gc) lo = mid; else hi = mid - 1u;
    }
    return lo;
}

// Stage the derived table of stream sp into LDS; flags invalid entries.
__device__ void stage_table(const StreamPlan& sp, uint32_t* lds, uint32_t* status) {
    int bad = 0;
    for (uint32_t i = threadIdx.x; i < sp.L; i += blockDim.x) {
        uint32_t e = sp.table[i];
        lds[i] = derive_entry(e, sp.L, i, &bad);
    }
    if (bad) atomicOr(&status[0], 1u);
    __syncthreads();
}

__global__ void k_init(Work w, uint32_t n_status) {
    for (uint32_t i = threadIdx.x; i < n_status; i += blockDim.x) w.status[i] = 0u;
    if (threadIdx.x < 4) w.ctick[threadIdx.x] = 0u;
}

__global__ __launch_bounds__(kWG) void k_spec(KArgs a) {
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t s_stream;
    const uint32_t t = blockIdx.x;
    if (threadIdx.x == 0) s_stream = stream_of_task<false>(a, t);
    __syncthreads();
    const StreamPlan& sp = *plan_of(a, s_stream);
    stage_table(sp, lds, a.w.status);
    const uint32_t c = (t - sp.task_chunk) * kWG + threadIdx.x;
    if (c >= sp.n_chunks) return;
    LdsTab tab{lds};
    lane_spec(sp, a.w, tab, c);
}

__global__ __launch_bounds__(kWG) void k_walk(KArgs a) {
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t s_stream;
    const uint32_t t = blockIdx.x;
    if (threadIdx.x == 0) s_stream = stream_of_task<false>(a, t);
    __syncthreads();
    const StreamPlan& sp = *plan_of(a, s_stream);
    stage_table(sp, lds, a.w.status);
    const uint32_t c = (t - sp.task_chunk) * kWG + threadIdx.x;
    if (c == 0u || c >= sp.n_chunks) return;
    LdsTab tab{lds};
    const uint32_t gc = sp.chunk_base + c;
    uint32_t ok = lane_walk(sp, a.w, tab, c, a.w.exitv[gc - 1u]);
    a.w.fflag[gc] = 0u;
    if (!ok && c + 1u < sp.n_chunks && a.w.wexit[gc] != a.w.exitv[gc]) {
        // the walk into c+1 started from exitv[c], which is not chunk c's exit
        uint32_t slot = atomicAdd(&a.w.status[1], 1u);
        a.w.flist[slot] = gc + 1u;
    }
}

// Fixup rounds: one workgroup; tables are read from global memory (rare path).
struct GlobalTabView {
    const uint32_t* raw;
    uint32_t L;
    __device__ __forceinline__ uint32_t operator[](uint32_t i) const {
        int bad = 0;
        return derive_entry(raw[i], L, i, &bad);
    }
};

__global__ __launch_bounds__(kWG) void k_fixup(KArgs a) {
    // The list holds chunks whose recorded walk start (wentry) may differ from the
    // current exit of their predecessor.  Each round walks them again from that exit
    // and lists any chunk (itself, or its successor) still inconsistent afterwards.
    // The leftmost listed chunk always starts from a proven exit, so the rounds end.
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t s_n, s_uni, s_first;
    uint32_t* list = a.w.flist;
    uint32_t* next = a.w.flist2;
    uint32_t* status = a.w.status;
    for (uint32_t round = 0;; ++round) {
        __syncthreads();
        if (threadIdx.x == 0)
            s_n = __hip_atomic_load(&status[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const uint32_t n = s_n;
        if (n == 0u) break;
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {      // capture + mark
            uint32_t g = list[i];
            const StreamPlan& sp = *plan_of(a, stream_of_chunk(a, g));
            next[i] = chunk_exit(sp, a.w, g - sp.chunk_base - 1u);
            a.w.fflag[g] = 1u;
        }
        __syncthreads();
        // one stream listed (always so for a single stream): stage its table in LDS
        if (threadIdx.x == 0) { s_uni = 1u; s_first = stream_of_chunk(a, list[0]); }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
            if (stream_of_chunk(a, list[i]) != s_first) s_uni = 0u;
        __syncthreads();
        if (s_uni) {
            const StreamPlan& sp = *plan_of(a, s_first);
            stage_table(sp, lds, status);
            LdsTab tab{lds};
            for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
                lane_walk(sp, a.w, tab, list[i] - sp.chunk_base, next[i]);
        } else {
            for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
                uint32_t g = list[i];
                const StreamPlan& sp = *plan_of(a, stream_of_chunk(a, g));
                GlobalTabView tab{sp.table, sp.L};
                lane_walk(sp, a.w, tab, g - sp.chunk_base, next[i]);
            }
        }
        __threadfence();
        __syncthreads();
        if (threadIdx.x == 0) status[2] = 0u;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {      // consistency
            uint32_t g = list[i];
            const StreamPlan& sp = *plan_of(a, stream_of_chunk(a, g));
            uint32_t c = g - sp.chunk_base;
            if (a.w.wentry[g] != chunk_exit(sp, a.w, c - 1u)) {
                uint32_t slot = atomicAdd(&status[2], 1u);
                next[slot] = g;
            }
            if (c + 1u < sp.n_chunks && !a.w.fflag[g + 1u] &&
                a.w.wentry[g + 1u] != chunk_exit(sp, a.w, c)) {
                uint32_t slot = atomicAdd(&status[2], 1u);
                next[slot] = g + 1u;
            }
        }
        __threadfence();
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) a.w.fflag[list[i]] = 0u;
        if (threadIdx.x == 0) {
            status[1] = __hip_atomic_load(&status[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            status[3] += 1u;                                              // rounds
        }
        __threadfence();
        uint32_t* tmp = list; list = next; next = tmp;
    }
}

// Per-stream exclusive scan of chunk step counts -> coff (u64).  One workgroup per
// stream.
__global__ __launch_bounds__(kScanWG) void k_scan(KArgs a) {
    __shared__ uint64_t s_part[kScanWG];
    const StreamPlan& sp = *plan_of(a, blockIdx.x);
    const uint32_t n = sp.n_chunks;
    const uint32_t per = (n + kScanWG - 1u) / kScanWG;
    const uint32_t b = threadIdx.x * per;
    const uint32_t e = min(n, b + per);
    uint64_t sum = 0;
    for (uint32_t c = b; c < e; ++c) sum += chunk_count(sp, a.w, c);
    s_part[threadIdx.x] = sum;
    __syncthreads();
    // Hillis-Steele inclusive scan over kScanWG partials
    for (uint32_t d = 1; d < kScanWG; d <<= 1) {
        uint64_t v = threadIdx.x >= d ? s_part[threadIdx.x - d] : 0ull;
        __syncthreads();
        s_part[threadIdx.x] += v;
        __syncthreads();
    }
    uint64_t run = threadIdx.x ? s_part[threadIdx.x - 1] : 0ull;
    for (uint32_t c = b; c < e; ++c) {
        a.w.coff[sp.chunk_base + c] = run;
        run += chunk_count(sp, a.w, c);
    }
}

__global__ __launch_bounds__(kWG) void k_write(KArgs a) {
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t s_stream;
    const uint32_t t = blockIdx.x;
    if (threadIdx.x == 0) s_stream = stream_of_task<true>(a, t);
    __syncthreads();
    const StreamPlan& sp = *plan_of(a, s_stream);
    stage_table(sp, lds, a.w.status);
    const uint32_t q = (t - sp.task_item) * kWG + threadIdx.x;
    if (q >= sp.n_ck) return;
    uint32_t entry, entry2;
    uint64_t off = checkpoint_offset(sp, a.w, q, &entry);
    uint64_t end = (q + 1u < sp.n_ck) ? checkpoint_offset(sp, a.w, q + 1u, &entry2) : sp.n_out;
    if (end > sp.n_out) end = sp.n_out;
    if (off >= end) return;
    uint64_t cnt = end - off;
    LdsTab tab{lds};
    Cursor cur;
    cursor_init(cur, sp.units, sp.n_units, q << sp.log2K, entry_off(entry));
    uint32_t xi = entry_state(entry);
    uint8_t* o = sp.out + off;
    for (uint64_t i = 0; i < cnt; ++i) {
        uint32_t crossed;
        uint32_t wd = step(tab, cur, xi, crossed);
        o[i] = (uint8_t)(wd >> 8);
    }
}

}  // namespace

// ------------------------------------------------------------------ host side
namespace mans {

static uint32_t ilog2(uint32_t v) { uint32_t r = 0; while ((1u << (r + 1)) <= v) ++r; return r; }

// Chunk/checkpoint lengths for one stream.  Sync chunks must be much longer than
// the typical distance over which a guessed-state decode merges with the true
// path (tens to hundreds of units at L = 4096), so the spec and walk passes stay
// close to one decode of the stream.
void choose_geometry(uint64_t n_units, uint64_t n_out, uint32_t L, uint32_t req_S,
                     uint32_t req_K, uint32_t* S, uint32_t* K) {
    uint32_t k = req_K ? req_K : 16u;
    if (k < 1u) k = 1u;
    k = 1u << ilog2(k);
    uint32_t s = req_S ? req_S : 0u;
    if (!s) {
        // symbols per unit drives the per-unit decode cost and the merge distance
        double spu = n_units ? (double)n_out / (double)n_units : 4.0;
        s = spu > 12.0 ? 4096u : (spu > 6.0 ? 2048u : 1024u);
        (void)L;
        // keep enough chunks to fill the device for mid-size streams
        while (s > 256u && n_units / s < 2048u) s >>= 1;
    }
    s = 1u << ilog2(s < k ? k : s);
    *S = s;
    *K = k;
}

// Workspace layout for a set of plans (shared by single and batch paths).
struct Layout {
    uint64_t chunks = 0, cks = 0, bytes = 0;
    uint64_t o_exitv, o_own, o_went, o_wexit, o_wtot, o_mq, o_wcm, o_ff, o_coff;
    uint64_t o_ck, o_ckc, o_wb, o_wbc, o_status, o_fl, o_fl2, o_tick, o_plans;
};

static uint64_t align256(uint64_t v) { return (v + 255u) & ~255ull; }

Layout make_layout(uint64_t chunks, uint64_t cks, uint64_t n_plans) {
    Layout l;
    l.chunks = chunks;
    l.cks = cks;
    uint64_t o = 0;
    auto take = [&](uint64_t bytes) { uint64_t r = o; o = align256(o + bytes); return r; };
    l.o_status = take(4 * 64);   // first: mans_decode_status reads offset 0
    l.o_exitv = take(4 * chunks);
    l.o_own = take(4 * chunks);
    l.o_went = take(4 * chunks);
    l.o_wexit = take(4 * chunks);
    l.o_wtot = take(4 * chunks);
    l.o_mq = take(4 * chunks);
    l.o_wcm = take(4 * chunks);
    l.o_ff = take(4 * chunks);
    l.o_coff = take(8 * chunks);
    l.o_ck = take(4 * cks);
    l.o_ckc = take(4 * cks);
    l.o_wb = take(4 * cks);
    l.o_wbc = take(4 * cks);
    l.o_fl = take(4 * chunks + 4);
    l.o_fl2 = take(4 * chunks + 4);
    l.o_tick = take(4 * 8);
    l.o_plans = take(sizeof(StreamPlan) * n_plans);
    l.bytes = o;
    return l;
}

Work make_work(void* base, const Layout& l) {
    char* b = (char*)base;
    Work w;
    w.exitv = (uint32_t*)(b + l.o_exitv);
    w.own_total = (uint32_t*)(b + l.o_own);
    w.wentry = (uint32_t*)(b + l.o_went);
    w.wexit = (uint32_t*)(b + l.o_wexit);
    w.wtotal = (uint32_t*)(b + l.o_wtot);
    w.mq = (uint32_t*)(b + l.o_mq);
    w.wcm = (uint32_t*)(b + l.o_wcm);
    w.fflag = (uint32_t*)(b + l.o_ff);
    w.coff = (uint64_t*)(b + l.o_coff);
    w.ck = (uint32_t*)(b + l.o_ck);
    w.ckc = (uint32_t*)(b + l.o_ckc);
    w.wb = (uint32_t*)(b + l.o_wb);
    w.wbc = (uint32_t*)(b + l.o_wbc);
    w.status = (uint32_t*)(b + l.o_status);
    w.flist = (uint32_t*)(b + l.o_fl);
    w.flist2 = (uint32_t*)(b + l.o_fl2);
    w.ctick = (uint32_t*)(b + l.o_tick);
    return w;
}

static int launch_pipeline(const KArgs& a, uint32_t n_chunk_tasks, uint32_t n_item_tasks,
                           uint32_t n_scan_blocks, uint32_t maxL, hipStream_t st,
                           void** ev = nullptr) {
    size_t lds = (size_t)maxL * 4u;
    auto mark = [&](int i) { if (ev) hipEventRecord((hipEvent_t)ev[i], st); };
    mark(0);
    hipLaunchKernelGGL(k_init, dim3(1), dim3(64), 0, st, a.w, 64u);
    mark(1);
    hipLaunchKernelGGL(k_spec, dim3(n_chunk_tasks), dim3(kWG), lds, st, a);
    mark(2);
    hipLaunchKernelGGL(k_walk, dim3(n_chunk_tasks), dim3(kWG), lds, st, a);
    mark(3);
    hipLaunchKernelGGL(k_fixup, dim3(1), dim3(kWG), lds, st, a);
    mark(4);
    hipLaunchKernelGGL(k_scan, dim3(n_scan_blocks), dim3(kScanWG), 0, st, a);
    mark(5);
    hipLaunchKernelGGL(k_write, dim3(n_item_tasks), dim3(kWG), lds, st, a);
    mark(6);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(MANS_EHIP, "kernel launch: %s", hipGetErrorString(e));
    return MANS_OK;
}

static int check_stream_args(const uint32_t* units, uint64_t n_units, const uint8_t* out,
                             uint64_t n_out, const uint32_t* table, uint32_t L,
                             uint32_t first_state, uint32_t first_bit) {
    if (!table) return set_error(MANS_EINVAL, "table is NULL");
    if (L < 2u || L > 16384u) return set_error(MANS_EINVAL, "L=%u outside [2, 16384]", L);
    if (n_out && (!units || !out)) return set_error(MANS_EINVAL, "units/out is NULL");
    if (n_out && n_units == 0) return set_error(MANS_EINVAL, "n_units is 0 but n_out > 0");
    if (n_units >= 0xFFFFFF00ull) return set_error(MANS_EINVAL, "n_units too large");
    if (first_bit > 32u) return set_error(MANS_EINVAL, "first_bit=%u > 32", first_bit);
    if (n_out && (first_state < L || first_state >= 2u * L))
        return set_error(MANS_ECORRUPT, "first_state %u outside [L, 2L)", first_state);
    return MANS_OK;
}

}  // namespace mans

extern "C" int mans_workspace_bytes(uint64_t n_units, uint32_t L, uint64_t* bytes) {
    (void)L;
    if (!bytes) return set_error(MANS_EINVAL, "bytes is NULL");
    // smallest geometry any tuning choice uses: K = 8, S = 256
    uint64_t chunks = n_units / 256u + 2u, cks = n_units / 8u + 2u;
    *bytes = make_layout(chunks, cks, 1).bytes;
    return MANS_OK;
}

extern "C" int mans_workspace_bytes_geometry(uint64_t n_units, uint32_t chunk_units,
                                             uint32_t checkpoint_units, uint64_t* bytes) {
    if (!bytes || !chunk_units || !checkpoint_units) return set_error(MANS_EINVAL, "bad geometry");
    uint64_t chunks = n_units / chunk_units + 2u, cks = n_units / checkpoint_units + 2u;
    *bytes = make_layout(chunks, cks, 1).bytes;
    return MANS_OK;
}

extern "C" int mans_decode(const mans_decode_args* d) {
    if (!d) return set_error(MANS_EINVAL, "args is NULL");
    int rc = check_stream_args(d->units, d->n_units, d->out, d->n_out, d->table, d->L,
                               d->first_state, d->first_bit);
    if (rc) return rc;
    if (d->n_out == 0) return MANS_OK;
    uint32_t S, K;
    choose_geometry(d->n_units, d->n_out, d->L, d->chunk_units, d->checkpoint_units, &S, &K);
    if (K < 8u && !d->checkpoint_units) K = 8u;
    StreamPlan sp;
    std::memset(&sp, 0, sizeof(sp));
    sp.units = d->units;
    sp.out = d->out;
    sp.table = d->table;
    sp.n_out = d->n_out;
    sp.n_units = (uint32_t)d->n_units;
    sp.L = d->L;
    sp.first_state = d->first_state;
    sp.first_bit = d->first_bit;
    sp.S = S;
    sp.K = K;
    sp.log2K = ilog2(K);
    sp.n_chunks = (uint32_t)((d->n_units + S - 1) / S);
    sp.n_ck = (uint32_t)((d->n_units + K - 1) / K);
    sp.chunk_base = 0;
    sp.ck_base = 0;
    sp.task_chunk = 0;
    sp.task_item = 0;
    Layout l = make_layout(sp.n_chunks, sp.n_ck, 1);
    if (!d->workspace || d->workspace_bytes < l.bytes)
        return set_error(MANS_ENOMEM, "workspace of %llu bytes < %llu needed",
                         (unsigned long long)d->workspace_bytes, (unsigned long long)l.bytes);
    KArgs a;
    std::memset(&a, 0, sizeof(a));
    a.one = sp;
    a.plans = nullptr;
    a.n_streams = 1;
    a.w = make_work(d->workspace, l);
    uint32_t nct = (sp.n_chunks + kWG - 1) / kWG;
    uint32_t nit = (sp.n_ck + kWG - 1) / kWG;
    a.n_tasks = nct;
    rc = launch_pipeline(a, nct, nit, 1, sp.L, (hipStream_t)d->stream, d->kernel_events);
    if (rc) return rc;
    if (d->flags & MANS_F_SYNC_ERRORS) {
        int status = 0;
        rc = mans_decode_status(d->workspace, d->stream, &status);
        if (rc) return rc;
        if (status) return set_error(status, "device reported status %d", status);
    }
    return MANS_OK;
}

extern "C" int mans_decode_status(void* workspace, void* stream, int* status) {
    if (!workspace || !status) return set_error(MANS_EINVAL, "NULL argument");
    // status words sit at offset 0 of every layout
    const uint32_t* st = (const uint32_t*)workspace;
    uint32_t v = 0;
    hipError_t e = hipMemcpyAsync(&v, st, 4, hipMemcpyDeviceToHost, (hipStream_t)stream);
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) return set_error(MANS_EHIP, "status read: %s", hipGetErrorString(e));
    *status = (v & 1u) ? MANS_ETABLE : 0;
    return MANS_OK;
}

// --------------------------------------------------------------------- batch
namespace mans {

// Plans for a batch: per-stream geometry, global chunk/checkpoint bases, task bases.
static int build_plans(const mans_stream_desc* d, size_t n, std::vector<StreamPlan>& plans,
                       uint64_t* chunks, uint64_t* cks, uint32_t* chunk_tasks,
                       uint32_t* item_tasks, uint32_t* maxL, uint64_t* units, uint64_t* syms) {
    plans.resize(n);
    uint64_t cb = 0, qb = 0, tc = 0, ti = 0, tu = 0, ts = 0;
    uint32_t ml = 2;
    for (size_t i = 0; i < n; ++i) {
        int rc = check_stream_args(d[i].units, d[i].n_units, d[i].out, d[i].n_out, d[i].table,
                                   d[i].L, d[i].first_state, d[i].first_bit);
        if (rc) return set_error(rc, "stream %zu: %s", i, mans_last_error());
        StreamPlan& sp = plans[i];
        std::memset(&sp, 0, sizeof(sp));
        uint32_t S, K;
        choose_geometry(d[i].n_units, d[i].n_out, d[i].L, 0, 0, &S, &K);
        if (K < 8u) K = 8u;
        sp.units = d[i].units;
        sp.out = d[i].out;
        sp.table = d[i].table;
        sp.n_out = d[i].n_out;
        sp.n_units = (uint32_t)d[i].n_units;
        sp.L = d[i].L;
        sp.first_state = d[i].first_state;
        sp.first_bit = d[i].first_bit;
        sp.S = S;
        sp.K = K;
        sp.log2K = ilog2(K);
        // streams with nothing to decode keep one idle chunk/item so tasks stay valid
        uint64_t nu = d[i].n_out ? d[i].n_units : 1;
        sp.n_chunks = (uint32_t)((nu + S - 1) / S);
        sp.n_ck = d[i].n_out ? (uint32_t)((nu + K - 1) / K) : 0u;
        sp.chunk_base = (uint32_t)cb;
        sp.ck_base = qb;
        sp.task_chunk = (uint32_t)tc;
        sp.task_item = (uint32_t)ti;
        if (!d[i].n_out) sp.n_chunks = 0;
        cb += std::max<uint32_t>(sp.n_chunks, 1u);
        qb += std::max<uint32_t>(sp.n_ck, 1u);
        tc += std::max<uint64_t>((sp.n_chunks + kWG - 1) / kWG, 1);
        ti += std::max<uint64_t>((sp.n_ck + kWG - 1) / kWG, 1);
        tu += d[i].n_units;
        ts += d[i].n_out;
        ml = std::max(ml, sp.L);
    }
    if (cb >= 0xFFFFFFF0ull || tc >= 0x7FFFFFFFull || ti >= 0x7FFFFFFFull)
        return set_error(MANS_EINVAL, "batch too large");
    *chunks = cb;
    *cks = qb;
    *chunk_tasks = (uint32_t)tc;
    *item_tasks = (uint32_t)ti;
    *maxL = ml;
    *units = tu;
    *syms = ts;
    return MANS_OK;
}

}  // namespace mans

extern "C" int mans_batch_workspace_bytes(const mans_stream_desc* descs, size_t n, uint64_t* bytes) {
    if (!descs || !n || !bytes) return set_error(MANS_EINVAL, "bad batch");
    std::vector<StreamPlan> plans;
    uint64_t chunks, cks, u, s;
    uint32_t ct, it, ml;
    int rc = build_plans(descs, n, plans, &chunks, &cks, &ct, &it, &ml, &u, &s);
    if (rc) return rc;
    *bytes = make_layout(chunks, cks, n).bytes;
    return MANS_OK;
}

extern "C" int mans_decode_batch(const mans_stream_desc* descs, size_t n,
                                 const mans_batch_opts* opts, mans_batch_stats* stats) {
    if (!descs || !n || !opts) return set_error(MANS_EINVAL, "bad batch");
    std::vector<StreamPlan> plans;
    uint64_t chunks, cks, units, syms;
    uint32_t ct, it, ml;
    int rc = build_plans(descs, n, plans, &chunks, &cks, &ct, &it, &ml, &units, &syms);
    if (rc) return rc;
    Layout l = make_layout(chunks, cks, n);
    if (!opts->workspace || opts->workspace_bytes < l.bytes)
        return set_error(MANS_ENOMEM, "batch workspace of %llu bytes < %llu needed",
                         (unsigned long long)opts->workspace_bytes, (unsigned long long)l.bytes);
    hipStream_t st = (hipStream_t)opts->stream;
    StreamPlan* dplans = (StreamPlan*)((char*)opts->workspace + l.o_plans);
    hipError_t e = hipMemcpyAsync(dplans, plans.data(), sizeof(StreamPlan) * n,
                                  hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return set_error(MANS_EHIP, "plan upload: %s", hipGetErrorString(e));
    KArgs a;
    std::memset(&a, 0, sizeof(a));
    a.plans = dplans;
    a.n_streams = (uint32_t)n;
    a.n_tasks = ct;
    a.w = make_work(opts->workspace, l);
    rc = launch_pipeline(a, ct, it, (uint32_t)n, ml, st);
    if (rc) return rc;
    if (stats) {
        stats->total_units = units;
        stats->total_symbols = syms;
        stats->n_streams = (uint32_t)n;
        stats->n_chunks = (uint32_t)chunks;
        stats->n_items = (uint32_t)cks;
        stats->status = 0;
    }
    return MANS_OK;
}
