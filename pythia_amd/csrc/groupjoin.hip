// Join + group-by sum over device-resident columns: the C3 query shape
//   SELECT D.grp, SUM(F.val) FROM F JOIN D ON F.key = D.key GROUP BY D.grp
// fused into one probe pass.  Replaces the plan hashjoin(D, F) ->
// aggregate_sum(field = D.grp, sumfield = F.val): HashJoinOp's probe loop
// (operators/join.cpp:503-562) feeding GenericAggregate::remember +
// AggregateSum::fold (operators/genericaggregate.cpp:240-282,
// operators/aggregatesum.cpp:81-107).
//
//   k_dim_clear / k_dim_build   D.key -> D.grp open-addressing table in HBM
//                               (multimap: duplicate keys keep one slot each)
//   k_grp_clear                 group table: key -> FP64 sum (+ count)
//   k_groupjoin                 each fact row: look up its key, then fold its
//                               value into the group slot with hardware FP64
//                               atomics; a workgroup first merges equal
//                               groups of its own rows in LDS
//   k_grp_compact               ready slots -> (grp, sum, count) arrays
// FP64 sums are order-independent only up to rounding: results are checked
// within 1e-12 relative of exact integer-cent sums.
#include <algorithm>
#include <cstring>

#include "pgpu_internal.h"

namespace {

constexpr int32_t kEmpty32 = INT32_MIN;

__device__ __forceinline__ uint32_t hash32(int32_t k) {
    uint64_t z = (uint64_t)(int64_t)k * 0x9E3779B97F4A7C15ull;
    return (uint32_t)(z >> 32);
}

struct DSlot {
    int32_t key;
    int32_t grp;
};

struct GSlot {
    int32_t key;
    uint32_t cnt;
    double sum;
};

__global__ __launch_bounds__(256) void k_dim_clear(DSlot* t, uint64_t cap, unsigned long long* special) {
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < cap; i += stride) {
        t[i].key = kEmpty32;
        t[i].grp = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) special[0] = 0;
}

// Dimension rows whose key equals the empty marker are kept in a side list
// (at most `special_cap` of them; more is reported as an overflow).
__global__ __launch_bounds__(256) void k_dim_build(const int32_t* __restrict__ dkey, const int32_t* __restrict__ dgrp,
                                                   uint64_t n, DSlot* __restrict__ t, uint32_t mask,
                                                   unsigned long long* __restrict__ special, int32_t* special_grp,
                                                   uint32_t special_cap) {
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
        const int32_t k = dkey[i];
        if (k == kEmpty32) {
            const unsigned long long s = atomicAdd(special, 1ull);
            if (s < special_cap) special_grp[s] = dgrp[i];
            continue;
        }
        uint32_t s = hash32(k) & mask;
        while (atomicCAS((int*)&t[s].key, kEmpty32, k) != kEmpty32) s = (s + 1) & mask;
        t[s].grp = dgrp[i];
    }
}

__global__ __launch_bounds__(256) void k_grp_clear(GSlot* t, uint64_t cap) {
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < cap; i += stride) {
        t[i].key = kEmpty32;
        t[i].cnt = 0;
        t[i].sum = 0.0;
    }
}

__device__ __forceinline__ void grp_fold(GSlot* __restrict__ g, uint32_t gmask, int32_t grp, double v, uint32_t c,
                                         unsigned long long* __restrict__ special) {
    if (grp == kEmpty32) {  // the empty marker as a real group key: side slot
        atomicAdd((double*)&special[2], v);
        atomicAdd(&special[1], (unsigned long long)c);
        return;
    }
    uint32_t s = hash32(grp) & gmask;
    while (true) {
        const int32_t prev = atomicCAS((int*)&g[s].key, kEmpty32, grp);
        if (prev == kEmpty32 || prev == grp) {
            atomicAdd(&g[s].sum, v);
            atomicAdd(&g[s].cnt, c);
            return;
        }
        s = (s + 1) & gmask;
    }
}

constexpr int kGJThreads = 256;
constexpr int kLdsGroups = 2048;  // per-workgroup pre-aggregation slots

__global__ __launch_bounds__(kGJThreads) void k_groupjoin(const int32_t* __restrict__ fkey,
                                                         const double* __restrict__ fval, uint64_t n,
                                                         const DSlot* __restrict__ t, uint32_t dmask,
                                                         const unsigned long long* __restrict__ dspecial_n,
                                                         const int32_t* __restrict__ special_grp,
                                                         GSlot* __restrict__ g, uint32_t gmask,
                                                         unsigned long long* __restrict__ gspecial,
                                                         unsigned long long* __restrict__ joined) {
    // LDS pre-aggregation: rows of this workgroup that hit the same group are
    // summed in LDS first, then flushed with one global atomic per group.
    __shared__ int32_t lk[kLdsGroups];
    __shared__ double ls[kLdsGroups];
    __shared__ uint32_t lc[kLdsGroups];
    __shared__ unsigned long long ljoined;
    for (int i = threadIdx.x; i < kLdsGroups; i += kGJThreads) {
        lk[i] = kEmpty32;
        ls[i] = 0.0;
        lc[i] = 0;
    }
    if (threadIdx.x == 0) ljoined = 0;
    __syncthreads();
    const uint64_t nspecial = *dspecial_n;
    unsigned long long my_joined = 0;
    const uint64_t stride = (uint64_t)gridDim.x * kGJThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kGJThreads + threadIdx.x; i < n; i += stride) {
        const int32_t k = fkey[i];
        const double v = fval[i];
        auto fold_one = [&](int32_t grp) {
            ++my_joined;
            if (grp == kEmpty32) {
                grp_fold(g, gmask, grp, v, 1, gspecial);
                return;
            }
            // try the LDS slots (bounded linear probe), else go global
            uint32_t s = hash32(grp) & (kLdsGroups - 1);
            for (int p = 0; p < 8; ++p) {
                const int32_t prev = atomicCAS(&lk[s], kEmpty32, grp);
                if (prev == kEmpty32 || prev == grp) {
                    atomicAdd(&ls[s], v);
                    atomicAdd(&lc[s], 1u);
                    return;
                }
                s = (s + 1) & (kLdsGroups - 1);
            }
            grp_fold(g, gmask, grp, v, 1, gspecial);
        };
        if (k == kEmpty32) {
            for (uint64_t j = 0; j < nspecial; ++j) fold_one(special_grp[j]);
            continue;
        }
        for (uint32_t s = hash32(k) & dmask; t[s].key != kEmpty32; s = (s + 1) & dmask)
            if (t[s].key == k) fold_one(t[s].grp);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kLdsGroups; i += kGJThreads)
        if (lk[i] != kEmpty32) grp_fold(g, gmask, lk[i], ls[i], lc[i], gspecial);
    atomicAdd(&ljoined, my_joined);
    __syncthreads();
    if (threadIdx.x == 0 && ljoined) atomicAdd(joined, ljoined);
}

__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t ws[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) ws[w] = x;
    __syncthreads();
    uint32_t pre = 0;
    for (int k = 0; k < w; ++k) pre += ws[k];
    *total = ws[0] + ws[1] + ws[2] + ws[3];
    __syncthreads();
    return pre + x - v;
}

__global__ __launch_bounds__(256) void k_grp_compact(const GSlot* __restrict__ g, uint64_t cap,
                                                     const unsigned long long* __restrict__ gspecial,
                                                     int32_t* __restrict__ okey, double* __restrict__ osum,
                                                     uint64_t* __restrict__ ocnt,
                                                     unsigned long long* __restrict__ cursor) {
    __shared__ unsigned long long base;
    for (uint64_t b0 = (uint64_t)blockIdx.x * 256; b0 < cap; b0 += (uint64_t)gridDim.x * 256) {
        const uint64_t i = b0 + threadIdx.x;
        const uint32_t live = (i < cap && g[i].key != kEmpty32) ? 1u : 0u;
        uint32_t total;
        const uint32_t off = block_scan(live, &total);
        if (threadIdx.x == 0) base = total ? atomicAdd(cursor, (unsigned long long)total) : 0ull;
        __syncthreads();
        if (live) {
            okey[base + off] = g[i].key;
            osum[base + off] = g[i].sum;
            ocnt[base + off] = g[i].cnt;
        }
        __syncthreads();
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && gspecial[1]) {
        const unsigned long long s = atomicAdd(cursor, 1ull);
        okey[s] = kEmpty32;
        osum[s] = __longlong_as_double((long long)gspecial[2]);
        ocnt[s] = gspecial[1];
    }
}

__global__ void k_zero(unsigned long long* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

inline unsigned grid_for(uint64_t n, uint64_t cap = 8192) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(pgpu::ceil_div(n, 256), cap));
}

constexpr uint32_t kSpecialCap = 1024;

}  // namespace

struct pgpu_groupjoin {
    pgpu_ctx* ctx = nullptr;
    uint64_t maxD = 0, maxF = 0, maxG = 0;
    DSlot* dt = nullptr;
    uint32_t dmask = 0;
    GSlot* gt = nullptr;
    uint32_t gmask = 0;
    unsigned long long* counters = nullptr;  // [0] dim special count, [1..2] group special (cnt, sum bits), [3] joined, [4] cursor
    unsigned long long* h_counters = nullptr;
    int32_t* special_grp = nullptr;
    int32_t* okey = nullptr;
    double* osum = nullptr;
    uint64_t* ocnt = nullptr;
    uint64_t ngroups = 0, joined = 0;
    pgpu::KernelTimer* timer = nullptr;
    std::vector<void*> owned;
};

using namespace pgpu;

namespace {
int gj_alloc(pgpu_groupjoin* p, uint64_t bytes, void** out) {
    PGPU_TRY(ctx_alloc(p->ctx, bytes, out));
    p->owned.push_back(*out);
    return PGPU_OK;
}
struct Timed {
    KernelTimer* t;
    int idx;
    Timed(KernelTimer* t_, const char* n) : t(t_), idx(t_->start(n)) {}
    ~Timed() { t->stop(idx); }
};
}  // namespace

extern "C" {

int pgpu_groupjoin_create(pgpu_ctx* ctx, uint64_t max_dim, uint64_t max_fact, uint64_t max_groups,
                          pgpu_groupjoin** out) {
    if (!ctx || !out) return fail(PGPU_EINVAL, "pgpu_groupjoin_create: null argument");
    *out = nullptr;
    if (max_dim >= (1ull << 31) || max_groups >= (1ull << 31))
        return fail(PGPU_EINVAL, "pgpu_groupjoin_create: dimension / groups must be < 2^31");
    PGPU_TRY(ctx_activate(ctx));
    auto* p = new pgpu_groupjoin();
    p->ctx = ctx;
    p->maxD = max_dim;
    p->maxF = max_fact;
    p->maxG = std::max<uint64_t>(max_groups, 1);
    p->timer = new KernelTimer(ctx->stream);
    const uint64_t dcap = 1ull << ilog2_ceil(std::max<uint64_t>(2 * max_dim, 64));
    const uint64_t gcap = 1ull << ilog2_ceil(std::max<uint64_t>(2 * p->maxG, 64));
    p->dmask = (uint32_t)(dcap - 1);
    p->gmask = (uint32_t)(gcap - 1);
    int rc = PGPU_OK;
    if (rc == PGPU_OK) rc = gj_alloc(p, dcap * sizeof(DSlot), (void**)&p->dt);
    if (rc == PGPU_OK) rc = gj_alloc(p, gcap * sizeof(GSlot), (void**)&p->gt);
    if (rc == PGPU_OK) rc = gj_alloc(p, 8 * sizeof(unsigned long long), (void**)&p->counters);
    if (rc == PGPU_OK) rc = gj_alloc(p, kSpecialCap * sizeof(int32_t), (void**)&p->special_grp);
    if (rc == PGPU_OK) rc = gj_alloc(p, (p->maxG + 1) * sizeof(int32_t), (void**)&p->okey);
    if (rc == PGPU_OK) rc = gj_alloc(p, (p->maxG + 1) * sizeof(double), (void**)&p->osum);
    if (rc == PGPU_OK) rc = gj_alloc(p, (p->maxG + 1) * sizeof(uint64_t), (void**)&p->ocnt);
    if (rc == PGPU_OK && hipHostMalloc((void**)&p->h_counters, 8 * sizeof(unsigned long long), 0) != hipSuccess)
        rc = fail(PGPU_ENOMEM, "hipHostMalloc failed");
    if (rc != PGPU_OK) {
        pgpu_groupjoin_destroy(p);
        return rc;
    }
    *out = p;
    return PGPU_OK;
}

int pgpu_groupjoin_destroy(pgpu_groupjoin* p) {
    if (!p) return PGPU_OK;
    (void)ctx_activate(p->ctx);
    (void)hipStreamSynchronize(p->ctx->stream);
    delete p->timer;
    for (void* x : p->owned) (void)ctx_free(p->ctx, x);
    if (p->h_counters) (void)hipHostFree(p->h_counters);
    delete p;
    return PGPU_OK;
}

int pgpu_groupjoin_run(pgpu_groupjoin* p, const int32_t* d_dkey, const int32_t* d_dgrp, uint64_t nD,
                       const int32_t* d_fkey, const double* d_fval, uint64_t nF, uint64_t* n_groups,
                       uint64_t* n_joined) {
    if (!p || !n_groups) return fail(PGPU_EINVAL, "pgpu_groupjoin_run: null argument");
    if (nD > p->maxD || nF > p->maxF) return fail(PGPU_EINVAL, "pgpu_groupjoin_run: inputs exceed the plan");
    if ((nD && (!d_dkey || !d_dgrp)) || (nF && (!d_fkey || !d_fval)))
        return fail(PGPU_EINVAL, "pgpu_groupjoin_run: null column");
    PGPU_TRY(ctx_activate(p->ctx));
    hipStream_t st = p->ctx->stream;
    const uint64_t dcap = (uint64_t)p->dmask + 1, gcap = (uint64_t)p->gmask + 1;
    {
        Timed t(p->timer, "gj_build");
        hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, st, p->counters, 8);
        hipLaunchKernelGGL(k_dim_clear, dim3(grid_for(dcap)), dim3(256), 0, st, p->dt, dcap, p->counters);
        hipLaunchKernelGGL(k_grp_clear, dim3(grid_for(gcap)), dim3(256), 0, st, p->gt, gcap);
        if (nD)
            hipLaunchKernelGGL(k_dim_build, dim3(grid_for(nD)), dim3(256), 0, st, d_dkey, d_dgrp, nD, p->dt, p->dmask,
                               p->counters, p->special_grp, kSpecialCap);
        PGPU_LAUNCH_CHECK();
    }
    if (nF && nD) {
        Timed t(p->timer, "gj_probe");
        hipLaunchKernelGGL(k_groupjoin, dim3(grid_for(nF, 2048)), dim3(kGJThreads), 0, st, d_fkey, d_fval, nF, p->dt,
                           p->dmask, p->counters, p->special_grp, p->gt, p->gmask, p->counters, p->counters + 3);
        PGPU_LAUNCH_CHECK();
    }
    {
        Timed t(p->timer, "gj_compact");
        hipLaunchKernelGGL(k_grp_compact, dim3(grid_for(gcap)), dim3(256), 0, st, p->gt, gcap, p->counters, p->okey,
                           p->osum, p->ocnt, p->counters + 4);
        PGPU_LAUNCH_CHECK();
    }
    PGPU_HIP(hipMemcpyAsync(p->h_counters, p->counters, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    PGPU_HIP(hipStreamSynchronize(st));
    if (p->h_counters[0] > kSpecialCap)
        return fail(PGPU_EOVERFLOW, "more than %u dimension rows carry the key INT32_MIN", kSpecialCap);
    p->ngroups = p->h_counters[4];
    p->joined = p->h_counters[3];
    if (p->ngroups > p->maxG + 1)
        return fail(PGPU_EOVERFLOW, "%llu groups exceed the plan's %llu", (unsigned long long)p->ngroups,
                    (unsigned long long)p->maxG);
    *n_groups = p->ngroups;
    if (n_joined) *n_joined = p->joined;
    return PGPU_OK;
}

int pgpu_groupjoin_fetch(pgpu_groupjoin* p, int32_t* keys, double* sums, uint64_t* counts, uint64_t n) {
    if (!p) return fail(PGPU_EINVAL, "pgpu_groupjoin_fetch: null plan");
    if (n > p->ngroups) return fail(PGPU_EINVAL, "pgpu_groupjoin_fetch: %llu > %llu groups", (unsigned long long)n,
                                    (unsigned long long)p->ngroups);
    PGPU_TRY(ctx_activate(p->ctx));
    if (keys && n) PGPU_HIP(hipMemcpy(keys, p->okey, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (sums && n) PGPU_HIP(hipMemcpy(sums, p->osum, n * sizeof(double), hipMemcpyDeviceToHost));
    if (counts && n) PGPU_HIP(hipMemcpy(counts, p->ocnt, n * sizeof(uint64_t), hipMemcpyDeviceToHost));
    return PGPU_OK;
}

int pgpu_groupjoin_timing(pgpu_groupjoin* p, int enable) {
    if (!p) return fail(PGPU_EINVAL, "null plan");
    p->timer->enable(enable != 0);
    return PGPU_OK;
}

int pgpu_groupjoin_times(pgpu_groupjoin* p, pgpu_kernel_time* out, int cap, int* n) {
    if (!p || !n) return fail(PGPU_EINVAL, "null argument");
    p->timer->harvest();
    const auto& t = p->timer->totals();
    *n = (int)t.size();
    for (int i = 0; i < (int)t.size() && i < cap; ++i) {
        memset(&out[i], 0, sizeof out[i]);
        snprintf(out[i].name, sizeof out[i].name, "%s", t[i].name.c_str());
        out[i].launches = t[i].launches;
        out[i].total_ms = t[i].ms;
    }
    return PGPU_OK;
}

}  // extern "C"
