// Row-format range partitioning behind the gpupartition operator: a stable
// counting sort of packed rows by their exactrange bucket, optionally sorted
// inside each bucket by a second column.
//
// Replaces PartitionOp (operators/partition.cpp:166-359): the per-thread
// histograms, the idxstart prefix over (thread, partition), repartition()
// and the optional per-partition sortAllInPage().  Bucket b of the output
// holds exactly the rows the reference writes to partition b, in the same
// order (all of source 0's rows in input order, then source 1's, ...), since
// every pass below is stable.  With sorting, rows of a bucket are ordered by
// the sort column (ties in input order, one of the orders std::sort may
// produce).
//
//   k_pt_prep    bucket = (key - min) / ((max - min + buckets) / buckets) as
//                ExactRangeValueHasher computes it (hash.h:444-479, C
//                division); payload = bucket << 32 | row; sort key mapped to
//                an order-preserving unsigned 64-bit value
//   k_dhist      per chunk of 8192 elements: counts of one 8-bit digit
//   k_dscan      digit-major exclusive scan over chunks -> write offsets
//   k_dscatter   per chunk, tiles of 256 elements in order: rank among the
//                wavefront's lanes with the same digit (ballots on the 8 digit
//                bits), wavefront prefix per digit in LDS -> stable position
//   k_pt_gather  row j of the output = input row (payload_j & 0xffffffff)
// Passes: LSD over the sort key's bytes (4 for int, 8 for long / date / dec),
// then one pass on the bucket.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "pgpu_internal.h"

namespace {

constexpr int kPT = 256;               // threads per block
constexpr uint32_t kChunk = 8192;      // elements per histogram / scatter block
constexpr int kBucketShift = 64;       // digit pass on the bucket (payload >> 32)

__device__ __forceinline__ int64_t pt_load_int(const char* p, uint32_t bytes) {
    if (bytes == 4) {
        int32_t v;
        __builtin_memcpy(&v, p, 4);
        return v;
    }
    int64_t v;
    __builtin_memcpy(&v, p, 8);
    return v;
}

__device__ __forceinline__ uint64_t sort_bits(const char* p, uint32_t type) {
    if (type == PGPU_COL_INT32) {
        int32_t v;
        __builtin_memcpy(&v, p, 4);
        return (uint64_t)((uint32_t)v ^ 0x80000000u);
    }
    uint64_t b;
    __builtin_memcpy(&b, p, 8);
    if (type == PGPU_COL_INT64) return b ^ (1ull << 63);
    if ((b << 1) == 0) b = 0;  // -0.0 sorts as +0.0 (equal under operator<)
    return (b >> 63) ? ~b : (b | (1ull << 63));  // FP64: IEEE order as unsigned order
}

__global__ __launch_bounds__(kPT) void k_pt_prep(const char* __restrict__ rows, uint64_t n, pgpu_part_desc d,
                                                 int64_t bucket_range, uint64_t* __restrict__ skey,
                                                 uint64_t* __restrict__ pay, int32_t* __restrict__ err) {
    const uint64_t stride = (uint64_t)gridDim.x * kPT;
    for (uint64_t i = (uint64_t)blockIdx.x * kPT + threadIdx.x; i < n; i += stride) {
        const char* r = rows + i * d.row_bytes;
        const int64_t val = pt_load_int(r + d.key_offset, d.key_bytes) - d.range_min;
        const int64_t b = val / bucket_range;  // truncating, as the reference's CtLong division
        if (b < 0 || b >= (int64_t)d.buckets) *err = 1;
        pay[i] = ((uint64_t)(b < 0 || b >= (int64_t)d.buckets ? 0 : b) << 32) | i;
        if (skey) skey[i] = sort_bits(r + d.sort_offset, d.sort_type);
    }
}

__device__ __forceinline__ uint32_t digit_of(const uint64_t* k, const uint64_t* p, uint64_t i, int shift) {
    return shift == kBucketShift ? (uint32_t)(p[i] >> 32) : (uint32_t)(k[i] >> shift) & 0xFFu;
}

__global__ __launch_bounds__(kPT) void k_dhist(const uint64_t* __restrict__ k, const uint64_t* __restrict__ p,
                                               uint64_t n, int shift, uint32_t* __restrict__ H) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t r0 = (uint64_t)blockIdx.x * kChunk, r1 = min<uint64_t>(n, r0 + kChunk);
    for (uint64_t i = r0 + threadIdx.x; i < r1; i += kPT) atomicAdd(&h[digit_of(k, p, i, shift)], 1u);
    __syncthreads();
    H[(uint64_t)blockIdx.x * 256 + threadIdx.x] = h[threadIdx.x];
}

// H[c][d] <- position of chunk c's first element with digit d; tot[d] =
// elements with digit d.  One block of 256 threads (thread = digit).
__global__ __launch_bounds__(kPT) void k_dscan(uint32_t* __restrict__ H, uint32_t nchunks, uint32_t* __restrict__ tot) {
    __shared__ uint32_t s[kPT];
    const uint32_t d = threadIdx.x;
    uint32_t t = 0;
    for (uint32_t c = 0; c < nchunks; ++c) t += H[(uint64_t)c * 256 + d];
    tot[d] = t;
    s[d] = t;
    __syncthreads();
    if (d == 0) {
        uint32_t run = 0;
        for (int i = 0; i < kPT; ++i) {
            const uint32_t x = s[i];
            s[i] = run;
            run += x;
        }
    }
    __syncthreads();
    uint32_t run = s[d];
    for (uint32_t c = 0; c < nchunks; ++c) {
        uint32_t* q = &H[(uint64_t)c * 256 + d];
        const uint32_t x = *q;
        *q = run;
        run += x;
    }
}

// Stable scatter of one chunk by one digit.  kout may be null (the key is not
// carried in the bucket pass).
__global__ __launch_bounds__(kPT) void k_dscatter(const uint64_t* __restrict__ kin, const uint64_t* __restrict__ pin,
                                                  uint64_t* __restrict__ kout, uint64_t* __restrict__ pout,
                                                  uint64_t n, int shift, const uint32_t* __restrict__ H) {
    constexpr int NW = kPT / 64;
    __shared__ uint32_t cur[256];
    __shared__ uint32_t wpre[NW][256];
    __shared__ uint32_t ttot[256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    cur[threadIdx.x] = H[(uint64_t)blockIdx.x * 256 + threadIdx.x];
    for (int j = 0; j < NW; ++j) wpre[j][threadIdx.x] = 0;
    __syncthreads();
    const uint64_t r0 = (uint64_t)blockIdx.x * kChunk, r1 = min<uint64_t>(n, r0 + kChunk);
    for (uint64_t t0 = r0; t0 < r1; t0 += kPT) {
        const uint64_t i = t0 + threadIdx.x;
        const bool ok = i < r1;
        const uint32_t dg = ok ? digit_of(kin, pin, i, shift) : 0u;
        uint64_t peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (dg >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & ((1ull << lane) - 1));
        if (ok && rank == 0) wpre[w][dg] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            uint32_t run = 0;
            for (int j = 0; j < NW; ++j) {
                const uint32_t x = wpre[j][threadIdx.x];
                wpre[j][threadIdx.x] = run;
                run += x;
            }
            ttot[threadIdx.x] = run;
        }
        __syncthreads();
        if (ok) {
            const uint32_t dst = cur[dg] + wpre[w][dg] + rank;
            if (kout) kout[dst] = kin[i];
            pout[dst] = pin[i];
        }
        __syncthreads();
        cur[threadIdx.x] += ttot[threadIdx.x];
        for (int j = 0; j < NW; ++j) wpre[j][threadIdx.x] = 0;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kPT) void k_pt_gather(const char* __restrict__ in, uint32_t tsz,
                                                   const uint64_t* __restrict__ pay, uint64_t n, char* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * kPT;
    const bool w8 = tsz % 8 == 0 && ((uintptr_t)in & 7) == 0 && ((uintptr_t)out & 7) == 0;
    const bool w4 = tsz % 4 == 0 && ((uintptr_t)in & 3) == 0 && ((uintptr_t)out & 3) == 0;
    for (uint64_t j = (uint64_t)blockIdx.x * kPT + threadIdx.x; j < n; j += stride) {
        const char* s = in + (pay[j] & 0xFFFFFFFFull) * tsz;
        char* o = out + j * tsz;
        if (w8) {
            for (uint32_t b = 0; b < tsz; b += 8) *(uint64_t*)(o + b) = *(const uint64_t*)(s + b);
        } else if (w4) {
            for (uint32_t b = 0; b < tsz; b += 4) *(uint32_t*)(o + b) = *(const uint32_t*)(s + b);
        } else {
            for (uint32_t b = 0; b < tsz; ++b) o[b] = s[b];
        }
    }
}

inline unsigned pt_grid(uint64_t n) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(pgpu::ceil_div(n, kPT), 8192));
}

}  // namespace

struct pgpu_part {
    pgpu_ctx* ctx = nullptr;
    pgpu_part_desc d{};
    int64_t bucket_range = 1;
    std::mutex mu;
    std::vector<std::vector<char>> staged;  // per source
    // device buffers, grown to the largest input seen
    uint64_t cap = 0;
    char *in = nullptr, *out = nullptr;
    uint64_t *k0 = nullptr, *k1 = nullptr, *p0 = nullptr, *p1 = nullptr;
    uint32_t* H = nullptr;
    uint32_t* tot = nullptr;
    int32_t* err = nullptr;
    int32_t* h_err = nullptr;
    uint32_t* h_tot = nullptr;
    uint64_t n = 0;
    std::vector<uint64_t> start;  // buckets + 1 output bounds
    bool finished = false;
};

using namespace pgpu;

namespace {

void pt_free_buffers(pgpu_part* p) {
    for (void* x : {(void*)p->in, (void*)p->out, (void*)p->k0, (void*)p->k1, (void*)p->p0, (void*)p->p1,
                    (void*)p->H})
        if (x) (void)ctx_free(p->ctx, x);
    p->in = p->out = nullptr;
    p->k0 = p->k1 = p->p0 = p->p1 = nullptr;
    p->H = nullptr;
    p->cap = 0;
}

int pt_reserve(pgpu_part* p, uint64_t n) {
    if (n <= p->cap) return PGPU_OK;
    pt_free_buffers(p);
    const uint64_t cap = std::max<uint64_t>(n, 1024);
    const uint64_t chunks = ceil_div(cap, kChunk);
    const bool sort = p->d.sort != 0;
    PGPU_TRY(ctx_alloc(p->ctx, cap * p->d.row_bytes, (void**)&p->in));
    PGPU_TRY(ctx_alloc(p->ctx, cap * p->d.row_bytes, (void**)&p->out));
    PGPU_TRY(ctx_alloc(p->ctx, cap * 8, (void**)&p->p0));
    PGPU_TRY(ctx_alloc(p->ctx, cap * 8, (void**)&p->p1));
    if (sort) {
        PGPU_TRY(ctx_alloc(p->ctx, cap * 8, (void**)&p->k0));
        PGPU_TRY(ctx_alloc(p->ctx, cap * 8, (void**)&p->k1));
    }
    PGPU_TRY(ctx_alloc(p->ctx, chunks * 256 * 4, (void**)&p->H));
    p->cap = cap;
    return PGPU_OK;
}

// One stable digit pass (kin, pin) -> (kout, pout).
int pt_pass(pgpu_part* p, const uint64_t* kin, const uint64_t* pin, uint64_t* kout, uint64_t* pout, int shift) {
    hipStream_t st = p->ctx->stream;
    const uint32_t chunks = (uint32_t)ceil_div(p->n, kChunk);
    hipLaunchKernelGGL(k_dhist, dim3(chunks), dim3(kPT), 0, st, kin, pin, p->n, shift, p->H);
    hipLaunchKernelGGL(k_dscan, dim3(1), dim3(kPT), 0, st, p->H, chunks, p->tot);
    hipLaunchKernelGGL(k_dscatter, dim3(chunks), dim3(kPT), 0, st, kin, pin, kout, pout, p->n, shift, p->H);
    PGPU_LAUNCH_CHECK();
    return PGPU_OK;
}

}  // namespace

extern "C" {

int pgpu_part_create(pgpu_ctx* ctx, const pgpu_part_desc* desc, pgpu_part** out) {
    if (!ctx || !desc || !out) return fail(PGPU_EINVAL, "pgpu_part_create: null argument");
    *out = nullptr;
    const pgpu_part_desc& d = *desc;
    if (d.row_bytes == 0 || (d.key_bytes != 4 && d.key_bytes != 8) || d.key_offset + d.key_bytes > d.row_bytes)
        return fail(PGPU_EINVAL, "pgpu_part_create: key must be a 4- or 8-byte column inside the row");
    if (d.buckets == 0 || d.buckets > PGPU_PART_MAX_BUCKETS)
        return fail(PGPU_EINVAL, "pgpu_part_create: buckets must be 1..%d", PGPU_PART_MAX_BUCKETS);
    if (d.range_max < d.range_min) return fail(PGPU_EINVAL, "pgpu_part_create: range max < min");
    if (d.sort) {
        const uint32_t w = d.sort_type == PGPU_COL_INT32 ? 4 : 8;
        if (d.sort_type != PGPU_COL_INT32 && d.sort_type != PGPU_COL_INT64 && d.sort_type != PGPU_COL_F64)
            return fail(PGPU_EUNSUPPORTED, "pgpu_part_create: sort column must be int, long / date or dec");
        if (d.sort_offset + w > d.row_bytes) return fail(PGPU_EINVAL, "pgpu_part_create: sort column outside row");
    }
    PGPU_TRY(ctx_activate(ctx));
    auto* p = new pgpu_part();
    p->ctx = ctx;
    p->d = d;
    p->bucket_range = (d.range_max - d.range_min + (int64_t)d.buckets) / (int64_t)d.buckets;
    p->staged.resize(PGPU_PART_MAX_SOURCES);
    int rc = ctx_alloc(ctx, 256 * 4, (void**)&p->tot);
    if (rc == PGPU_OK) rc = ctx_alloc(ctx, 256, (void**)&p->err);
    if (rc == PGPU_OK && (hipHostMalloc((void**)&p->h_err, 4, 0) != hipSuccess ||
                          hipHostMalloc((void**)&p->h_tot, 256 * 4, 0) != hipSuccess))
        rc = fail(PGPU_ENOMEM, "hipHostMalloc failed");
    if (rc != PGPU_OK) {
        pgpu_part_destroy(p);
        return rc;
    }
    *out = p;
    return PGPU_OK;
}

int pgpu_part_destroy(pgpu_part* p) {
    if (!p) return PGPU_OK;
    (void)ctx_activate(p->ctx);
    (void)hipStreamSynchronize(p->ctx->stream);
    pt_free_buffers(p);
    if (p->tot) (void)ctx_free(p->ctx, p->tot);
    if (p->err) (void)ctx_free(p->ctx, p->err);
    if (p->h_err) (void)hipHostFree(p->h_err);
    if (p->h_tot) (void)hipHostFree(p->h_tot);
    delete p;
    return PGPU_OK;
}

int pgpu_part_append(pgpu_part* p, uint32_t source, const void* rows, uint64_t n) {
    if (!p || (n && !rows)) return fail(PGPU_EINVAL, "pgpu_part_append: null argument");
    if (source >= PGPU_PART_MAX_SOURCES)
        return fail(PGPU_EINVAL, "pgpu_part_append: source %u >= %d", source, PGPU_PART_MAX_SOURCES);
    std::lock_guard<std::mutex> g(p->mu);
    if (p->finished) return fail(PGPU_ESTATE, "pgpu_part_append: after finish (call pgpu_part_reset)");
    auto& v = p->staged[source];
    const char* c = (const char*)rows;
    v.insert(v.end(), c, c + n * p->d.row_bytes);
    return PGPU_OK;
}

int pgpu_part_finish(pgpu_part* p, uint64_t* counts) {
    if (!p) return fail(PGPU_EINVAL, "pgpu_part_finish: null handle");
    std::lock_guard<std::mutex> g(p->mu);
    if (p->finished) return fail(PGPU_ESTATE, "pgpu_part_finish: called twice");
    PGPU_TRY(ctx_activate(p->ctx));
    hipStream_t st = p->ctx->stream;
    const uint32_t tsz = p->d.row_bytes;
    uint64_t n = 0;
    for (const auto& v : p->staged) n += v.size() / tsz;
    if (n >= (1ull << 32)) return fail(PGPU_EINVAL, "pgpu_part_finish: %llu rows (limit 2^32 - 1)", (unsigned long long)n);
    p->n = n;
    p->start.assign(p->d.buckets + 1, 0);
    if (n) {
        PGPU_TRY(pt_reserve(p, n));
        uint64_t at = 0;
        for (const auto& v : p->staged) {
            if (v.empty()) continue;
            PGPU_HIP(hipMemcpyAsync(p->in + at, v.data(), v.size(), hipMemcpyHostToDevice, st));
            at += v.size();
        }
        PGPU_HIP(hipMemsetAsync(p->err, 0, 4, st));
        hipLaunchKernelGGL(k_pt_prep, dim3(pt_grid(n)), dim3(kPT), 0, st, p->in, n, p->d, p->bucket_range,
                           p->d.sort ? p->k0 : (uint64_t*)nullptr, p->p0, p->err);
        PGPU_LAUNCH_CHECK();
        uint64_t *ka = p->k0, *kb = p->k1, *pa = p->p0, *pb = p->p1;
        if (p->d.sort) {
            const int bits = p->d.sort_type == PGPU_COL_INT32 ? 32 : 64;
            for (int sh = 0; sh < bits; sh += 8) {
                PGPU_TRY(pt_pass(p, ka, pa, kb, pb, sh));
                std::swap(ka, kb);
                std::swap(pa, pb);
            }
        }
        PGPU_TRY(pt_pass(p, nullptr, pa, nullptr, pb, kBucketShift));
        hipLaunchKernelGGL(k_pt_gather, dim3(pt_grid(n)), dim3(kPT), 0, st, p->in, tsz, pb, n, p->out);
        PGPU_LAUNCH_CHECK();
        PGPU_HIP(hipMemcpyAsync(p->h_tot, p->tot, 256 * 4, hipMemcpyDeviceToHost, st));
        PGPU_HIP(hipMemcpyAsync(p->h_err, p->err, 4, hipMemcpyDeviceToHost, st));
        PGPU_HIP(hipStreamSynchronize(st));
        if (*p->h_err)
            return fail(PGPU_EOVERFLOW, "pgpu_part_finish: a key falls outside the %u buckets of [%lld, %lld]",
                        p->d.buckets, (long long)p->d.range_min, (long long)p->d.range_max);
        for (uint32_t b = 0; b < p->d.buckets; ++b) p->start[b + 1] = p->start[b] + p->h_tot[b];
    }
    for (auto& v : p->staged) std::vector<char>().swap(v);
    if (counts)
        for (uint32_t b = 0; b < p->d.buckets; ++b) counts[b] = p->start[b + 1] - p->start[b];
    p->finished = true;
    return PGPU_OK;
}

int pgpu_part_pull(pgpu_part* p, uint32_t bucket, uint64_t first, uint64_t n, void* dst) {
    if (!p || (n && !dst)) return fail(PGPU_EINVAL, "pgpu_part_pull: null argument");
    if (!p->finished) return fail(PGPU_ESTATE, "pgpu_part_pull: before finish");
    if (bucket >= p->d.buckets) return fail(PGPU_EINVAL, "pgpu_part_pull: bucket %u out of range", bucket);
    const uint64_t cnt = p->start[bucket + 1] - p->start[bucket];
    if (first > cnt || n > cnt - first) return fail(PGPU_EINVAL, "pgpu_part_pull: rows past the bucket");
    if (!n) return PGPU_OK;
    PGPU_TRY(ctx_activate(p->ctx));
    const uint64_t tsz = p->d.row_bytes;
    PGPU_HIP(hipMemcpy(dst, p->out + (p->start[bucket] + first) * tsz, n * tsz, hipMemcpyDeviceToHost));
    return PGPU_OK;
}

int pgpu_part_reset(pgpu_part* p) {
    if (!p) return fail(PGPU_EINVAL, "pgpu_part_reset: null handle");
    std::lock_guard<std::mutex> g(p->mu);
    for (auto& v : p->staged) std::vector<char>().swap(v);
    p->n = 0;
    p->start.clear();
    p->finished = false;
    return PGPU_OK;
}

}  // extern "C"
