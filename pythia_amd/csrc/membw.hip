// Streaming read and copy kernels that calibrate the roofline: what one
// MI355X's HBM sustains for a pure read (the probe / histogram pattern) and
// a read + write copy (the partition pass pattern), timed with HIP events on
// the context stream.  Also the reference size for scaling rocprofv3's
// FETCH_SIZE / WRITE_SIZE counters.
#include "join_common.h"
#include "pgpu_internal.h"

namespace {

using pgpu::dev::nt_load;
using pgpu::dev::nt_store;

// The fastest streaming forms measured on MI355X: four 16-byte accesses in
// flight per thread, non-temporal cache policy, many small workgroups.
constexpr int kU = 4;

__global__ __launch_bounds__(256) void k_read(const uint4* __restrict__ in, uint64_t n,
                                              unsigned long long* __restrict__ sink) {
    const uint64_t stride = (uint64_t)gridDim.x * 256 * kU;
    uint32_t acc = 0;
    for (uint64_t b = (uint64_t)blockIdx.x * 256 * kU + threadIdx.x; b < n; b += stride) {
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const uint64_t i = b + (uint64_t)u * 256;
            if (i < n) {
                const uint4 v = nt_load(&in[i]);
                acc ^= v.x ^ v.y ^ v.z ^ v.w;
            }
        }
    }
    if (acc == 0x9E3779B9u) atomicAdd(sink, 1ull);  // keeps the loads alive
}

__global__ __launch_bounds__(256) void k_copy(const uint4* __restrict__ in, uint4* __restrict__ out, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * 256 * kU;
    for (uint64_t b = (uint64_t)blockIdx.x * 256 * kU + threadIdx.x; b < n; b += stride) {
        uint4 v[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (b + (uint64_t)u * 256 < n) v[u] = nt_load(&in[b + (uint64_t)u * 256]);
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (b + (uint64_t)u * 256 < n) nt_store(&out[b + (uint64_t)u * 256], v[u]);
    }
}

}  // namespace

using namespace pgpu;

extern "C" int pgpu_membw(pgpu_ctx* ctx, uint64_t bytes, int iters, double* read_gbps, double* copy_gbps) {
    if (!ctx || !read_gbps || !copy_gbps || iters < 1 || bytes < 4096)
        return fail(PGPU_EINVAL, "pgpu_membw: bad argument");
    PGPU_TRY(ctx_activate(ctx));
    bytes &= ~uint64_t(15);
    void *a = nullptr, *b = nullptr, *sink = nullptr;
    PGPU_TRY(ctx_alloc(ctx, bytes, &a));
    int rc = ctx_alloc(ctx, bytes, &b);
    if (rc == PGPU_OK) rc = ctx_alloc(ctx, 8, &sink);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == PGPU_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = fail(PGPU_EDEVICE, "hipEventCreate failed");
    if (rc == PGPU_OK) {
        hipStream_t st = ctx->stream;
        const uint64_t n = bytes / 16;
        const unsigned grid = 256 * 128;  // enough workgroups to fill every CU many times
        (void)hipMemsetAsync(a, 1, bytes, st);
        hipLaunchKernelGGL(k_read, dim3(grid), dim3(256), 0, st, (const uint4*)a, n, (unsigned long long*)sink);
        (void)hipEventRecord(e0, st);
        for (int i = 0; i < iters; ++i)
            hipLaunchKernelGGL(k_read, dim3(grid), dim3(256), 0, st, (const uint4*)a, n, (unsigned long long*)sink);
        (void)hipEventRecord(e1, st);
        float ms = 0;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            rc = fail(PGPU_EDEVICE, "event timing failed");
        *read_gbps = (double)bytes * iters / (ms * 1e-3) / 1e9;
        hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, st, (const uint4*)a, (uint4*)b, n);
        (void)hipEventRecord(e0, st);
        for (int i = 0; i < iters; ++i)
            hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, st, (const uint4*)a, (uint4*)b, n);
        (void)hipEventRecord(e1, st);
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            rc = fail(PGPU_EDEVICE, "event timing failed");
        *copy_gbps = 2.0 * (double)bytes * iters / (ms * 1e-3) / 1e9;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)ctx_free(ctx, a);
    if (b) (void)ctx_free(ctx, b);
    if (sink) (void)ctx_free(ctx, sink);
    return rc;
}
