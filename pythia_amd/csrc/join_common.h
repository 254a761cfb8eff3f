// Device helpers shared by the batched and served join kernels: the radix
// hash, row element layouts, the LDS table slot and wavefront / block scans.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pgpu {
namespace dev {

constexpr uint64_t kFib = 0x9E3779B97F4A7C15ull;

// 64-bit multiplicative hash of a signed key (int32 keys are sign-extended so
// an int key and a long key of the same value hash alike).  Partitions take
// the top bits, LDS slots the bits just below them.
template <typename KT>
__device__ __forceinline__ uint64_t hmix(KT k) {
    return (uint64_t)(int64_t)k * kFib;
}

template <typename KT>
struct KeyEmpty;
template <>
struct KeyEmpty<int32_t> {
    static constexpr int32_t v = INT32_MIN;
};
template <>
struct KeyEmpty<int64_t> {
    static constexpr int64_t v = INT64_MIN;
};

// A partitioned row: key and (when HV) payload, packed.
template <typename KT, bool HV>
struct Row;
template <typename KT>
struct Row<KT, true> {
    KT k;
    KT v;
};
template <typename KT>
struct Row<KT, false> {
    KT k;
};

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Exclusive scan of one value per thread over a block of T threads (T a
// multiple of 64, at most 1024).  `ws` is LDS scratch of T/64 + 1 words.
// Returns the thread's exclusive prefix; *total receives the block sum.
template <int T, typename V>
__device__ __forceinline__ V block_excl_scan(V v, V* ws, V* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    V x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        V y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) ws[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        V run = 0;
        for (int i = 0; i < T / 64; ++i) {
            V t = ws[i];
            ws[i] = run;
            run += t;
        }
        ws[T / 64] = run;
    }
    __syncthreads();
    const V r = ws[w] + x - v;
    *total = ws[T / 64];
    __syncthreads();
    return r;
}

}  // namespace dev
}  // namespace pgpu
