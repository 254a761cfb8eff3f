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

// Streaming loads / stores with the non-temporal cache policy: partition
// passes read every byte once, so keeping the lines in the caches only
// evicts useful ones (measured on MI355X: streaming read 6.3 -> 7.0 TB/s,
// copy 5.4 -> 6.0 TB/s).  Stores of short scattered runs must not use it:
// they lose the L2's merging of partial lines (the radix scatters slowed by
// 10-50 %).  T is 4, 8 or 16 bytes, naturally aligned.
template <typename T>
__device__ __forceinline__ T nt_load(const T* p) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8 || sizeof(T) == 16, "nt_load width");
    T r;
    if constexpr (sizeof(T) == 16) {
        typedef long long v2 __attribute__((ext_vector_type(2)));
        const v2 x = __builtin_nontemporal_load((const v2*)p);
        __builtin_memcpy(&r, &x, 16);
    } else if constexpr (sizeof(T) == 8) {
        const long long x = __builtin_nontemporal_load((const long long*)p);
        __builtin_memcpy(&r, &x, 8);
    } else {
        const int x = __builtin_nontemporal_load((const int*)p);
        __builtin_memcpy(&r, &x, 4);
    }
    return r;
}

template <typename T>
__device__ __forceinline__ void nt_store(T* p, const T& v) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8 || sizeof(T) == 16, "nt_store width");
    if constexpr (sizeof(T) == 16) {
        typedef long long v2 __attribute__((ext_vector_type(2)));
        v2 x;
        __builtin_memcpy(&x, &v, 16);
        __builtin_nontemporal_store(x, (v2*)p);
    } else if constexpr (sizeof(T) == 8) {
        long long x;
        __builtin_memcpy(&x, &v, 8);
        __builtin_nontemporal_store(x, (long long*)p);
    } else {
        int x;
        __builtin_memcpy(&x, &v, 4);
        __builtin_nontemporal_store(x, (int*)p);
    }
}

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
