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

// Workgroups of a launch are dispatched to the 8 XCDs (each with its own L2)
// in turn: workgroup i runs on XCD i % 8.  xcd_remap(i, n) renumbers them so
// that XCD x takes one contiguous range of logical blocks [0, n): blocks whose
// output runs meet in one 128-byte line then complete it in the same L2.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_remap(uint32_t i, uint32_t n) {
    const uint32_t q = n / kXcds, r = n % kXcds;
    const uint32_t x = i % kXcds, k = i / kXcds;  // XCD x runs q + (x < r) of the blocks
    return x * q + min(x, r) + k;
}

// As xcd_remap, in chunks of K consecutive logical blocks: chunk c runs on
// XCD c % 8, so neighbours within a chunk share an L2 while a range of
// unusually heavy blocks (a hot partition's units) still spreads over all
// XCDs.  Blocks past the last whole round of 8 chunks keep their index.
template <uint32_t K>
__device__ __forceinline__ uint32_t xcd_remap_chunked(uint32_t i, uint32_t n) {
    const uint32_t full = n / (kXcds * K) * (kXcds * K);
    if (i >= full) return i;
    const uint32_t x = i % kXcds, k = i / kXcds;  // k-th block on XCD x
    return ((k / K) * kXcds + x) * K + k % K;
}

// Reads through a buffer resource: one 32-bit offset register per load in
// flight instead of a 64-bit address, so more loads stay in registers.  The
// base and `bytes` must be the same for the whole wavefront (both are read
// from its first lane, so the resource lives in scalar registers even where
// the compiler cannot prove it uniform: otherwise every load is wrapped in a
// loop over the distinct resources of the wavefront); offsets at or past
// `bytes` read zero.  Word 3 = 0x00020000 selects the untyped 32-bit data
// format of gfx9 buffer resources.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, uint32_t bytes) {
    const uint64_t a = (uint64_t)base;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    const uint32_t nb = __builtin_amdgcn_readfirstlane(bytes);
    return __builtin_amdgcn_make_buffer_rsrc((void*)(((uint64_t)hi << 32) | lo), (short)0, (int)nb, 0x00020000);
}

// NT: the non-temporal cache policy (cache-policy operand bit 1, the same
// `nt` bit nt_load sets on a global load).
template <typename T, bool NT = false>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8 || sizeof(T) == 16, "buf_load width");
    constexpr int kPol = NT ? 2 : 0;
    T v;
    if constexpr (sizeof(T) == 16) {
        typedef int v4i __attribute__((ext_vector_type(4)));
        const v4i x = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, kPol);
        __builtin_memcpy(&v, &x, 16);
    } else if constexpr (sizeof(T) == 8) {
        typedef int v2i __attribute__((ext_vector_type(2)));
        const v2i x = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, kPol);
        __builtin_memcpy(&v, &x, 8);
    } else {
        const int x = __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, kPol);
        __builtin_memcpy(&v, &x, 4);
    }
    return v;
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
