// Join + group-by sum over device-resident columns: the C3 query shape
//   SELECT D.grp, SUM(F.val) FROM F JOIN D ON F.key = D.key GROUP BY D.grp
// Replaces the plan hashjoin(D, F) -> aggregate_sum(field = D.grp,
// sumfield = F.val): HashJoinOp's build and probe loops
// (operators/join.cpp:396-464, :503-562) feeding GenericAggregate::remember +
// AggregateSum::fold (operators/genericaggregate.cpp:240-282,
// operators/aggregatesum.cpp:81-107).
//
// PARTITIONED path (dimension keys unique within a range of at most 2^25
// values, group keys within a range of at most 2^22): no global atomics on
// the row streams.  Every fact row moves twice at streaming bandwidth instead
// of taking a random HBM lookup (60 M random 4-byte gathers from a 60 MB
// table cost ~1.1 ms on MI355X) and a random FP64 atomic (~2.6 ms):
//   k_minmax               key and group ranges of D (host decides the path)
//   k_rhist / k_cscan /    D and F radix-partitioned by key range: partition
//   k_tscan / k_rscatter   p = (key - kmin) >> 15 holds 32 K consecutive keys;
//                          rows are stored as (16-bit local key, payload) with
//                          D's payload turned into a dense group id
//                          gid = grp - gmin; the histogram is per chunk of
//                          16 tiles, with each tile's offsets in its chunk,
//                          and the scatter runs one workgroup per tile
//   k_items                per partition, work items of <= 64 K rows
//   k_phist                per item: the partition's 32 K-entry slice
//                          local key -> gid is built in LDS from D's rows,
//                          every fact row looks its group up there; writes
//                          the row's gid, counts rows per group range
//                          (gid >> 13) and flags duplicate dimension keys
//   k_pscatter             one workgroup per 6144-row tile of a probe item:
//                          (gid & 8191, value) rows into group-range
//                          partitions at offsets of its own (k_phist counts
//                          each tile's group ranges per wavefront)
//   k_gagg                 per item of a group range: LDS FP64 / count tables
//                          of 8192 groups, one partial table per item
//   k_gcompact             partial tables summed per group, non-empty groups
//                          written out
// A duplicate dimension key (a multi-match join) is detected in pass 1; the
// result of the partitioned path is then discarded and the general path runs.
//
// GENERAL path (any keys, duplicates, the INT32_MIN marker):
//   k_dim_clear / k_dim_build  D.key -> D.grp open-addressing table in HBM
//                              (multimap: duplicate keys keep one slot each)
//   k_grp_clear                group table: key -> FP64 sum (+ count)
//   k_groupjoin                each fact row: look up its key, then fold its
//                              value into the group slot with hardware FP64
//                              atomics; a workgroup first merges equal groups
//                              of its own rows in LDS
//   k_grp_compact              used slots -> (grp, sum, count) arrays
// FP64 sums are order-independent only up to rounding: results are checked
// within 1e-12 relative of exact integer-cent sums.
#include <algorithm>
#include <cstring>

#include "join_common.h"
#include "pgpu_internal.h"

using pgpu::dev::block_excl_scan;
using pgpu::dev::nt_load;

namespace {

constexpr int32_t kEmpty32 = INT32_MIN;

// ---------------------------------------------------------------------------
// General path
// ---------------------------------------------------------------------------

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

// counters: [0] dimension rows with the marker key, [1] / [2] count and sum
// bits of the marker group, [3] joined fact rows, [4] output cursor,
// [5] group table full
enum { kCtDimSpecial = 0, kCtGrpCnt = 1, kCtGrpSum = 2, kCtJoined = 3, kCtCursor = 4, kCtFull = 5, kCtWords = 8 };

__global__ __launch_bounds__(256) void k_dim_clear(DSlot* t, uint64_t cap) {
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < cap; i += stride) {
        t[i].key = kEmpty32;
        t[i].grp = 0;
    }
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
        // the table holds at least twice the plan's dimension rows: never full
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

// Folds (v, c) into group grp.  A table with no free slot left sets
// ct[kCtFull] and drops the fold (the run then reports an overflow).
__device__ __forceinline__ void grp_fold(GSlot* __restrict__ g, uint32_t gmask, int32_t grp, double v, uint32_t c,
                                         unsigned long long* __restrict__ ct) {
    if (grp == kEmpty32) {  // the empty marker as a real group key: side slot
        atomicAdd((double*)&ct[kCtGrpSum], v);
        atomicAdd(&ct[kCtGrpCnt], (unsigned long long)c);
        return;
    }
    uint32_t s = hash32(grp) & gmask;
    for (uint32_t probes = 0; probes <= gmask; ++probes) {
        const int32_t prev = atomicCAS((int*)&g[s].key, kEmpty32, grp);
        if (prev == kEmpty32 || prev == grp) {
            atomicAdd(&g[s].sum, v);
            atomicAdd(&g[s].cnt, c);
            return;
        }
        s = (s + 1) & gmask;
    }
    ct[kCtFull] = 1;
}

constexpr int kGJThreads = 256;
constexpr int kLdsGroups = 2048;  // per-workgroup pre-aggregation slots

__global__ __launch_bounds__(kGJThreads) void k_groupjoin(const int32_t* __restrict__ fkey,
                                                         const double* __restrict__ fval, uint64_t n,
                                                         const DSlot* __restrict__ t, uint32_t dmask,
                                                         const int32_t* __restrict__ special_grp,
                                                         uint32_t special_cap, GSlot* __restrict__ g,
                                                         uint32_t gmask, unsigned long long* __restrict__ ct) {
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
    const uint64_t nspecial = min<unsigned long long>(ct[kCtDimSpecial], special_cap);
    unsigned long long my_joined = 0;
    const uint64_t stride = (uint64_t)gridDim.x * kGJThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kGJThreads + threadIdx.x; i < n; i += stride) {
        const int32_t k = fkey[i];
        const double v = fval[i];
        auto fold_one = [&](int32_t grp) {
            ++my_joined;
            if (grp == kEmpty32) {
                grp_fold(g, gmask, grp, v, 1, ct);
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
            grp_fold(g, gmask, grp, v, 1, ct);
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
        if (lk[i] != kEmpty32) grp_fold(g, gmask, lk[i], ls[i], lc[i], ct);
    atomicAdd(&ljoined, my_joined);
    __syncthreads();
    if (threadIdx.x == 0 && ljoined) atomicAdd(&ct[kCtJoined], ljoined);
}

// Used slots -> output rows; positions past `out_cap` are counted, not
// written (the host reports the overflow).
__global__ __launch_bounds__(256) void k_grp_compact(const GSlot* __restrict__ g, uint64_t cap,
                                                     unsigned long long* __restrict__ ct, int32_t* __restrict__ okey,
                                                     double* __restrict__ osum, uint64_t* __restrict__ ocnt,
                                                     uint64_t out_cap) {
    __shared__ unsigned long long base;
    __shared__ uint32_t ws[256 / 64 + 1];
    for (uint64_t b0 = (uint64_t)blockIdx.x * 256; b0 < cap; b0 += (uint64_t)gridDim.x * 256) {
        const uint64_t i = b0 + threadIdx.x;
        const uint32_t live = (i < cap && g[i].key != kEmpty32) ? 1u : 0u;
        uint32_t total;
        const uint32_t off = block_excl_scan<256>(live, ws, &total);
        if (threadIdx.x == 0) base = total ? atomicAdd(&ct[kCtCursor], (unsigned long long)total) : 0ull;
        __syncthreads();
        if (live && base + off < out_cap) {
            okey[base + off] = g[i].key;
            osum[base + off] = g[i].sum;
            ocnt[base + off] = g[i].cnt;
        }
        __syncthreads();
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && ct[kCtGrpCnt]) {
        const unsigned long long s = atomicAdd(&ct[kCtCursor], 1ull);
        if (s < out_cap) {
            okey[s] = kEmpty32;
            osum[s] = __longlong_as_double((long long)ct[kCtGrpSum]);
            ocnt[s] = ct[kCtGrpCnt];
        }
    }
}

__global__ void k_zero(unsigned long long* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

inline unsigned grid_for(uint64_t n, uint64_t cap = 8192) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(pgpu::ceil_div(n, 256), cap));
}

constexpr uint32_t kSpecialCap = 1024;

// ---------------------------------------------------------------------------
// Partitioned path
// ---------------------------------------------------------------------------
namespace part {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kSK = 15;                                        // keys per range partition: 2^15
constexpr int kSG = 13;                                        // groups per group-range partition: 2^13
constexpr uint32_t kP1Max = 1024;                              // key-range partitions
constexpr uint32_t kP2Max = 512;                               // group-range partitions
constexpr uint64_t kMaxKeyRange = (uint64_t)kP1Max << kSK;     // 2^25
constexpr uint64_t kMaxGroupRange = (uint64_t)kP2Max << kSG;   // 2^22
constexpr uint32_t kCR = 65536;                                // rows per histogram chunk (16 scatter tiles)
constexpr uint32_t kIR1 = 196608;                              // rows per probe item
constexpr uint32_t kIR2 = 98304;                               // rows per aggregation item
constexpr int kT = 1024;                                       // threads per workgroup (scatter, probe, agg)
constexpr int kMinmaxBlocks = 1024;
constexpr int kRangeTile = 4096;                               // rows per LDS staging tile: key-range scatter
constexpr uint32_t kTilesPerChunk = kCR / kRangeTile;
constexpr int kRT = 512;                                       // key-range scatter threads (two workgroups per CU)
constexpr int kGroupTile = 6144;                               // rows per LDS staging tile: group-range scatter
constexpr uint32_t kTilesPerItem1 = kIR1 / kGroupTile;         // group-scatter tiles per probe item
constexpr int kPST = 512;                                      // group-range scatter threads (two workgroups per CU)
static_assert(kIR1 % kGroupTile == 0 && kTilesPerItem1 <= 2 * (kT / 64) && kTilesPerItem1 <= 256,
              "two tiles per probe-pass wavefront; tile numbers fit 8 bits");

struct Item {
    uint32_t part, r0, r1, pad;
};

// mm words: [0] min key, [1] max key, [2] min group, [3] max group,
// [4] duplicate dimension key seen, [5] / [6] item counts (probe, aggregation),
// [7] group-scatter tile count
enum { kMmKMin = 0, kMmKMax, kMmGMin, kMmGMax, kMmDup, kMmItems1, kMmItems2, kMmTiles1, kMmWords = 8 };

__global__ void k_mm_init(int32_t* mm) {
    if (threadIdx.x < kMmWords)
        mm[threadIdx.x] = (threadIdx.x == kMmKMin || threadIdx.x == kMmGMin)   ? INT32_MAX
                          : (threadIdx.x == kMmKMax || threadIdx.x == kMmGMax) ? INT32_MIN
                                                                               : 0;
}

// Key and group ranges of D in two stages (no same-address atomics): each of
// kMinmaxBlocks blocks writes its (min key, max key, min grp, max grp) to
// part[block], then one block folds them into mm.
__device__ __forceinline__ void mm_fold(int32_t* a, const int32_t* b) {
    a[0] = min(a[0], b[0]);
    a[1] = max(a[1], b[1]);
    a[2] = min(a[2], b[2]);
    a[3] = max(a[3], b[3]);
}

__device__ __forceinline__ void mm_block(int32_t* v, int32_t* __restrict__ out) {
    __shared__ int32_t sw[kT / 64][4];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        int32_t o[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) o[j] = __shfl_xor(v[j], d, 64);
        mm_fold(v, o);
    }
    if ((threadIdx.x & 63) == 0)
        for (int j = 0; j < 4; ++j) sw[threadIdx.x >> 6][j] = v[j];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) mm_fold(v, sw[w]);
        for (int j = 0; j < 4; ++j) out[j] = v[j];
    }
}

// Each lane folds groups of four rows (one 16-byte load per column, whole
// lines per load instruction), four groups in flight per step.
__global__ __launch_bounds__(kT) void k_minmax(const int32_t* __restrict__ dkey, const int32_t* __restrict__ dgrp,
                                               uint64_t n, int32_t* __restrict__ part) {
    int32_t v[4] = {INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN};
    auto fold = [&](int32_t k, int32_t g) {
        const int32_t o[4] = {k, k, g, g};
        mm_fold(v, o);
    };
    const uint64_t nthr = (uint64_t)gridDim.x * kT, tid = (uint64_t)blockIdx.x * kT + threadIdx.x;
    uint64_t done = 0;  // rows [0, done) are covered by the vector loop
    if (((uintptr_t)dkey & 15) == 0 && ((uintptr_t)dgrp & 15) == 0) {
        constexpr int U = 4;
        const uint64_t ng = n / 4;  // whole groups of four rows
        const int4* kv = (const int4*)dkey;
        const int4* gv = (const int4*)dgrp;
        for (uint64_t g0 = tid; g0 < ng; g0 += U * nthr) {
            int4 a[U], b[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint64_t g = g0 + (uint64_t)u * nthr;
                if (g < ng) {
                    a[u] = kv[g];  // cached: the histogram pass reads the keys next
                    b[u] = gv[g];
                } else {
                    a[u] = make_int4(INT32_MAX, INT32_MAX, INT32_MAX, INT32_MAX);
                    b[u] = make_int4(INT32_MAX, INT32_MAX, INT32_MAX, INT32_MAX);
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (g0 + (uint64_t)u * nthr < ng) {
                    fold(a[u].x, b[u].x);
                    fold(a[u].y, b[u].y);
                    fold(a[u].z, b[u].z);
                    fold(a[u].w, b[u].w);
                }
            }
        }
        done = ng * 4;
    }
    for (uint64_t i = done + tid; i < n; i += nthr) fold(dkey[i], dgrp[i]);
    mm_block(v, part + 4 * blockIdx.x);
}

__global__ __launch_bounds__(kT) void k_minmax2(const int32_t* __restrict__ part, uint32_t nblocks,
                                                int32_t* __restrict__ mm) {
    int32_t v[4] = {INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN};
    for (uint32_t b = threadIdx.x; b < nblocks; b += kT) mm_fold(v, part + 4 * b);
    mm_block(v, mm + kMmKMin);
}

// Key-range histogram of one chunk of kCR rows: H[chunk][p] counts the rows
// whose key lies in [kmin, kmin + range), p = (key - kmin) >> kSK, and
// toff[tile][p] the chunk's rows of partition p in its tiles (kRangeTile
// rows each) before `tile`.  Each wavefront counts whole tiles (tiles w and
// w + 8 of the chunk) in its own LDS row: no barrier during the sweep.
__global__ __launch_bounds__(512) void k_rhist(const int32_t* __restrict__ key, uint64_t n, int64_t kmin,
                                               uint64_t range, uint32_t Ppad, uint32_t* __restrict__ H,
                                               uint16_t* __restrict__ toff) {
    static_assert(kTilesPerChunk == 16 && kCR < 65536 + 1, "two tiles per wavefront; in-chunk offsets fit 16 bits");
    __shared__ uint32_t lt[8][kP1Max];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (uint32_t i = lane; i < Ppad; i += 64) lt[w][i] = 0;
    const uint64_t r0 = (uint64_t)blockIdx.x * kCR, r1 = min<uint64_t>(n, r0 + kCR);
    const bool vec = ((uintptr_t)key & 15) == 0;  // tile starts are multiples of 4 rows
    auto add = [&](int32_t k) {
        const uint64_t off = (uint64_t)((int64_t)k - kmin);
        if (off < range) atomicAdd(&lt[w][off >> kSK], 1u);
    };
    for (uint32_t t = w; t < kTilesPerChunk; t += 8) {
        const uint64_t t0 = r0 + (uint64_t)t * kRangeTile;
        if (t0 >= r1) break;
        const uint64_t te = min<uint64_t>(r1, t0 + kRangeTile);
        uint64_t r = t0;
        if (vec) {
            const uint64_t ng = (te - t0) / 4;
            const int4* kv = (const int4*)(key + t0);
#pragma unroll 4
            for (uint64_t g = lane; g < ng; g += 64) {
                const int4 q = nt_load(&kv[g]);
                add(q.x);
                add(q.y);
                add(q.z);
                add(q.w);
            }
            r = t0 + ng * 4;
        }
        for (r += lane; r < te; r += 64) add(key[r]);
        // one wavefront's LDS operations complete in order: its counts are final
        uint16_t* tc = toff + (uint64_t)(r0 / kRangeTile + t) * Ppad;
        for (uint32_t i = lane; i < Ppad; i += 64) {
            tc[i] = (uint16_t)lt[w][i];
            lt[w][i] = 0;
        }
    }
    __syncthreads();  // every tile's counts are in toff
    const uint64_t t1 = (r1 - r0 + kRangeTile - 1) / kRangeTile;
    uint16_t* tc = toff + (uint64_t)(r0 / kRangeTile) * Ppad;
    for (uint32_t i = threadIdx.x; i < Ppad; i += 512) {
        uint32_t c[kTilesPerChunk];  // all of the column's counts loaded before the first store
#pragma unroll
        for (uint32_t t = 0; t < kTilesPerChunk; ++t) c[t] = t < t1 ? tc[t * Ppad + i] : 0u;
        uint32_t run = 0;
#pragma unroll
        for (uint32_t t = 0; t < kTilesPerChunk; ++t) {
            if (t < t1) tc[t * Ppad + i] = (uint16_t)run;
            run += c[t];
        }
        H[(uint64_t)blockIdx.x * Ppad + i] = run;
    }
}

// Exclusive scan of every column of H[rows][Ppad] in place; tot[d] receives
// column d's sum.  One block per kScanCols columns, 64 row segments per
// column (Ppad a multiple of kScanCols).
constexpr uint32_t kScanCols = 16;
__global__ __launch_bounds__(1024) void k_cscan(uint32_t* __restrict__ H, uint32_t rows, uint32_t Ppad,
                                                uint32_t* __restrict__ tot) {
    constexpr uint32_t SG = 1024 / kScanCols;
    static_assert(SG == 64, "one wavefront scans one column's segments");
    __shared__ uint32_t s[SG][kScanCols + 1];
    const uint32_t c = threadIdx.x % kScanCols, seg = threadIdx.x / kScanCols;
    const uint32_t d = blockIdx.x * kScanCols + c;
    const uint32_t per = (rows + SG - 1) / SG;
    const uint32_t a = min(rows, seg * per), b = min(rows, a + per);
    constexpr int B = 8;  // rows per step: their counts are loaded together
    uint32_t sum = 0;
    for (uint32_t r0 = a; r0 < b; r0 += B) {
#pragma unroll
        for (int j = 0; j < B; ++j) sum += r0 + j < b ? H[(uint64_t)(r0 + j) * Ppad + d] : 0u;
    }
    s[seg][c] = sum;
    __syncthreads();
    {
        const uint32_t col = threadIdx.x >> 6, lane = threadIdx.x & 63;  // wave col scans column col
        const uint32_t v = s[lane][col];
        uint32_t x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (lane >= (uint32_t)o) x += y;
        }
        s[lane][col] = x - v;
        if (lane == 63) tot[blockIdx.x * kScanCols + col] = x;
    }
    __syncthreads();
    uint32_t run = s[seg][c];
    for (uint32_t r0 = a; r0 < b; r0 += B) {
        uint32_t v[B];
#pragma unroll
        for (int j = 0; j < B; ++j) v[j] = r0 + j < b ? H[(uint64_t)(r0 + j) * Ppad + d] : 0u;
#pragma unroll
        for (int j = 0; j < B; ++j) {
            if (r0 + j < b) H[(uint64_t)(r0 + j) * Ppad + d] = run;
            run += v[j];
        }
    }
}

// base[d] = sum of tot[0..d) for d <= Ppad (Ppad <= 2048).  One block.
__global__ __launch_bounds__(1024) void k_tscan(const uint32_t* __restrict__ tot, uint32_t Ppad,
                                                uint32_t* __restrict__ base) {
    __shared__ uint32_t ws[1024 / 64 + 1];
    constexpr int PER = kP1Max / 1024;
    uint32_t loc[PER], sum = 0;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const uint32_t idx = threadIdx.x * PER + u;
        loc[u] = idx < Ppad ? tot[idx] : 0u;
        sum += loc[u];
    }
    uint32_t total;
    uint32_t ex = block_excl_scan<1024>(sum, ws, &total);
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const uint32_t idx = threadIdx.x * PER + u;
        if (idx < Ppad) base[idx] = ex;
        ex += loc[u];
    }
    if (threadIdx.x == 0) base[Ppad] = total;
}

// Work items of at most `ir` rows over the partitions' row ranges
// [base[p], base[p+1]), p < P (P <= 2048).  first[p] = partition p's first
// item, first[P] = *nitems = the item count.  With tile_rows (ir a multiple
// of it), the items are also cut into tiles of tile_rows rows: tile0[i] =
// item i's first tile, titem[t] = (item << 8) | tile-in-item, *ntiles = the
// tile count.  One block.
__global__ __launch_bounds__(1024) void k_items(const uint32_t* __restrict__ base, uint32_t P, uint32_t ir,
                                                Item* __restrict__ items, uint32_t* __restrict__ first,
                                                int32_t* __restrict__ nitems, uint32_t tile_rows,
                                                uint32_t* __restrict__ tile0, uint32_t* __restrict__ titem,
                                                int32_t* __restrict__ ntiles) {
    __shared__ uint32_t ws[1024 / 64 + 1];
    constexpr int PER = kP1Max / 1024;
    uint32_t cnt[PER], tcnt[PER], sum = 0, tsum = 0;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const uint32_t p = threadIdx.x * PER + u;
        const uint32_t rows = p < P ? base[p + 1] - base[p] : 0u;
        cnt[u] = (rows + ir - 1) / ir;
        tcnt[u] = tile_rows ? rows / ir * (ir / tile_rows) + (rows % ir + tile_rows - 1) / tile_rows : 0u;
        sum += cnt[u];
        tsum += tcnt[u];
    }
    uint32_t total, ttotal;
    uint32_t ex = block_excl_scan<1024>(sum, ws, &total);
    uint32_t tex = block_excl_scan<1024>(tsum, ws, &ttotal);
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const uint32_t p = threadIdx.x * PER + u;
        if (p < P) {
            first[p] = ex;
            const uint32_t b0 = base[p], b1 = base[p + 1];
            for (uint32_t j = 0; j < cnt[u]; ++j) {
                Item it;
                it.part = p;
                it.r0 = b0 + j * ir;
                it.r1 = min(b1, it.r0 + ir);
                it.pad = 0;
                items[ex + j] = it;
                if (tile_rows) {
                    const uint32_t tb = tex + j * (ir / tile_rows);
                    tile0[ex + j] = tb;
                    const uint32_t nt = (it.r1 - it.r0 + tile_rows - 1) / tile_rows;
                    for (uint32_t k = 0; k < nt; ++k) titem[tb + k] = ((ex + j) << 8) | k;
                }
            }
        }
        ex += cnt[u];
        tex += tcnt[u];
    }
    if (threadIdx.x == 0) {
        first[P] = total;
        *nitems = (int32_t)total;
        if (tile_rows) *ntiles = (int32_t)ttotal;
    }
}

template <typename VIn, typename VOut>
__device__ __forceinline__ VOut conv(VIn v, int64_t gmin);
template <>
__device__ __forceinline__ double conv<double, double>(double v, int64_t) {
    return v;
}
template <>
__device__ __forceinline__ uint32_t conv<int32_t, uint32_t>(int32_t v, int64_t gmin) {
    return (uint32_t)((int64_t)v - gmin);
}

// Two consecutive rows (one vector load when aligned and complete).  Lane l
// of a wavefront loads rows 2l, 2l + 1 of a 128-row stretch, so each load
// instruction covers whole 128-byte lines: with four or eight consecutive
// rows per lane, each instruction touched every line of the stretch in part,
// and the streaming (non-temporal) loads fetched some lines twice.
template <typename V>
__device__ __forceinline__ void load2(const V* __restrict__ p, uint64_t row, uint64_t end, bool vec, V* out) {
    if (vec && row + 2 <= end) {
        if constexpr (sizeof(V) == 8) {
            const uint4 a = nt_load((const uint4*)(p + row));
            memcpy(out, &a, 16);
        } else {
            const uint2 a = nt_load((const uint2*)(p + row));
            memcpy(out, &a, 8);
        }
        return;
    }
    out[0] = row < end ? p[row] : V();
    out[1] = row + 1 < end ? p[row + 1] : V();
}

// Scatter of one tile of kRangeTile rows into key-range partitions: rows
// (local key, conv(payload)) at exact offsets base[p] + H[chunk][p] +
// toff[tile][p] (H column-scanned), reordered by partition in LDS and written
// as contiguous runs.  Rows whose key is out of range are dropped.  Tiles t
// and t + 1 write adjacent runs of every partition: neighbours run side by
// side on one XCD (chunks of 32 tiles), so the L2 merges the lines they share.
template <typename VIn, typename VOut, int TR>
__global__ __launch_bounds__(kRT, 4) void k_rscatter(const int32_t* __restrict__ key, const VIn* __restrict__ pay,
                                                      uint64_t n, int64_t kmin, uint64_t range, int64_t gmin,
                                                      uint32_t Ppad, const uint32_t* __restrict__ H,
                                                      const uint16_t* __restrict__ toff,
                                                      const uint32_t* __restrict__ base, uint16_t* __restrict__ olk,
                                                      VOut* __restrict__ oval) {
    constexpr int PR = TR / kRT, PS = kP1Max / kRT;
    static_assert(TR == kRangeTile && PR % 2 == 0, "tile shape");
    __shared__ uint16_t slk[TR], spart[TR];
    __shared__ VOut sval[TR];
    __shared__ uint32_t cnt[kP1Max], start[kP1Max], cur[kP1Max];
    __shared__ uint32_t ws[kRT / 64 + 1];
    const uint32_t t = pgpu::dev::xcd_remap_chunked<32>(blockIdx.x, gridDim.x);
    const uint64_t t0 = (uint64_t)t * TR, te = min<uint64_t>(n, t0 + TR);
    const bool vec = ((uintptr_t)key & 7) == 0 && ((uintptr_t)pay & (2 * sizeof(VIn) - 1)) == 0;
    int32_t k[PR];
    VIn v[PR];
    uint64_t rows[PR / 2];
#pragma unroll
    for (int j = 0; j < PR / 2; ++j) {
        rows[j] = t0 + ((uint64_t)j * kRT + threadIdx.x) * 2;
        load2(key, rows[j], te, vec, k + 2 * j);
        load2(pay, rows[j], te, vec, v + 2 * j);
    }
    const uint32_t* h = H + (uint64_t)(t / kTilesPerChunk) * Ppad;
    const uint16_t* to = toff + (uint64_t)t * Ppad;
    for (uint32_t d = threadIdx.x; d < Ppad; d += kRT) {
        cur[d] = base[d] + h[d] + to[d];
        cnt[d] = 0;
    }
    __syncthreads();
    uint32_t part[PR], rk[PR];
    bool ok[PR];
#pragma unroll
    for (int u = 0; u < PR; ++u) {
        const uint64_t off = (uint64_t)((int64_t)k[u] - kmin);
        ok[u] = rows[u / 2] + (u & 1) < te && off < range;
        part[u] = (uint32_t)(off >> kSK);
        k[u] = (int32_t)(off & ((1u << kSK) - 1));
        if (ok[u]) rk[u] = atomicAdd(&cnt[part[u]], 1u);
    }
    __syncthreads();
    uint32_t loc[PS], sum = 0;
#pragma unroll
    for (int u = 0; u < PS; ++u) {
        const uint32_t idx = threadIdx.x * PS + u;
        loc[u] = idx < Ppad ? cnt[idx] : 0u;
        sum += loc[u];
    }
    uint32_t total;
    uint32_t ex = block_excl_scan<kRT>(sum, ws, &total);
#pragma unroll
    for (int u = 0; u < PS; ++u) {
        const uint32_t idx = threadIdx.x * PS + u;
        if (idx < Ppad) start[idx] = ex;
        ex += loc[u];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PR; ++u) {
        if (ok[u]) {
            const uint32_t pos = start[part[u]] + rk[u];
            slk[pos] = (uint16_t)k[u];
            spart[pos] = (uint16_t)part[u];
            sval[pos] = conv<VIn, VOut>(v[u], gmin);
        }
    }
    __syncthreads();
    for (uint32_t pos = threadIdx.x; pos < total; pos += kRT) {
        const uint32_t p = spart[pos];
        const uint32_t idx = cur[p] + (pos - start[p]);
        olk[idx] = slk[pos];
        oval[idx] = sval[pos];
    }
}

// The slice table of key-range partition p in LDS: local key -> gid from
// D's partition rows [b0, b1).  Equal keys keep the smallest gid, so both
// passes build identical tables even when a duplicate key makes the result
// void; *dup (pass 1 only) records the duplicate.
__device__ __forceinline__ void build_slice(uint32_t* tab, const uint16_t* __restrict__ dlk,
                                            const uint32_t* __restrict__ dgid, uint32_t b0, uint32_t b1,
                                            int32_t* __restrict__ dup) {
    for (uint32_t i = threadIdx.x; i < (1u << kSK); i += kT) tab[i] = kNone;
    __syncthreads();
    bool seen = false;
    constexpr int U = 8;  // rows per thread per step, all loads issued before the LDS updates
    for (uint32_t r0 = b0; r0 < b1; r0 += U * kT) {
        uint32_t lk[U], g[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t r = r0 + u * kT + threadIdx.x;
            lk[u] = r < b1 ? dlk[r] : 0u;
            g[u] = r < b1 ? dgid[r] : kNone;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (r0 + u * kT + threadIdx.x < b1) seen |= atomicMin(&tab[lk[u]], g[u]) != kNone;
    }
    if (dup && seen) *dup = 1;
    __syncthreads();
}

// Pass 1 over one probe item: every row's gid (kNone: no match) into fgid,
// H2[item][q] = rows of the item whose group lies in group-range partition
// q = gid >> kSG, and toff2[s][q] = the item's rows of range q before its
// half-tile s (kGroupTile / 2 rows; half-tiles 2t and 2t + 1 make up tile t,
// counted globally, so toff2[2t] holds tile t's offsets).  Each wavefront
// looks up whole half-tiles (w, w + 16, ... of the item) and counts them in
// its own LDS row (two 16-bit counters per word: a half-tile holds at most
// 3072 rows).  Half-tiles spread an item over the waves more evenly than
// tiles: an item of 22 tiles kept 6 waves busy with a second tile while the
// other 10 waited.
__global__ __launch_bounds__(kT) void k_phist(const Item* __restrict__ items, const int32_t* __restrict__ mm,
                                              const uint32_t* __restrict__ bD, const uint16_t* __restrict__ dlk,
                                              const uint32_t* __restrict__ dgid, const uint16_t* __restrict__ flk,
                                              uint32_t P2pad, uint32_t* __restrict__ H2, uint32_t* __restrict__ fgid,
                                              int32_t* __restrict__ dup, const uint32_t* __restrict__ tile0,
                                              uint32_t* __restrict__ toff2) {
    constexpr uint32_t kHalf = kGroupTile / 2;
    __shared__ uint32_t tab[1u << kSK];
    __shared__ uint32_t lt[kT / 64][kP2Max / 2];
    uint32_t* out = H2 + (uint64_t)blockIdx.x * P2pad;
    if ((int32_t)blockIdx.x >= mm[kMmItems1]) {  // past the item count: a zero row for the column scan
        for (uint32_t i = threadIdx.x; i < P2pad; i += kT) out[i] = 0;
        return;
    }
    const Item it = items[blockIdx.x];
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (uint32_t i = lane; i < kP2Max / 2; i += 64) lt[w][i] = 0;
    build_slice(tab, dlk, dgid, bD[it.part], bD[it.part + 1], dup);
    const uint32_t ns = 2 * ((it.r1 - it.r0 + kGroupTile - 1) / kGroupTile);  // half-tiles, the last may be empty
    uint32_t* const tcb = toff2 + (uint64_t)(2 * tile0[blockIdx.x]) * P2pad;
    constexpr int U = 16;  // rows per lane per step, all loads issued before the lookups
    for (uint32_t k = w; k < ns; k += kT / 64) {
        const uint32_t rs = it.r0 + k * kHalf, re = min(it.r1, rs + kHalf);
        for (uint32_t t0 = rs; t0 < re; t0 += U * 64) {
            uint32_t lk[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t r = t0 + u * 64 + lane;
                lk[u] = r < re ? flk[r] : kNone;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (lk[u] != kNone) {
                    const uint32_t g = tab[lk[u]];
                    if (g != kNone) {
                        const uint32_t q = g >> kSG;
                        atomicAdd(&lt[w][q >> 1], 1u << (16 * (q & 1)));
                    }
                    fgid[t0 + u * 64 + lane] = g;
                }
            }
        }
        // one wavefront's LDS operations complete in order: its counts are final
        uint32_t* tc = tcb + (uint64_t)k * P2pad;
        for (uint32_t q = lane; q < P2pad; q += 64) tc[q] = (lt[w][q >> 1] >> (16 * (q & 1))) & 0xFFFFu;
        for (uint32_t i = lane; i < kP2Max / 2; i += 64) lt[w][i] = 0;
    }
    __syncthreads();  // every half-tile's counts are in toff2
    constexpr int B = 8;  // half-tiles per step: their counts are loaded together
    for (uint32_t q = threadIdx.x; q < P2pad; q += kT) {
        uint32_t run = 0;
        for (uint32_t k0 = 0; k0 < ns; k0 += B) {
            uint32_t c[B];
#pragma unroll
            for (int b = 0; b < B; ++b) c[b] = k0 + b < ns ? tcb[(uint64_t)(k0 + b) * P2pad + q] : 0u;
#pragma unroll
            for (int b = 0; b < B; ++b) {
                if (k0 + b < ns) tcb[(uint64_t)(k0 + b) * P2pad + q] = run;
                run += c[b];
            }
        }
        out[q] = run;
    }
}

// Pass 2 over one tile of a probe item: (gid & 8191, value) rows into
// group-range partitions at base2[q] + H2[item][q] + toff2[2 tile][q] (H2
// column-scanned).  Tiles t and t + 1 write adjacent runs of every range:
// neighbours run side by side on one XCD (chunks of 32 tiles).
template <int TR>
__global__ __launch_bounds__(kPST, 4) void k_pscatter(const Item* __restrict__ items, const int32_t* __restrict__ mm,
                                                      const uint32_t* __restrict__ titem,
                                                      const uint32_t* __restrict__ fgid, const double* __restrict__ fval,
                                                      uint32_t P2pad, const uint32_t* __restrict__ H2,
                                                      const uint32_t* __restrict__ toff2,
                                                      const uint32_t* __restrict__ base2, uint16_t* __restrict__ opg,
                                                      double* __restrict__ opv) {
    constexpr int PR = TR / kPST;
    static_assert(TR % kPST == 0 && kP2Max <= kPST, "tile shape; one thread per range in the scan");
    __shared__ uint16_t sg[TR], sq[TR];
    __shared__ double sv[TR];
    __shared__ uint32_t cnt[kP2Max], start[kP2Max], cur[kP2Max];
    __shared__ uint32_t ws[kPST / 64 + 1];
    const uint32_t t = pgpu::dev::xcd_remap_chunked<32>(blockIdx.x, gridDim.x);
    if ((int32_t)t >= mm[kMmTiles1]) return;
    const uint32_t ik = titem[t], item = ik >> 8;
    const Item it = items[item];
    const uint32_t rs = it.r0 + (ik & 0xFFu) * TR, re = min(it.r1, rs + TR);
    uint32_t g[PR];
    double v[PR];
#pragma unroll
    for (int u = 0; u < PR; ++u) {
        const uint32_t r = rs + u * kPST + threadIdx.x;
        const bool ok = r < re;
        g[u] = ok ? fgid[r] : kNone;
        v[u] = ok ? fval[r] : 0.0;
    }
    const uint32_t* h = H2 + (uint64_t)item * P2pad;
    const uint32_t* to = toff2 + (uint64_t)(2 * t) * P2pad;  // the tile's first half-tile
    for (uint32_t d = threadIdx.x; d < P2pad; d += kPST) {
        cur[d] = base2[d] + h[d] + to[d];
        cnt[d] = 0;
    }
    __syncthreads();
    uint32_t rk[PR];
#pragma unroll
    for (int u = 0; u < PR; ++u)
        if (g[u] != kNone) rk[u] = atomicAdd(&cnt[g[u] >> kSG], 1u);
    __syncthreads();
    const uint32_t loc = threadIdx.x < P2pad ? cnt[threadIdx.x] : 0u;
    uint32_t total;
    const uint32_t ex = block_excl_scan<kPST>(loc, ws, &total);
    if (threadIdx.x < P2pad) start[threadIdx.x] = ex;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PR; ++u) {
        if (g[u] != kNone) {
            const uint32_t q = g[u] >> kSG;
            const uint32_t pos = start[q] + rk[u];
            sg[pos] = (uint16_t)(g[u] & ((1u << kSG) - 1));
            sq[pos] = (uint16_t)q;
            sv[pos] = v[u];
        }
    }
    __syncthreads();
    for (uint32_t pos = threadIdx.x; pos < total; pos += kPST) {
        const uint32_t q = sq[pos];
        const uint32_t idx = cur[q] + (pos - start[q]);
        opg[idx] = sg[pos];
        opv[idx] = sv[pos];
    }
}

// One aggregation item: rows [r0, r1) of a group-range partition summed into
// LDS tables of 8192 groups, written as the item's partial table.
__global__ __launch_bounds__(kT) void k_gagg(const Item* __restrict__ items, const int32_t* __restrict__ mm,
                                             const uint16_t* __restrict__ pg, const double* __restrict__ pv,
                                             double* __restrict__ psum, uint32_t* __restrict__ pcnt) {
    constexpr uint32_t NG = 1u << kSG;
    __shared__ double ls[NG];
    __shared__ uint32_t lc[NG];
    if ((int32_t)blockIdx.x >= mm[kMmItems2]) return;
    const Item it = items[blockIdx.x];
    for (uint32_t i = threadIdx.x; i < NG; i += kT) {
        ls[i] = 0.0;
        lc[i] = 0;
    }
    __syncthreads();
    constexpr int U = 8;  // rows per thread per step, all loads issued before the LDS adds
    for (uint32_t t0 = it.r0; t0 < it.r1; t0 += U * kT) {
        uint32_t g[U];
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t r = t0 + u * kT + threadIdx.x;
            g[u] = r < it.r1 ? pg[r] : kNone;
            v[u] = r < it.r1 ? pv[r] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (g[u] != kNone) {
                atomicAdd(&ls[g[u]], v[u]);
                atomicAdd(&lc[g[u]], 1u);
            }
        }
    }
    __syncthreads();
    double* os = psum + (uint64_t)blockIdx.x * NG;
    uint32_t* oc = pcnt + (uint64_t)blockIdx.x * NG;
    for (uint32_t i = threadIdx.x; i < NG; i += kT) {
        os[i] = ls[i];
        oc[i] = lc[i];
    }
}

// Group gid = sum of its partition's partial tables; non-empty groups are
// written out (positions past out_cap counted, not written).  kCompactSplit
// blocks per group-range partition, each over a slice of its groups: one
// cursor atomic and one joined-count atomic each.
constexpr uint32_t kCompactSplit = 4;
__global__ __launch_bounds__(kT) void k_gcompact(const uint32_t* __restrict__ first2, uint64_t Rg, int64_t gmin,
                                                 const double* __restrict__ psum, const uint32_t* __restrict__ pcnt,
                                                 int32_t* __restrict__ okey, double* __restrict__ osum,
                                                 uint64_t* __restrict__ ocnt, uint64_t out_cap,
                                                 unsigned long long* __restrict__ ct) {
    constexpr uint32_t NG = 1u << kSG, NS = NG / kCompactSplit;
    constexpr int U = NS / kT;
    static_assert(NS % kT == 0, "whole groups per thread");
    __shared__ unsigned long long base;
    __shared__ uint32_t ws[kT / 64 + 1];
    __shared__ unsigned long long wj[kT / 64];
    const uint32_t q = blockIdx.x / kCompactSplit;
    const uint32_t i0 = first2[q], i1 = first2[q + 1];
    const uint64_t g0 = ((uint64_t)q << kSG) + (uint64_t)(blockIdx.x % kCompactSplit) * NS;
    const uint32_t s0 = (blockIdx.x % kCompactSplit) * NS;  // first slot of this block in a partial table
    uint32_t c[U];
    double sm[U];
    uint32_t live = 0;
    unsigned long long joined = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        c[u] = 0;
        sm[u] = 0.0;
    }
    constexpr int B = 4;  // partial tables read per step, all loads issued before the adds
    for (uint32_t i = i0; i < i1; i += B) {
        uint32_t cc[B][U];
        double ss[B][U];
#pragma unroll
        for (int b = 0; b < B; ++b)
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t sl = s0 + u * kT + threadIdx.x;
                const bool ok = i + b < i1 && g0 + u * kT + threadIdx.x < Rg;
                cc[b][u] = ok ? pcnt[(uint64_t)(i + b) * NG + sl] : 0u;
                ss[b][u] = ok ? psum[(uint64_t)(i + b) * NG + sl] : 0.0;
            }
#pragma unroll
        for (int b = 0; b < B; ++b)
#pragma unroll
            for (int u = 0; u < U; ++u) {
                c[u] += cc[b][u];
                sm[u] += ss[b][u];
            }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        live += c[u] > 0;
        joined += c[u];
    }
    uint32_t total;
    uint32_t off = block_excl_scan<kT>(live, ws, &total);
    joined = pgpu::dev::wave_sum(joined);
    if ((threadIdx.x & 63) == 0) wj[threadIdx.x >> 6] = joined;
    if (threadIdx.x == 0) base = total ? atomicAdd(&ct[kCtCursor], (unsigned long long)total) : 0ull;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long j = 0;
        for (int w = 0; w < kT / 64; ++w) j += wj[w];
        if (j) atomicAdd(&ct[kCtJoined], j);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (c[u] > 0) {
            const uint64_t o = base + off++;
            if (o < out_cap) {
                okey[o] = (int32_t)(gmin + (int64_t)(g0 + u * kT + threadIdx.x));
                osum[o] = sm[u];
                ocnt[o] = c[u];
            }
        }
    }
}

inline uint32_t pad64(uint64_t p) { return (uint32_t)((p + 63) & ~63ull); }

}  // namespace part

}  // namespace

// Output rows of an aggregate over the group join: the int32 group key then
// an 8-byte accumulator (the FP64 sum, or the int64 count), packed into 12
// bytes with no padding (schema.cpp:82-117), for result rows [first, first + n).
__global__ __launch_bounds__(256) void k_gj_rows(const int32_t* __restrict__ okey, const double* __restrict__ osum,
                                                 const uint64_t* __restrict__ ocnt, int count, uint64_t first,
                                                 uint64_t n, uint32_t* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
        const uint64_t a = count ? ocnt[first + i] : (uint64_t)__double_as_longlong(osum[first + i]);
        uint32_t* r = out + 3 * i;
        r[0] = (uint32_t)okey[first + i];
        r[1] = (uint32_t)a;
        r[2] = (uint32_t)(a >> 32);
    }
}

// As k_gj_rows, written straight into pinned host memory (dst 16-byte
// aligned): a block's 256 rows are packed in LDS and leave as 16-byte stores,
// so every write that crosses the host link is a whole aligned chunk.
__global__ __launch_bounds__(256) void k_gj_rows_host(const int32_t* __restrict__ okey,
                                                      const double* __restrict__ osum,
                                                      const uint64_t* __restrict__ ocnt, int count, uint64_t first,
                                                      uint64_t n, uint4* __restrict__ dst) {
    __shared__ __align__(16) uint32_t buf[256 * 3];
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t b0 = (uint64_t)blockIdx.x * 256; b0 < n; b0 += stride) {
        const uint64_t i = b0 + threadIdx.x;
        if (i < n) {
            const uint64_t a = count ? ocnt[first + i] : (uint64_t)__double_as_longlong(osum[first + i]);
            buf[3 * threadIdx.x] = (uint32_t)okey[first + i];
            buf[3 * threadIdx.x + 1] = (uint32_t)a;
            buf[3 * threadIdx.x + 2] = (uint32_t)(a >> 32);
        }
        __syncthreads();
        const uint32_t words = (uint32_t)min<uint64_t>(256, n - b0) * 3;
        const uint32_t vecs = words / 4;
        uint4* o = dst + b0 * 3 / 4;  // 768 words per block: b0 * 3 is a multiple of 4
        if (threadIdx.x < vecs) o[threadIdx.x] = reinterpret_cast<const uint4*>(buf)[threadIdx.x];
        if (threadIdx.x < words - vecs * 4)
            reinterpret_cast<uint32_t*>(o + vecs)[threadIdx.x] = buf[vecs * 4 + threadIdx.x];
        __syncthreads();
    }
}

struct pgpu_groupjoin {
    pgpu_ctx* ctx = nullptr;
    uint64_t maxD = 0, maxF = 0, maxG = 0;
    // general path
    DSlot* dt = nullptr;
    uint32_t dmask = 0;
    GSlot* gt = nullptr;
    uint32_t gmask = 0;
    unsigned long long* counters = nullptr;  // kCt* words
    unsigned long long* h_counters = nullptr;
    int32_t* special_grp = nullptr;
    // outputs (maxG + 1 groups: the marker group of the general path fits too)
    int32_t* okey = nullptr;
    double* osum = nullptr;
    uint64_t* ocnt = nullptr;
    uint32_t* orow = nullptr;  // packed (key, accumulator) rows for pgpu_groupjoin_rows
    // partitioned path (absent when the plan's fact rows reach 2^31)
    bool part_ok = false;
    int32_t* mm = nullptr;
    int32_t* mmpart = nullptr;
    int32_t* h_mm = nullptr;
    uint32_t* H = nullptr;
    uint16_t* toff = nullptr;  // per scatter tile and key-range partition: rows in the chunk's earlier tiles
    uint32_t *tot = nullptr, *baseD = nullptr, *baseF = nullptr, *base2 = nullptr;
    uint32_t *first1 = nullptr, *first2 = nullptr;
    part::Item *items1 = nullptr, *items2 = nullptr;
    uint32_t *tile0 = nullptr, *titem = nullptr, *toff2 = nullptr;  // group-range scatter tiles of the probe items
    uint64_t tiles1_cap = 0;
    uint64_t items1_cap = 0, items2_cap = 0;
    uint16_t *dlk = nullptr, *flk = nullptr, *pg = nullptr;
    uint32_t *dgid = nullptr, *fgid = nullptr;
    double *fv = nullptr, *pv = nullptr, *psum = nullptr;
    uint32_t* pcnt = nullptr;
    int last_path = 0;  // 1 partitioned, 2 general
    uint64_t ngroups = 0, joined = 0;
    pgpu::KernelTimer* timer = nullptr;
    std::vector<void*> owned;
};

using namespace pgpu;

namespace {
int gj_alloc(pgpu_groupjoin* p, uint64_t bytes, void** out) {
    PGPU_TRY(ctx_alloc(p->ctx, std::max<uint64_t>(bytes, 256), out));
    p->owned.push_back(*out);
    return PGPU_OK;
}
struct Timed {
    KernelTimer* t;
    int idx;
    Timed(KernelTimer* t_, const char* n) : t(t_), idx(t_->start(n)) {}
    ~Timed() { t->stop(idx); }
};

int alloc_partitioned(pgpu_groupjoin* p) {
    using namespace part;
    const uint64_t chunksD = ceil_div(p->maxD, kCR), chunksF = ceil_div(p->maxF, kCR);
    p->items1_cap = kP1Max + ceil_div(p->maxF, kIR1);
    p->items2_cap = kP2Max + ceil_div(p->maxF, kIR2);
    const uint64_t hwords =
        std::max<uint64_t>({chunksD * kP1Max, chunksF * kP1Max, p->items1_cap * kP2Max});
    const uint64_t NG = 1ull << kSG;
    PGPU_TRY(gj_alloc(p, kMmWords * sizeof(int32_t), (void**)&p->mm));
    PGPU_TRY(gj_alloc(p, kMinmaxBlocks * 4 * sizeof(int32_t), (void**)&p->mmpart));
    PGPU_TRY(gj_alloc(p, hwords * 4, (void**)&p->H));
    PGPU_TRY(gj_alloc(p, (std::max(chunksD, chunksF) + 1) * kTilesPerChunk * kP1Max * 2, (void**)&p->toff));
    PGPU_TRY(gj_alloc(p, kP1Max * 4, (void**)&p->tot));
    PGPU_TRY(gj_alloc(p, (kP1Max + 1) * 4, (void**)&p->baseD));
    PGPU_TRY(gj_alloc(p, (kP1Max + 1) * 4, (void**)&p->baseF));
    PGPU_TRY(gj_alloc(p, (kP2Max + 1) * 4, (void**)&p->base2));
    PGPU_TRY(gj_alloc(p, (kP1Max + 1) * 4, (void**)&p->first1));
    PGPU_TRY(gj_alloc(p, (kP2Max + 1) * 4, (void**)&p->first2));
    PGPU_TRY(gj_alloc(p, p->items1_cap * sizeof(Item), (void**)&p->items1));
    PGPU_TRY(gj_alloc(p, p->items2_cap * sizeof(Item), (void**)&p->items2));
    p->tiles1_cap = ceil_div(p->maxF, kGroupTile) + kP1Max + 1;  // whole tiles + one partial tile per partition
    PGPU_TRY(gj_alloc(p, p->items1_cap * 4, (void**)&p->tile0));
    PGPU_TRY(gj_alloc(p, p->tiles1_cap * 4, (void**)&p->titem));
    PGPU_TRY(gj_alloc(p, 2 * p->tiles1_cap * kP2Max * 4, (void**)&p->toff2));  // two half-tiles per tile
    PGPU_TRY(gj_alloc(p, p->maxD * 2, (void**)&p->dlk));
    PGPU_TRY(gj_alloc(p, p->maxD * 4, (void**)&p->dgid));
    PGPU_TRY(gj_alloc(p, p->maxF * 2, (void**)&p->flk));
    PGPU_TRY(gj_alloc(p, p->maxF * 4, (void**)&p->fgid));
    PGPU_TRY(gj_alloc(p, p->maxF * 8, (void**)&p->fv));
    PGPU_TRY(gj_alloc(p, p->maxF * 2, (void**)&p->pg));
    PGPU_TRY(gj_alloc(p, p->maxF * 8, (void**)&p->pv));
    PGPU_TRY(gj_alloc(p, p->items2_cap * NG * 8, (void**)&p->psum));
    PGPU_TRY(gj_alloc(p, p->items2_cap * NG * 4, (void**)&p->pcnt));
    if (hipHostMalloc((void**)&p->h_mm, kMmWords * sizeof(int32_t), 0) != hipSuccess)
        return fail(PGPU_ENOMEM, "hipHostMalloc failed");
    p->part_ok = true;
    return PGPU_OK;
}

// Key-range partitioning of one side (hist, column scan, total scan,
// scatter).  `base` receives the Ppad + 1 partition bounds.
template <typename VIn, typename VOut, int TR>
int partition_side(pgpu_groupjoin* p, const char* name, const int32_t* key, const VIn* pay, uint64_t n, int64_t kmin,
                   uint64_t range, int64_t gmin, uint32_t Ppad, uint32_t* base, uint16_t* olk, VOut* oval) {
    using namespace part;
    hipStream_t st = p->ctx->stream;
    const uint32_t chunks = (uint32_t)std::max<uint64_t>(1, ceil_div(n, kCR));
    char nm[32];
    {
        snprintf(nm, sizeof nm, "%s_hist", name);
        Timed t(p->timer, nm);
        hipLaunchKernelGGL(k_rhist, dim3(chunks), dim3(512), 0, st, key, n, kmin, range, Ppad, p->H, p->toff);
        PGPU_LAUNCH_CHECK();
    }
    {
        snprintf(nm, sizeof nm, "%s_scan", name);
        Timed t(p->timer, nm);
        hipLaunchKernelGGL(k_cscan, dim3(Ppad / kScanCols), dim3(1024), 0, st, p->H, chunks, Ppad, p->tot);
        hipLaunchKernelGGL(k_tscan, dim3(1), dim3(1024), 0, st, p->tot, Ppad, base);
        PGPU_LAUNCH_CHECK();
    }
    {
        snprintf(nm, sizeof nm, "%s_scatter", name);
        Timed t(p->timer, nm);
        const uint32_t tiles = (uint32_t)ceil_div(n, (uint64_t)TR);
        if (tiles) {
            hipLaunchKernelGGL((k_rscatter<VIn, VOut, TR>), dim3(tiles), dim3(kRT), 0, st, key, pay, n, kmin, range,
                               gmin, Ppad, p->H, p->toff, base, olk, oval);
            PGPU_LAUNCH_CHECK();
        }
    }
    return PGPU_OK;
}

// The partitioned path; the caller has checked the ranges.  Leaves the
// duplicate flag and the counters in pinned memory after a sync.
int run_partitioned(pgpu_groupjoin* p, const int32_t* dkey, const int32_t* dgrp, uint64_t nD, const int32_t* fkey,
                    const double* fval, uint64_t nF, int64_t kmin, uint64_t Rk, int64_t gmin, uint64_t Rg) {
    using namespace part;
    hipStream_t st = p->ctx->stream;
    const uint32_t P1 = (uint32_t)ceil_div(Rk, 1ull << kSK), P1pad = pad64(P1);
    const uint32_t P2 = (uint32_t)ceil_div(Rg, 1ull << kSG), P2pad = pad64(P2);
    const uint32_t it1 = (uint32_t)(P1 + ceil_div(nF, kIR1)), it2 = (uint32_t)(P2 + ceil_div(nF, kIR2));
    if (P1pad > kP1Max || P2pad > kP2Max || it1 > p->items1_cap || it2 > p->items2_cap)
        return fail(PGPU_EOVERFLOW, "groupjoin: partitioned shape exceeds its allocation");
    int rc = partition_side<int32_t, uint32_t, kRangeTile>(p, "gj_dim", dkey, dgrp, nD, kmin, Rk, gmin, P1pad, p->baseD, p->dlk,
                                               p->dgid);
    if (rc == PGPU_OK)
        rc = partition_side<double, double, kRangeTile>(p, "gj_fact", fkey, fval, nF, kmin, Rk, 0, P1pad, p->baseF, p->flk, p->fv);
    PGPU_TRY(rc);
    {
        Timed t(p->timer, "gj_probe_hist");
        hipLaunchKernelGGL(k_items, dim3(1), dim3(1024), 0, st, p->baseF, P1, kIR1, p->items1, p->first1,
                           p->mm + kMmItems1, (uint32_t)kGroupTile, p->tile0, p->titem, p->mm + kMmTiles1);
        hipLaunchKernelGGL(k_phist, dim3(it1), dim3(kT), 0, st, p->items1, p->mm, p->baseD, p->dlk, p->dgid, p->flk,
                           P2pad, p->H, p->fgid, p->mm + kMmDup, p->tile0, p->toff2);
        PGPU_LAUNCH_CHECK();
    }
    {
        Timed t(p->timer, "gj_probe_scan");
        hipLaunchKernelGGL(k_cscan, dim3(P2pad / kScanCols), dim3(1024), 0, st, p->H, it1, P2pad, p->tot);
        hipLaunchKernelGGL(k_tscan, dim3(1), dim3(1024), 0, st, p->tot, P2pad, p->base2);
        hipLaunchKernelGGL(k_items, dim3(1), dim3(1024), 0, st, p->base2, P2, kIR2, p->items2, p->first2,
                           p->mm + kMmItems2, 0u, (uint32_t*)nullptr, (uint32_t*)nullptr, (int32_t*)nullptr);
        PGPU_LAUNCH_CHECK();
    }
    {
        Timed t(p->timer, "gj_probe_scatter");
        // tiles: at most the whole tiles of the fact rows plus one partial tile per partition
        const uint32_t tiles = (uint32_t)(ceil_div(nF, (uint64_t)kGroupTile) + P1);
        if (tiles > p->tiles1_cap) return fail(PGPU_EOVERFLOW, "groupjoin: tiles exceed the allocation");
        hipLaunchKernelGGL(k_pscatter<kGroupTile>, dim3(tiles), dim3(kPST), 0, st, p->items1, p->mm, p->titem,
                           p->fgid, p->fv, P2pad, p->H, p->toff2, p->base2, p->pg, p->pv);
        PGPU_LAUNCH_CHECK();
    }
    {
        Timed t(p->timer, "gj_agg");
        hipLaunchKernelGGL(k_gagg, dim3(it2), dim3(kT), 0, st, p->items2, p->mm, p->pg, p->pv, p->psum, p->pcnt);
        PGPU_LAUNCH_CHECK();
    }
    {
        Timed t(p->timer, "gj_compact");
        hipLaunchKernelGGL(k_gcompact, dim3(P2 * kCompactSplit), dim3(kT), 0, st, p->first2, Rg, gmin, p->psum,
                           p->pcnt, p->okey, p->osum, p->ocnt, p->maxG + 1, p->counters);
        PGPU_LAUNCH_CHECK();
    }
    PGPU_HIP(hipMemcpyAsync(p->h_counters, p->counters, kCtWords * 8, hipMemcpyDeviceToHost, st));
    PGPU_HIP(hipMemcpyAsync(p->h_mm, p->mm, kMmWords * 4, hipMemcpyDeviceToHost, st));
    PGPU_HIP(hipStreamSynchronize(st));
    return PGPU_OK;
}

int run_general(pgpu_groupjoin* p, const int32_t* d_dkey, const int32_t* d_dgrp, uint64_t nD, const int32_t* d_fkey,
                const double* d_fval, uint64_t nF) {
    hipStream_t st = p->ctx->stream;
    const uint64_t dcap = (uint64_t)p->dmask + 1, gcap = (uint64_t)p->gmask + 1;
    {
        Timed t(p->timer, "gj_build");
        hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, st, p->counters, kCtWords);
        hipLaunchKernelGGL(k_dim_clear, dim3(grid_for(dcap)), dim3(256), 0, st, p->dt, dcap);
        hipLaunchKernelGGL(k_grp_clear, dim3(grid_for(gcap)), dim3(256), 0, st, p->gt, gcap);
        if (nD)
            hipLaunchKernelGGL(k_dim_build, dim3(grid_for(nD)), dim3(256), 0, st, d_dkey, d_dgrp, nD, p->dt, p->dmask,
                               p->counters + kCtDimSpecial, p->special_grp, kSpecialCap);
        PGPU_LAUNCH_CHECK();
    }
    if (nF && nD) {
        Timed t(p->timer, "gj_probe");
        hipLaunchKernelGGL(k_groupjoin, dim3(grid_for(nF, 2048)), dim3(kGJThreads), 0, st, d_fkey, d_fval, nF, p->dt,
                           p->dmask, p->special_grp, kSpecialCap, p->gt, p->gmask, p->counters);
        PGPU_LAUNCH_CHECK();
    }
    {
        Timed t(p->timer, "gj_compact");
        hipLaunchKernelGGL(k_grp_compact, dim3(grid_for(gcap)), dim3(256), 0, st, p->gt, gcap, p->counters, p->okey,
                           p->osum, p->ocnt, p->maxG + 1);
        PGPU_LAUNCH_CHECK();
    }
    PGPU_HIP(hipMemcpyAsync(p->h_counters, p->counters, kCtWords * 8, hipMemcpyDeviceToHost, st));
    PGPU_HIP(hipStreamSynchronize(st));
    if (p->h_counters[kCtDimSpecial] > kSpecialCap)
        return fail(PGPU_EOVERFLOW, "more than %u dimension rows carry the key INT32_MIN", kSpecialCap);
    if (p->h_counters[kCtFull])
        return fail(PGPU_EOVERFLOW, "groupjoin: more groups than the plan's group table holds (max_groups %llu)",
                    (unsigned long long)p->maxG);
    return PGPU_OK;
}
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
    if (rc == PGPU_OK) rc = gj_alloc(p, kCtWords * sizeof(unsigned long long), (void**)&p->counters);
    if (rc == PGPU_OK) rc = gj_alloc(p, kSpecialCap * sizeof(int32_t), (void**)&p->special_grp);
    if (rc == PGPU_OK) rc = gj_alloc(p, (p->maxG + 1) * sizeof(int32_t), (void**)&p->okey);
    if (rc == PGPU_OK) rc = gj_alloc(p, (p->maxG + 1) * sizeof(double), (void**)&p->osum);
    if (rc == PGPU_OK) rc = gj_alloc(p, (p->maxG + 1) * sizeof(uint64_t), (void**)&p->ocnt);
    if (rc == PGPU_OK) rc = gj_alloc(p, (p->maxG + 1) * 12, (void**)&p->orow);
    if (rc == PGPU_OK && hipHostMalloc((void**)&p->h_counters, kCtWords * sizeof(unsigned long long), 0) != hipSuccess)
        rc = fail(PGPU_ENOMEM, "hipHostMalloc failed");
    if (rc == PGPU_OK && max_fact < (1ull << 31)) rc = alloc_partitioned(p);
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
    if (p->h_mm) (void)hipHostFree(p->h_mm);
    delete p;
    return PGPU_OK;
}

int pgpu_groupjoin_run(pgpu_groupjoin* p, const int32_t* d_dkey, const int32_t* d_dgrp, uint64_t nD,
                       const int32_t* d_fkey, const double* d_fval, uint64_t nF, uint64_t* n_groups,
                       uint64_t* n_joined) {
    using namespace part;
    if (!p || !n_groups) return fail(PGPU_EINVAL, "pgpu_groupjoin_run: null argument");
    if (nD > p->maxD || nF > p->maxF) return fail(PGPU_EINVAL, "pgpu_groupjoin_run: inputs exceed the plan");
    if ((nD && (!d_dkey || !d_dgrp)) || (nF && (!d_fkey || !d_fval)))
        return fail(PGPU_EINVAL, "pgpu_groupjoin_run: null column");
    PGPU_TRY(ctx_activate(p->ctx));
    hipStream_t st = p->ctx->stream;
    bool done = false;
    if (p->part_ok && nD && nF) {
        {
            Timed t(p->timer, "gj_minmax");
            hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, st, p->counters, kCtWords);
            hipLaunchKernelGGL(k_mm_init, dim3(1), dim3(64), 0, st, p->mm);
            hipLaunchKernelGGL(k_minmax, dim3(kMinmaxBlocks), dim3(kT), 0, st, d_dkey, d_dgrp, nD, p->mmpart);
            hipLaunchKernelGGL(k_minmax2, dim3(1), dim3(kT), 0, st, p->mmpart, (uint32_t)kMinmaxBlocks, p->mm);
            PGPU_LAUNCH_CHECK();
        }
        PGPU_HIP(hipMemcpyAsync(p->h_mm, p->mm, kMmWords * 4, hipMemcpyDeviceToHost, st));
        PGPU_HIP(hipStreamSynchronize(st));
        const int64_t kmin = p->h_mm[kMmKMin], gmin = p->h_mm[kMmGMin];
        const uint64_t Rk = (uint64_t)((int64_t)p->h_mm[kMmKMax] - kmin + 1);
        const uint64_t Rg = (uint64_t)((int64_t)p->h_mm[kMmGMax] - gmin + 1);
        if (Rk <= kMaxKeyRange && Rg <= kMaxGroupRange) {
            PGPU_TRY(run_partitioned(p, d_dkey, d_dgrp, nD, d_fkey, d_fval, nF, kmin, Rk, gmin, Rg));
            done = p->h_mm[kMmDup] == 0;  // a duplicate key voids the result
            p->last_path = 1;
        }
    }
    if (!done) {
        PGPU_TRY(run_general(p, d_dkey, d_dgrp, nD, d_fkey, d_fval, nF));
        p->last_path = 2;
    }
    p->ngroups = p->h_counters[kCtCursor];
    p->joined = p->h_counters[kCtJoined];
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

int pgpu_groupjoin_rows(pgpu_groupjoin* p, int acc, uint64_t first, uint64_t n, void* dst) {
    if (!p || (n && !dst)) return fail(PGPU_EINVAL, "pgpu_groupjoin_rows: null argument");
    if (acc != PGPU_GJ_ROW_SUM && acc != PGPU_GJ_ROW_COUNT)
        return fail(PGPU_EINVAL, "pgpu_groupjoin_rows: unknown accumulator %d", acc);
    if (first > p->ngroups || n > p->ngroups - first)
        return fail(PGPU_EINVAL, "pgpu_groupjoin_rows: rows [%llu, +%llu) past %llu groups", (unsigned long long)first,
                    (unsigned long long)n, (unsigned long long)p->ngroups);
    if (!n) return PGPU_OK;
    PGPU_TRY(ctx_activate(p->ctx));
    hipStream_t st = p->ctx->stream;
    const int count = acc == PGPU_GJ_ROW_COUNT ? 1 : 0;
    if (((uintptr_t)dst & 15) == 0 && host_alloc_holds(dst, n * 12)) {
        // pinned memory from pgpu_host_alloc: the kernel writes it directly
        // (one launch, no staging copy, no DMA set-up)
        hipLaunchKernelGGL(k_gj_rows_host, dim3(grid_for(n)), dim3(256), 0, st, p->okey, p->osum, p->ocnt, count,
                           first, n, (uint4*)dst);
        PGPU_LAUNCH_CHECK();
    } else {
        hipLaunchKernelGGL(k_gj_rows, dim3(grid_for(n)), dim3(256), 0, st, p->okey, p->osum, p->ocnt, count, first,
                           n, p->orow);
        PGPU_LAUNCH_CHECK();
        PGPU_HIP(hipMemcpyAsync(dst, p->orow, n * 12, hipMemcpyDeviceToHost, st));
    }
    PGPU_HIP(hipStreamSynchronize(st));
    return PGPU_OK;
}

int pgpu_groupjoin_path(pgpu_groupjoin* p, int* path) {
    if (!p || !path) return fail(PGPU_EINVAL, "pgpu_groupjoin_path: null argument");
    *path = p->last_path;
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
