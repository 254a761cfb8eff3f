// Batched equi-joins: many independent (build, probe) pairs of uneven sizes
// joined with fused count + sum consumers in one launch sequence.
//
// Reference path replaced, once per join: the queryhashjoin plan shape
// (unit_tests/queryhashjoin.cpp) of HashJoinOp build + probe
// (operators/join.cpp:396-464, :503-562, :641-667) over a chained-bucket
// HashTable (util/hashtable.cpp:230-273), feeding AggregateCount /
// AggregateSum (operators/aggregatecount.cpp:37-52,
// operators/aggregatesum.cpp:81-107).  The reference runs one such plan per
// query; here a whole batch shares every launch:
//
//   k_bhist     one block per chunk (the segment's chunk_rows rows) of any side of any
//               join: LDS histogram over that join's 2^bits partitions
//               (bits from the join's build size, one pass, <= 12 bits)
//   k_bscan     one block per side of a join: exclusive scans of the
//               histograms -> each chunk's write offsets + partition bounds
//   k_bscatter  one block per chunk: tiles reordered by partition in LDS and
//               written as contiguous runs at exact offsets (no atomics)
//   k_bjoin     persistent: every workgroup pulls (join, partition) items
//               from one atomic queue, builds the partition's LDS table and
//               streams its probe rows; per-join results by atomics
//
// The batch plan captures the sequence once into a hipGraph and replays it.
// The autotuner times every (workgroup shape, partition size) candidate on
// the batch itself and keeps the fastest.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <type_traits>

#include "batch_internal.h"
#include "join_common.h"

using namespace pgpu::dev;
using pgpu::batch::Chunk;
using pgpu::batch::Job;
using pgpu::batch::kChunkRows;
using pgpu::batch::kMaxBits;
using pgpu::batch::kScatterTileBytes;
using pgpu::batch::Seg;

namespace {

constexpr int kND = 1 << kMaxBits;

template <typename KT>
__global__ __launch_bounds__(512) void k_bhist(const Seg* __restrict__ segs, const Chunk* __restrict__ chunks,
                                               uint32_t chunk_base, uint32_t* __restrict__ H) {
    __shared__ uint32_t h[kND];
    const uint32_t ci = chunk_base + blockIdx.x;
    const Chunk c = chunks[ci];
    const Seg s = segs[c.seg];
    const uint64_t r0 = c.row0, r1 = min<uint64_t>(s.n, r0 + s.chunk_rows);
    const uint32_t nd = 1u << s.bits;
    uint32_t* out = H + s.hist + (uint64_t)(ci - s.chunk0) * nd;
    if (s.bits == 0) {
        if (threadIdx.x == 0) out[0] = (uint32_t)(r1 - r0);
        return;
    }
    for (uint32_t i = threadIdx.x; i < nd; i += 512) h[i] = 0;
    __syncthreads();
    const KT* key = (const KT*)s.key;
    const int sh = 64 - (int)s.bits;
    constexpr int G = 16 / sizeof(KT);
    uint64_t r = r0;
    if (((uintptr_t)key & 15) == 0) {
        const uint64_t ng = (r1 - r0) / G;
        const uint4* kv = (const uint4*)(key + r0);
#pragma unroll 4
        for (uint64_t g = threadIdx.x; g < ng; g += 512) {
            const uint4 q = nt_load(&kv[g]);
            KT ks[G];
            memcpy(ks, &q, 16);
#pragma unroll
            for (int u = 0; u < G; ++u) atomicAdd(&h[hmix(ks[u]) >> sh], 1u);
        }
        r = r0 + ng * G;
    }
    for (r += threadIdx.x; r < r1; r += 512) atomicAdd(&h[hmix(key[r]) >> sh], 1u);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nd; i += 512) out[i] = h[i];
}

// Offsets of every segment in three launches, spread over the device (one
// large segment, e.g. a single served query of 2^12 partitions x 256 chunks,
// is 4 MB of histogram):
//   k_bcol   block (g, s): column totals of bins [64 g, 64 g + 64) of
//            segment s over its chunks (16 waves, a sixteenth of the chunks
//            each), into bnd[d]
//   k_bbins  block s: exclusive scan of the totals in place; bnd[nd] = n
//   k_bapply block (g, s): h[c][d] = bnd[d] + sum of h[c'][d] for c' < c
constexpr int kColBins = 64;
constexpr int kColWaves = 16;

// Sum of column d over chunks [c0, c1): 32 loads in flight per step (a
// large segment's column is ~128 chunks per wave).
__device__ __forceinline__ uint32_t col_sum(const uint32_t* h, uint32_t nd, uint32_t d, uint32_t c0, uint32_t c1) {
    constexpr int B = 32;
    uint32_t t = 0;
    for (uint32_t cb = c0; cb < c1; cb += B) {
        uint32_t v[B];
#pragma unroll
        for (int j = 0; j < B; ++j) v[j] = cb + j < c1 ? h[(uint64_t)(cb + j) * nd + d] : 0u;
#pragma unroll
        for (int j = 0; j < B; ++j) t += v[j];
    }
    return t;
}

__global__ __launch_bounds__(64 * kColWaves) void k_bcol(const Seg* __restrict__ segs, uint32_t seg_base,
                                                          const uint32_t* __restrict__ H, uint32_t* __restrict__ bnd) {
    __shared__ uint32_t part[kColWaves][kColBins];
    const Seg s = segs[seg_base + blockIdx.y];
    const uint32_t nd = 1u << s.bits, nc = s.nchunks;
    if (blockIdx.x * kColBins >= nd) return;  // whole block: fewer bins than groups
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t d = blockIdx.x * kColBins + lane;
    const uint32_t c0 = (uint32_t)((uint64_t)nc * w / kColWaves), c1 = (uint32_t)((uint64_t)nc * (w + 1) / kColWaves);
    part[w][lane] = d < nd ? col_sum(H + s.hist, nd, d, c0, c1) : 0u;
    __syncthreads();
    if (w == 0 && d < nd) {
        uint32_t t = 0;
#pragma unroll
        for (int q = 0; q < kColWaves; ++q) t += part[q][lane];
        bnd[s.bnd + d] = t;
    }
}

__global__ __launch_bounds__(1024) void k_bbins(const Seg* __restrict__ segs, uint32_t seg_base,
                                                uint32_t* __restrict__ bnd) {
    __shared__ uint32_t ws[1024 / 64 + 1];
    const Seg s = segs[seg_base + blockIdx.x];
    const uint32_t nd = 1u << s.bits;
    uint32_t* b = bnd + s.bnd;
    constexpr int PER = kND / 1024;
    uint32_t loc[PER], sum = 0;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const uint32_t idx = threadIdx.x * PER + u;
        loc[u] = idx < nd ? b[idx] : 0u;
        sum += loc[u];
    }
    uint32_t total;
    uint32_t ex = block_excl_scan<1024>(sum, ws, &total);
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const uint32_t idx = threadIdx.x * PER + u;
        if (idx < nd) b[idx] = ex;
        ex += loc[u];
    }
    if (threadIdx.x == 0) b[nd] = (uint32_t)s.n;
}

__global__ __launch_bounds__(64 * kColWaves) void k_bapply(const Seg* __restrict__ segs, uint32_t seg_base,
                                                            uint32_t* __restrict__ H, const uint32_t* __restrict__ bnd) {
    __shared__ uint32_t part[kColWaves][kColBins];
    const Seg s = segs[seg_base + blockIdx.y];
    const uint32_t nd = 1u << s.bits, nc = s.nchunks;
    if (blockIdx.x * kColBins >= nd) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t d = blockIdx.x * kColBins + lane;
    const uint32_t c0 = (uint32_t)((uint64_t)nc * w / kColWaves), c1 = (uint32_t)((uint64_t)nc * (w + 1) / kColWaves);
    uint32_t* h = H + s.hist;
    part[w][lane] = d < nd ? col_sum(h, nd, d, c0, c1) : 0u;
    __syncthreads();
    if (d >= nd) return;
    uint32_t run = bnd[s.bnd + d];
    for (int q = 0; q < w; ++q) run += part[q][lane];
    constexpr int B = 16;  // chunks per step: their counts are loaded before the first store
    for (uint32_t cb = c0; cb < c1; cb += B) {
        uint32_t v[B];
#pragma unroll
        for (int j = 0; j < B; ++j) v[j] = cb + j < c1 ? h[(uint64_t)(cb + j) * nd + d] : 0u;
#pragma unroll
        for (int j = 0; j < B; ++j) {
            if (cb + j < c1) h[(uint64_t)(cb + j) * nd + d] = run;
            run += v[j];
        }
    }
}

// Loads rows [g*G, g*G+G) of a chunk-relative group into k/v (bounds per row).
template <typename KT, bool HV>
__device__ __forceinline__ void load_group(const KT* key, const KT* pay, bool vec, uint64_t row, uint64_t r1, KT* k,
                                           KT* v) {
    constexpr int G = 16 / sizeof(KT);
    if (vec && row + G <= r1) {
        const uint4 q = *(const uint4*)(key + row);
        memcpy(k, &q, 16);
        if constexpr (HV) {
            if (pay) {
                const uint4 p = *(const uint4*)(pay + row);
                memcpy(v, &p, 16);
            } else {
#pragma unroll
                for (int u = 0; u < G; ++u) v[u] = 0;
            }
        }
    } else {
#pragma unroll
        for (int u = 0; u < G; ++u) {
            if (row + u < r1) {
                k[u] = key[row + u];
                if constexpr (HV) v[u] = pay ? pay[row + u] : (KT)0;
            }
        }
    }
}

// TB bytes of LDS staging per tile, ND partitions at most (the launch's
// largest 2^bits), T threads: fan-outs up to 1024 take 128 KB tiles, twice
// the rows per partition run of the 64 KB tile that 4096 partitions leave
// room for (64 KB tiles at two workgroups per CU leave runs of 8 rows, half a
// line: slower).
template <typename KT, bool HV, int TB, int ND, int T>
__global__ __launch_bounds__(T, T == 512 ? 4 : 1) void k_bscatter(const Seg* __restrict__ segs, const Chunk* __restrict__ chunks,
                                                uint32_t chunk_base, const uint32_t* __restrict__ H,
                                                Row<KT, HV>* __restrict__ out) {
    using E = Row<KT, HV>;
    constexpr int G = 16 / sizeof(KT);            // rows per 16-byte key load
    constexpr int TR = TB / sizeof(E);            // rows per tile
    constexpr int NG = TR / (T * G);              // row groups per thread per tile
    static_assert(NG >= 1 && TR % (T * G) == 0, "tile shape");
    static_assert(ND >= T && ND <= kND, "partition table");
    __shared__ E stage[TR];
    // cnt: the tile's rows per partition, then (after the scan) each
    // partition's first stage row; cur: next output row of each partition
    // (segment-relative); delta: the output row of stage row j of partition d
    // is delta[d] + j (mod 2^32)
    __shared__ uint32_t cnt[ND], cur[ND], delta[ND];
    static_assert(ND <= 4096 && TR < (1 << 20), "digit (12 bits) and rank (20 bits) share a register");
    __shared__ uint32_t ws[T / 64 + 1];
    // consecutive chunks of a segment write adjacent runs of every partition:
    // neighbours run side by side on one XCD (chunks of 32), so the L2 merges
    // the lines they share
    const uint32_t ci = chunk_base + xcd_remap_chunked<32>(blockIdx.x, gridDim.x);
    const Chunk c = chunks[ci];
    const Seg s = segs[c.seg];
    const uint64_t r0 = c.row0, r1 = min<uint64_t>(s.n, r0 + s.chunk_rows);
    const uint32_t nd = 1u << s.bits;
    const KT* key = (const KT*)s.key;
    const KT* pay = (const KT*)s.pay;
    E* dst = out + s.dst;
    const uint32_t* h = H + s.hist + (uint64_t)(ci - s.chunk0) * nd;
    const bool vec = ((uintptr_t)key & 15) == 0 && (!pay || ((uintptr_t)pay & 15) == 0);
    if (s.bits == 0) {  // one partition: copy into the row layout
        const uint32_t base = h[0];
        for (uint64_t g = r0 + (uint64_t)threadIdx.x * G; g < r1; g += (uint64_t)T * G) {
            KT k[G], v[G];
            load_group<KT, HV>(key, pay, vec, g, r1, k, v);
#pragma unroll
            for (int u = 0; u < G; ++u) {
                if (g + u < r1) {
                    E e;
                    e.k = k[u];
                    if constexpr (HV) e.v = v[u];
                    dst[base + (g + u - r0)] = e;
                }
            }
        }
        return;
    }
    for (uint32_t d = threadIdx.x; d < nd; d += T) {
        cur[d] = h[d];
        cnt[d] = 0;
    }
    __syncthreads();
    const int sh = 64 - (int)s.bits;
    constexpr int PER = ND / T;
    for (uint64_t t0 = r0; t0 < r1; t0 += TR) {
        const uint64_t te = min<uint64_t>(r1, t0 + TR);
        const uint32_t tn = (uint32_t)(te - t0);
        KT k[NG][G], v[NG][G];
        uint32_t dr[NG][G];  // partition | rank in the tile << 12
#pragma unroll
        for (int q = 0; q < NG; ++q) {
            const uint64_t row = t0 + ((uint64_t)q * T + threadIdx.x) * G;
            load_group<KT, HV>(key, pay, vec, row, te, k[q], v[q]);
#pragma unroll
            for (int u = 0; u < G; ++u)
                if (row + u < te) {
                    const uint32_t d = (uint32_t)(hmix(k[q][u]) >> sh);
                    dr[q][u] = d | (atomicAdd(&cnt[d], 1u) << 12);
                }
        }
        __syncthreads();
        uint32_t loc[PER], sum = 0;
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const uint32_t idx = threadIdx.x * PER + u;
            loc[u] = idx < nd ? cnt[idx] : 0u;
            sum += loc[u];
        }
        uint32_t total;
        uint32_t ex = block_excl_scan<T>(sum, ws, &total);
        // each thread owns bins [tid * PER, tid * PER + PER): their counts
        // become the partitions' first stage rows
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const uint32_t idx = threadIdx.x * PER + u;
            if (idx < nd) {
                cnt[idx] = ex;
                delta[idx] = cur[idx] - ex;
                cur[idx] += loc[u];
            }
            ex += loc[u];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NG; ++q) {
            const uint64_t row = t0 + ((uint64_t)q * T + threadIdx.x) * G;
#pragma unroll
            for (int u = 0; u < G; ++u) {
                if (row + u < te) {
                    E e;
                    e.k = k[q][u];
                    if constexpr (HV) e.v = v[q][u];
                    stage[cnt[dr[q][u] & 0xFFF] + (dr[q][u] >> 12)] = e;
                }
            }
        }
        __syncthreads();
        // every stage write is done: the owned counters restart at 0 (the
        // next tile ranks after the barrier at the end of this one)
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const uint32_t idx = threadIdx.x * PER + u;
            if (idx < nd) cnt[idx] = 0;
        }
        // runs of each partition at its exact offsets (plain stores: the L2
        // merges the partial lines at run ends).
        if (tn == (uint32_t)TR) {
            constexpr int SB = 4;  // rows whose LDS reads are issued together
#pragma unroll
            for (int b0 = 0; b0 < TR / T; b0 += SB) {
                E e[SB];
                uint32_t dl[SB];
#pragma unroll
                for (int u = 0; u < SB; ++u) e[u] = stage[threadIdx.x + (b0 + u) * T];
#pragma unroll
                for (int u = 0; u < SB; ++u) dl[u] = delta[(uint32_t)(hmix(e[u].k) >> sh)];
#pragma unroll
                for (int u = 0; u < SB; ++u) dst[dl[u] + (uint32_t)(threadIdx.x + (b0 + u) * T)] = e[u];
            }
        } else {
            for (uint32_t pos = threadIdx.x; pos < tn; pos += T) {
                const E e = stage[pos];
                dst[delta[(uint32_t)(hmix(e.k) >> sh)] + pos] = e;
            }
        }
        if (te < r1) __syncthreads();  // chunks of several tiles: counters reset, stage read out
    }
}

// LDS table slot: key, match count and folded payload sum in 16 bytes for
// int32 keys; int64 keys keep their counts in a packed 16-bit array.
template <typename KT>
struct LSlot;
template <>
struct __align__(16) LSlot<int32_t> {
    int32_t key;
    uint32_t cnt;
    unsigned long long sum;
};
template <>
struct __align__(16) LSlot<int64_t> {
    long long key;
    unsigned long long sum;
};

template <int CAP>
struct Log2;
template <>
struct Log2<4096> {
    static constexpr int v = 12;
};
template <>
struct Log2<8192> {
    static constexpr int v = 13;
};

// ilist (optional): the work items to run are ilist[0, *ilist_n) instead of
// 0 .. n_items - 1 (the partitions the unique-key kernel handed back).
template <typename KT, bool HVS, int CAP, int T>
__global__ __launch_bounds__(T) void k_bjoin(const Job* __restrict__ jobs, const uint32_t* __restrict__ item_job,
                                             uint32_t n_items, const Row<KT, true>* __restrict__ R,
                                             const Row<KT, HVS>* __restrict__ S, const uint32_t* __restrict__ bndR,
                                             const uint32_t* __restrict__ bndS, uint32_t* __restrict__ queue,
                                             unsigned long long* __restrict__ res,
                                             const uint32_t* __restrict__ ilist, const uint32_t* __restrict__ ilist_n) {
    constexpr int L = Log2<CAP>::v;
    constexpr uint32_t PIECE = CAP / 2;  // load factor <= 0.5
    constexpr bool W = sizeof(KT) == 8;
    static_assert(PIECE < 65536, "16-bit LDS match counters");
    constexpr KT EMPTY = KeyEmpty<KT>::v;
    using SR = Row<KT, HVS>;
    constexpr int U = sizeof(SR) >= 16 ? 3 : 4;  // probe rows per thread per batch (two batches in flight)
    constexpr int RB = (PIECE + T - 1) / T;       // build rows per thread per piece
    __shared__ LSlot<KT> tab[CAP];
    __shared__ uint32_t tc[W ? CAP / 2 : 1];
    __shared__ unsigned long long sp_cnt, sp_sum;
    __shared__ unsigned long long red[3][T / 64];
    __shared__ uint32_t s_item;
    const int tid = threadIdx.x;
    const unsigned short* tc16 = (const unsigned short*)tc;  // little-endian halves of tc[]
    const uint32_t n_run = ilist ? *ilist_n : n_items;
    while (true) {
        if (tid == 0) s_item = atomicAdd(queue, 1u);
        __syncthreads();
        if (s_item >= n_run) break;
        const uint32_t item = ilist ? ilist[s_item] : s_item;
        const uint32_t j = item_job[item];
        const Job jb = jobs[j];
        const uint32_t K = jb.split > 1 ? jb.split : 1u;
        const uint32_t d = (item - jb.item0) / K, ks = (item - jb.item0) % K;
        const uint64_t rb = jb.rdst + bndR[jb.rbnd + d];
        const uint32_t nR = bndR[jb.rbnd + d + 1] - bndR[jb.rbnd + d];
        // slice ks of K of the partition's probe rows (every slice builds the
        // partition's table: large probes of small tables spread over the CUs)
        const uint64_t s0 = bndS[jb.sbnd + d], sl = bndS[jb.sbnd + d + 1] - s0;
        const uint64_t sb = jb.sdst + s0 + sl * ks / K;
        const uint64_t se = jb.sdst + s0 + sl * (ks + 1) / K;
        // LDS slot = bits [64 - bits - L, 64 - bits) of hmix(k): they lie in
        // the high word of the product (bits + L <= 32), computed without its
        // low half.
        const int sh = 32 - (int)jb.bits - L;
        auto slot_of = [&](KT key) -> uint32_t {
            const uint64_t kk = (uint64_t)(int64_t)key;
            const uint32_t lo = (uint32_t)kk, hi = (uint32_t)(kk >> 32);
            const uint32_t h = __umulhi(lo, (uint32_t)kFib) + lo * (uint32_t)(kFib >> 32) + hi * (uint32_t)kFib;
            return (h >> sh) & (CAP - 1);
        };
        unsigned long long acc_c = 0, acc_b = 0, acc_p = 0;
        SR x[U] = {};
        // rows of the batch at b0 this thread holds: b0 + tid + u*T < se
        auto rows_in = [&](uint64_t b0) -> uint32_t {
            const uint64_t r0 = b0 + tid;
            return r0 < se ? (uint32_t)min<uint64_t>(U, (se - r0 + T - 1) / T) : 0u;
        };
        // A batch's probe rows through a buffer resource based at the batch:
        // 32-bit offsets (fewer registers in flight) and rows past the item's
        // end read as zero (masked by the probe).
        auto batch_rsrc = [&](uint64_t b0) {
            const uint64_t left = b0 < se ? (se - b0) * sizeof(SR) : 0;
            return buf_rsrc(S + b0, left < 0xFFFFFFF0ull ? (uint32_t)left : 0xFFFFFFF0u);
        };
        auto fetch = [&](uint64_t b0) {
            const uint32_t cnt = rows_in(b0);
            const __amdgpu_buffer_rsrc_t rs = batch_rsrc(b0);
#pragma unroll
            for (int u = 0; u < U; ++u)
                if ((uint32_t)u < cnt) x[u] = buf_load<SR, true>(rs, (uint32_t)(tid + u * T) * (uint32_t)sizeof(SR));
        };
        // Unpredicated fetch of the batch at b0 into z.
        auto fetch_clamped = [&](SR(&z)[U], uint64_t b0) {
            const __amdgpu_buffer_rsrc_t rs = batch_rsrc(b0);
#pragma unroll
            for (int u = 0; u < U; ++u) z[u] = buf_load<SR, true>(rs, (uint32_t)(tid + u * T) * (uint32_t)sizeof(SR));
        };
        Row<KT, true> br[RB];
        auto fetch_build = [&](uint32_t p0) {
            const uint32_t pn = min(PIECE, nR - p0);
#pragma unroll
            for (int b = 0; b < RB; ++b) {
                const uint32_t i = tid + b * T;
                if (i < pn) br[b] = nt_load(&R[rb + p0 + i]);
            }
        };
        const bool work = nR && se > sb;
        // Neither the build rows nor the first probe batch depend on the
        // table: both are in flight while it is cleared.
        if (work) {
            fetch_build(0);
            fetch(sb);
        }
        for (uint32_t p0 = 0; work && p0 < nR; p0 += PIECE) {
            if (p0) fetch_build(p0);  // later pieces (skewed build partitions) fetch their rows here
            for (int i = tid; i < CAP; i += T) {
                tab[i].key = EMPTY;
                tab[i].sum = 0;
                if constexpr (W) {
                    if (i < CAP / 2) tc[i] = 0;
                } else {
                    tab[i].cnt = 0;
                }
            }
            if (tid == 0) sp_cnt = sp_sum = 0;
            __syncthreads();
            const uint32_t pn = min(PIECE, nR - p0);
#pragma unroll
            for (int b = 0; b < RB; ++b) {
                if ((uint32_t)(tid + b * T) >= pn) continue;
                const Row<KT, true> r = br[b];
                if (r.k == EMPTY) {
                    atomicAdd(&sp_cnt, 1ull);
                    atomicAdd(&sp_sum, (unsigned long long)(int64_t)r.v);
                    continue;
                }
                uint32_t slot = slot_of(r.k);
                while (true) {
                    KT prev;
                    if constexpr (W)
                        prev = (KT)atomicCAS((unsigned long long*)&tab[slot].key, (unsigned long long)EMPTY,
                                             (unsigned long long)r.k);
                    else
                        prev = (KT)atomicCAS((int*)&tab[slot].key, (int)EMPTY, (int)r.k);
                    if (prev == EMPTY || prev == r.k) {
                        if constexpr (W)
                            atomicAdd(&tc[slot >> 1], 1u << (16 * (slot & 1)));
                        else
                            atomicAdd(&tab[slot].cnt, 1u);
                        atomicAdd(&tab[slot].sum, (unsigned long long)(int64_t)r.v);
                        break;
                    }
                    slot = (slot + 1) & (CAP - 1);
                }
            }
            __syncthreads();
            if (p0) fetch(sb);  // later pieces re-read the item's probe rows
            const unsigned long long spc = sp_cnt, sps = sp_sum;
            unsigned long long pe_n = 0, pe_v = 0;  // this piece's probe rows keyed by the empty marker (and payloads)
            // One batch: all U slot reads in flight at once, collision chains
            // walked only by the lanes that need them, selects instead of
            // branches; FULL batches (every thread holds U rows) skip the
            // row-count masks.
            auto probe_batch = [&](auto full_tag, const SR(&x)[U], uint32_t cnt) {
                constexpr bool FULL = decltype(full_tag)::value;
                uint32_t slot[U];
                LSlot<KT> t[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    slot[u] = slot_of(x[u].k);
                    t[u] = tab[slot[u]];
                }
                bool any = false;
#pragma unroll
                for (int u = 0; u < U; ++u)
                    any |= (FULL || (uint32_t)u < cnt) && t[u].key != x[u].k && t[u].key != EMPTY;
                if (any) {
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        if (!FULL && (uint32_t)u >= cnt) continue;
                        while (t[u].key != x[u].k && t[u].key != EMPTY) {
                            slot[u] = (slot[u] + 1) & (CAP - 1);
                            t[u] = tab[slot[u]];
                        }
                    }
                }
                // Each walk ended on the key's slot or on an empty one, whose
                // count and sum are 0: no hit / miss select.  A probe key
                // equal to the empty marker also ends on an empty slot; those
                // rows are tallied and take the side accumulator per piece.
                uint32_t bc = 0;  // fits 32 bits: U rows x PIECE matches
                unsigned long long bs = 0, bp = 0;
                bool em = false;
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const bool live = FULL || (uint32_t)u < cnt;
                    uint32_t c;
                    if constexpr (W)
                        c = tc16[slot[u]];
                    else
                        c = t[u].cnt;
                    unsigned long long sbv = t[u].sum;
                    if (!live) {
                        c = 0;
                        sbv = 0;
                    }
                    bc += c;
                    bs += sbv;
                    if constexpr (HVS) {
                        if constexpr (W)
                            bp += (unsigned long long)c * (unsigned long long)(int64_t)x[u].v;
                        else  // c <= PIECE < 2^31: one signed 32 x 32 -> 64-bit multiply-add
                            bp += (unsigned long long)((int64_t)(int32_t)c * (int64_t)x[u].v);
                    }
                    em |= live && x[u].k == EMPTY;
                }
                acc_c += bc;
                acc_b += bs;
                acc_p += bp;
                if (__any(em)) {  // rare: a wavefront holds a probe key equal to the empty marker
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        if ((FULL || (uint32_t)u < cnt) && x[u].k == EMPTY) {
                            pe_n += 1;
                            if constexpr (HVS) pe_v += (unsigned long long)(int64_t)x[u].v;
                        }
                    }
                }
            };
            auto probe = [&](const SR(&z)[U], uint64_t b0) {
                if (b0 + (uint64_t)U * T <= se)
                    probe_batch(std::true_type{}, z, (uint32_t)U);
                else
                    probe_batch(std::false_type{}, z, rows_in(b0));
            };
            // Two batches in flight: the next batch's rows load while this
            // one probes the LDS table.
            SR y[U];
            for (uint64_t b0 = sb; b0 < se; b0 += 2ull * U * T) {
                fetch_clamped(y, b0 + (uint64_t)U * T);
                probe(x, b0);
                fetch_clamped(x, b0 + 2ull * U * T);
                if (b0 + (uint64_t)U * T < se) probe(y, b0 + (uint64_t)U * T);
            }
            acc_c += pe_n * spc;
            acc_b += pe_n * sps;
            acc_p += pe_v * spc;
            __syncthreads();
        }
        acc_c = wave_sum(acc_c);
        acc_b = wave_sum(acc_b);
        acc_p = wave_sum(acc_p);
        const int lane = tid & 63, w = tid >> 6;
        if (lane == 0) {
            red[0][w] = acc_c;
            red[1][w] = acc_b;
            red[2][w] = acc_p;
        }
        __syncthreads();
        if (tid == 0) {
            unsigned long long cc = 0, bb = 0, pp = 0;
            for (int k = 0; k < T / 64; ++k) {
                cc += red[0][k];
                bb += red[1][k];
                pp += red[2][k];
            }
            unsigned long long* o = res + 4 * (uint64_t)j;
            if (cc) {
                atomicAdd(&o[0], cc);
                atomicAdd(&o[1], bb);
                atomicAdd(&o[2], pp);
            }
            const unsigned long long pieces = work ? (nR + PIECE - 1) / PIECE : 0;
            if (pieces > 1) atomicMax(&o[3], pieces);
        }
    }
}

// Unique-key LDS join for 4-byte keys (variant 3): slots of (key, build
// payload) in 8 bytes, kCap8 of them in 64 KB, so two 512-thread workgroups
// share a CU and one's table build overlaps the other's probe loads.  A
// partition whose build rows repeat a key, or number more than kPiece8, is
// appended to `redo` (count in queue[1]) and left to k_bjoin, which runs
// after this kernel over exactly those items; nothing of it is counted here.
// Results per job as k_bjoin: res[4 j + 0..2] = matches, sum of matched
// build payloads, sum of probe payload x matches.
constexpr int kCap8 = 8192;
constexpr uint32_t kPiece8 = 6144;  // load factor <= 0.75
constexpr int kT8 = 512;

template <bool HVS>
__global__ __launch_bounds__(kT8, 4) void k_bjoin8(const Job* __restrict__ jobs, const uint32_t* __restrict__ item_job,
                                                uint32_t n_items, const Row<int32_t, true>* __restrict__ R,
                                                const Row<int32_t, HVS>* __restrict__ S,
                                                const uint32_t* __restrict__ bndR, const uint32_t* __restrict__ bndS,
                                                uint32_t* __restrict__ queue, uint32_t* __restrict__ redo,
                                                unsigned long long* __restrict__ res) {
    constexpr int T = kT8, CAP = kCap8, L = Log2<kCap8>::v;
    constexpr int32_t EMPTY = INT32_MIN;
    using SR = Row<int32_t, HVS>;
    constexpr int U = 3;                                // probe rows per thread per batch (two batches in flight)
    constexpr int RB = (kPiece8 + T - 1) / T;           // build rows per thread
    __shared__ int2 tab[CAP];                           // (key, build payload); empty: (EMPTY, 0)
    __shared__ unsigned long long sp_cnt, sp_sum;       // build rows keyed by the empty marker
    __shared__ unsigned long long red[3][T / 64];
    __shared__ uint32_t s_item;
    __shared__ int s_dup;
    const int tid = threadIdx.x;
    while (true) {
        if (tid == 0) {
            s_item = atomicAdd(&queue[0], 1u);
            s_dup = 0;
            sp_cnt = sp_sum = 0;
        }
        __syncthreads();
        const uint32_t item = s_item;
        if (item >= n_items) break;
        const uint32_t j = item_job[item];
        const Job jb = jobs[j];
        const uint32_t K = jb.split > 1 ? jb.split : 1u;
        const uint32_t d = (item - jb.item0) / K, ks = (item - jb.item0) % K;
        const uint64_t rb = jb.rdst + bndR[jb.rbnd + d];
        const uint32_t nR = bndR[jb.rbnd + d + 1] - bndR[jb.rbnd + d];
        const uint64_t s0 = bndS[jb.sbnd + d], sl = bndS[jb.sbnd + d + 1] - s0;
        const uint64_t sb = jb.sdst + s0 + sl * ks / K;
        const uint64_t se = jb.sdst + s0 + sl * (ks + 1) / K;
        if (!nR || se <= sb) {  // nothing matches
            __syncthreads();
            continue;
        }
        if (nR > kPiece8) {  // more than one table's worth of build rows: the general kernel
            if (tid == 0) redo[atomicAdd(&queue[1], 1u)] = item;
            __syncthreads();
            continue;
        }
        const int sh = 32 - (int)jb.bits - L;
        auto slot_of = [&](int32_t key) -> uint32_t {
            const uint64_t kk = (uint64_t)(int64_t)key;
            const uint32_t lo = (uint32_t)kk, hi = (uint32_t)(kk >> 32);
            const uint32_t h = __umulhi(lo, (uint32_t)kFib) + lo * (uint32_t)(kFib >> 32) + hi * (uint32_t)kFib;
            return (h >> sh) & (CAP - 1);
        };
        auto rows_in = [&](uint64_t b0) -> uint32_t {
            const uint64_t r0 = b0 + tid;
            return r0 < se ? (uint32_t)min<uint64_t>(U, (se - r0 + T - 1) / T) : 0u;
        };
        auto batch_rsrc = [&](uint64_t b0) {
            const uint64_t left = b0 < se ? (se - b0) * sizeof(SR) : 0;
            return buf_rsrc(S + b0, left < 0xFFFFFFF0ull ? (uint32_t)left : 0xFFFFFFF0u);
        };
        auto fetch_clamped = [&](SR(&z)[U], uint64_t b0) {
            const __amdgpu_buffer_rsrc_t rs = batch_rsrc(b0);
#pragma unroll
            for (int u = 0; u < U; ++u) z[u] = buf_load<SR, true>(rs, (uint32_t)(tid + u * T) * (uint32_t)sizeof(SR));
        };
        // build rows and the first probe batch are in flight while the table is cleared
        Row<int32_t, true> br[RB];
#pragma unroll
        for (int b = 0; b < RB; ++b)
            if ((uint32_t)(tid + b * T) < nR) br[b] = nt_load(&R[rb + tid + b * T]);
        SR x[U];
        fetch_clamped(x, sb);
        for (int i = tid; i < CAP; i += T) tab[i] = make_int2(EMPTY, 0);
        __syncthreads();
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            if ((uint32_t)(tid + b * T) >= nR) continue;
            const Row<int32_t, true> r = br[b];
            if (r.k == EMPTY) {
                atomicAdd(&sp_cnt, 1ull);
                atomicAdd(&sp_sum, (unsigned long long)(int64_t)r.v);
                continue;
            }
            uint32_t slot = slot_of(r.k);
            while (true) {
                const int prev = atomicCAS(&tab[slot].x, EMPTY, r.k);
                if (prev == EMPTY) {
                    tab[slot].y = r.v;  // the slot is this row's alone
                    break;
                }
                if (prev == r.k) {  // a repeated key: the general kernel takes the partition
                    s_dup = 1;
                    break;
                }
                slot = (slot + 1) & (CAP - 1);
            }
        }
        __syncthreads();
        if (s_dup) {
            if (tid == 0) redo[atomicAdd(&queue[1], 1u)] = item;
            __syncthreads();
            continue;
        }
        unsigned long long acc_c = 0, acc_b = 0, acc_p = 0, pe_n = 0, pe_v = 0;
        // One batch: every walk ends on the key's slot or on an empty one
        // (payload 0); a hit is key equality.  A probe key equal to the empty
        // marker "hits" an empty slot: such rows are taken back out and
        // matched against the build rows of that key (sp_cnt, sp_sum).
        auto probe_batch = [&](auto full_tag, const SR(&z)[U], uint32_t cnt) {
            constexpr bool FULL = decltype(full_tag)::value;
            uint32_t slot[U];
            int2 t[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                slot[u] = slot_of(z[u].k);
                t[u] = tab[slot[u]];
            }
            bool any = false;
#pragma unroll
            for (int u = 0; u < U; ++u)
                any |= (FULL || (uint32_t)u < cnt) && t[u].x != z[u].k && t[u].x != EMPTY;
            if (any) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (!FULL && (uint32_t)u >= cnt) continue;
                    while (t[u].x != z[u].k && t[u].x != EMPTY) {
                        slot[u] = (slot[u] + 1) & (CAP - 1);
                        t[u] = tab[slot[u]];
                    }
                }
            }
            uint32_t bc = 0;
            long long bs = 0, bp = 0;
            bool em = false;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const bool live = FULL || (uint32_t)u < cnt;
                const bool hit = live && t[u].x == z[u].k;
                bc += hit ? 1u : 0u;
                bs += live ? (long long)t[u].y : 0ll;
                if constexpr (HVS) bp += hit ? (long long)z[u].v : 0ll;
                em |= live && z[u].k == EMPTY;
            }
            acc_c += bc;
            acc_b += (unsigned long long)bs;
            acc_p += (unsigned long long)bp;
            if (__any(em)) {  // rare: a wavefront holds a probe key equal to the empty marker
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if ((FULL || (uint32_t)u < cnt) && z[u].k == EMPTY) {
                        acc_c -= 1;  // its "hit" on an empty slot (payload 0)
                        pe_n += 1;
                        if constexpr (HVS) {
                            acc_p -= (unsigned long long)(int64_t)z[u].v;
                            pe_v += (unsigned long long)(int64_t)z[u].v;
                        }
                    }
                }
            }
        };
        auto probe = [&](const SR(&z)[U], uint64_t b0) {
            if (b0 + (uint64_t)U * T <= se)
                probe_batch(std::true_type{}, z, (uint32_t)U);
            else
                probe_batch(std::false_type{}, z, rows_in(b0));
        };
        SR y[U];
        for (uint64_t b0 = sb; b0 < se; b0 += 2ull * U * T) {
            fetch_clamped(y, b0 + (uint64_t)U * T);
            probe(x, b0);
            fetch_clamped(x, b0 + 2ull * U * T);
            if (b0 + (uint64_t)U * T < se) probe(y, b0 + (uint64_t)U * T);
        }
        const unsigned long long spc = sp_cnt, sps = sp_sum;
        acc_c += pe_n * spc;
        acc_b += pe_n * sps;
        acc_p += pe_v * spc;
        acc_c = wave_sum(acc_c);
        acc_b = wave_sum(acc_b);
        acc_p = wave_sum(acc_p);
        const int lane = tid & 63, w = tid >> 6;
        if (lane == 0) {
            red[0][w] = acc_c;
            red[1][w] = acc_b;
            red[2][w] = acc_p;
        }
        __syncthreads();
        if (tid == 0) {
            unsigned long long cc = 0, bb = 0, pp = 0;
            for (int k = 0; k < T / 64; ++k) {
                cc += red[0][k];
                bb += red[1][k];
                pp += red[2][k];
            }
            unsigned long long* o = res + 4 * (uint64_t)j;
            if (cc) {
                atomicAdd(&o[0], cc);
                atomicAdd(&o[1], bb);
                atomicAdd(&o[2], pp);
            }
        }
        __syncthreads();  // red and the shared item words are free for the next item
    }
}

__global__ __launch_bounds__(256) void k_bzero(unsigned long long* __restrict__ p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) p[i] = 0;
}

template <typename KT, bool HVS, int CAP, int T>
int launch_join_t(hipStream_t st, int grid, const Job* jobs, const uint32_t* item_job, uint32_t n_items,
                  const void* R, const void* S, const uint32_t* bndR, const uint32_t* bndS, uint32_t* queue,
                  unsigned long long* res) {
    hipLaunchKernelGGL((k_bjoin<KT, HVS, CAP, T>), dim3(grid), dim3(T), 0, st, jobs, item_job, n_items,
                       (const Row<KT, true>*)R, (const Row<KT, HVS>*)S, bndR, bndS, queue, res, nullptr, nullptr);
    PGPU_LAUNCH_CHECK();
    return PGPU_OK;
}

// Variant 3: the unique-key kernel over every item, then variant 2's kernel
// over the items it handed back (one workgroup per CU; its count is read on
// the device).  queue[0..2]: item queue, handed-back count, their queue.
template <bool HVS>
int launch_join8_t(hipStream_t st, int grid, const Job* jobs, const uint32_t* item_job, uint32_t n_items,
                   const void* R, const void* S, const uint32_t* bndR, const uint32_t* bndS, uint32_t* queue,
                   uint32_t* redo, unsigned long long* res) {
    int dev = 0, cus = 0;
    PGPU_HIP(hipGetDevice(&dev));
    PGPU_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    hipLaunchKernelGGL(k_bjoin8<HVS>, dim3(grid), dim3(kT8), 0, st, jobs, item_job, n_items,
                       (const Row<int32_t, true>*)R, (const Row<int32_t, HVS>*)S, bndR, bndS, queue, redo, res);
    PGPU_LAUNCH_CHECK();
    hipLaunchKernelGGL((k_bjoin<int32_t, HVS, 8192, 1024>), dim3(std::min<uint32_t>(cus, n_items)), dim3(1024), 0, st,
                       jobs, item_job, n_items, (const Row<int32_t, true>*)R, (const Row<int32_t, HVS>*)S, bndR, bndS,
                       queue + 2, res, redo, queue + 1);
    PGPU_LAUNCH_CHECK();
    return PGPU_OK;
}

template <typename KT, bool HVS, int CAP, int T>
int occupancy_t(int* blocks) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, (const void*)k_bjoin<KT, HVS, CAP, T>, T, 0) ==
                   hipSuccess
               ? PGPU_OK
               : pgpu::fail(PGPU_EDEVICE, "occupancy query failed");
}

}  // namespace

namespace pgpu {
namespace batch {

void Layout::add(const void* key, const void* pay, uint64_t n, uint32_t bits, uint32_t chunk_base,
                 uint32_t chunk_rows) {
    Seg s{};
    s.key = key;
    s.pay = pay;
    s.n = n;
    s.dst = rows;
    s.hist = hist;
    s.bnd = bnd;
    s.bits = bits;
    s.chunk_rows = chunk_rows;
    s.chunk0 = chunk_base + (uint32_t)chunks.size();
    s.nchunks = (uint32_t)ceil_div(n, chunk_rows);
    for (uint32_t c = 0; c < s.nchunks; ++c) chunks.push_back(Chunk{(uint32_t)segs.size(), 0, (uint64_t)c * chunk_rows});
    rows += n;
    hist += (uint64_t)s.nchunks << bits;
    bnd += (1ull << bits) + 1;
    segs.push_back(s);
}

int variant_piece(int variant) { return variant == 1 ? 2048 : 4096; }
uint32_t variant_target(int variant) { return variant == 1 ? 1536 : 3072; }
bool variant_ok(int variant, int key_bytes) { return variant == 1 || variant == 2 || (variant == 3 && key_bytes == 4); }

uint32_t bits_for(uint64_t n, uint32_t target) {
    if (n <= target) return 0;
    return std::min<uint32_t>(kMaxBits, ilog2_ceil(ceil_div(n, target)));
}

int launch_hist(hipStream_t st, int kb, const Seg* segs, const Chunk* chunks, uint32_t chunk_base, uint32_t n,
                uint32_t* H) {
    if (!n) return PGPU_OK;
    if (kb == 8)
        hipLaunchKernelGGL(k_bhist<int64_t>, dim3(n), dim3(512), 0, st, segs, chunks, chunk_base, H);
    else
        hipLaunchKernelGGL(k_bhist<int32_t>, dim3(n), dim3(512), 0, st, segs, chunks, chunk_base, H);
    PGPU_LAUNCH_CHECK();
    return PGPU_OK;
}

int launch_scan(hipStream_t st, uint32_t max_bits, const Seg* segs, uint32_t seg_base, uint32_t nsegs, uint32_t* H,
                uint32_t* bnd) {
    if (!nsegs) return PGPU_OK;
    if (max_bits > (uint32_t)kMaxBits) return fail(PGPU_EINVAL, "scan: %u partition bits (at most %d)", max_bits, kMaxBits);
    const uint32_t groups = (uint32_t)ceil_div(1u << max_bits, kColBins);
    hipLaunchKernelGGL(k_bcol, dim3(groups, nsegs), dim3(64 * kColWaves), 0, st, segs, seg_base, H, bnd);
    PGPU_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_bbins, dim3(nsegs), dim3(1024), 0, st, segs, seg_base, bnd);
    PGPU_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_bapply, dim3(groups, nsegs), dim3(64 * kColWaves), 0, st, segs, seg_base, H, bnd);
    PGPU_LAUNCH_CHECK();
    return PGPU_OK;
}

uint32_t scatter_tile_rows(int kb, bool hv, uint32_t max_bits) {
    const uint32_t esz = (uint32_t)kb * (hv ? 2u : 1u);
    const uint32_t tb = esz >= 8 && max_bits <= 10 ? 2 * kScatterTileBytes : kScatterTileBytes;
    return tb / esz;
}

template <typename KT, bool HV>
void launch_scatter_t(hipStream_t st, uint32_t max_bits, const Seg* segs, const Chunk* chunks, uint32_t chunk_base,
                      uint32_t n, const uint32_t* H, void* out) {
    // 4-byte rows already hold 16K rows in the 64 KB tile
    if constexpr (sizeof(Row<KT, HV>) >= 8) {
        if (max_bits <= 10) {
            hipLaunchKernelGGL((k_bscatter<KT, HV, 2 * kScatterTileBytes, 1024, 1024>), dim3(n), dim3(1024), 0, st,
                               segs, chunks, chunk_base, H, (Row<KT, HV>*)out);
            return;
        }
    }
    hipLaunchKernelGGL((k_bscatter<KT, HV, kScatterTileBytes, kND, 1024>), dim3(n), dim3(1024), 0, st, segs, chunks,
                       chunk_base, H, (Row<KT, HV>*)out);
}

int launch_scatter(hipStream_t st, int kb, bool hv, uint32_t max_bits, const Seg* segs, const Chunk* chunks,
                   uint32_t chunk_base, uint32_t n, const uint32_t* H, void* out) {
    if (!n) return PGPU_OK;
    if (max_bits > (uint32_t)kMaxBits) return fail(PGPU_EINVAL, "scatter: %u partition bits (at most %d)", max_bits, kMaxBits);
    if (kb == 8) {
        if (hv)
            launch_scatter_t<int64_t, true>(st, max_bits, segs, chunks, chunk_base, n, H, out);
        else
            launch_scatter_t<int64_t, false>(st, max_bits, segs, chunks, chunk_base, n, H, out);
    } else {
        if (hv)
            launch_scatter_t<int32_t, true>(st, max_bits, segs, chunks, chunk_base, n, H, out);
        else
            launch_scatter_t<int32_t, false>(st, max_bits, segs, chunks, chunk_base, n, H, out);
    }
    PGPU_LAUNCH_CHECK();
    return PGPU_OK;
}

#define PGPU_JOIN_VARIANTS(X)                               \
    X(int64_t, true, 1, 4096, 512) X(int64_t, true, 2, 8192, 1024)   \
    X(int64_t, false, 1, 4096, 512) X(int64_t, false, 2, 8192, 1024) \
    X(int32_t, true, 1, 4096, 512) X(int32_t, true, 2, 8192, 1024)   \
    X(int32_t, false, 1, 4096, 512) X(int32_t, false, 2, 8192, 1024)

int launch_join(hipStream_t st, int kb, bool hvs, int variant, int grid, const Job* jobs, const uint32_t* item_job,
                uint32_t n_items, const void* R, const void* S, const uint32_t* bndR, const uint32_t* bndS,
                uint32_t* queue, unsigned long long* res, uint32_t* redo) {
    if (!n_items) return PGPU_OK;
    if (variant == 3) {
        if (kb != 4 || !redo) return fail(PGPU_EINVAL, "join variant 3 needs 4-byte keys and a hand-back list");
        return hvs ? launch_join8_t<true>(st, grid, jobs, item_job, n_items, R, S, bndR, bndS, queue, redo, res)
                   : launch_join8_t<false>(st, grid, jobs, item_job, n_items, R, S, bndR, bndS, queue, redo, res);
    }
#define X(KT, HV, V, CAP, T)                                                                                      \
    if ((kb == 8) == (sizeof(KT) == 8) && hvs == HV && variant == V)                                              \
        return launch_join_t<KT, HV, CAP, T>(st, grid, jobs, item_job, n_items, R, S, bndR, bndS, queue, res);
    PGPU_JOIN_VARIANTS(X)
#undef X
    return fail(PGPU_EINVAL, "unknown join variant %d", variant);
}

int join_grid(int kb, bool hvs, int variant, int* grid) {
    int dev = 0, cus = 0, blocks = 0;
    PGPU_HIP(hipGetDevice(&dev));
    PGPU_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (variant == 3 && kb == 4) {
        const void* f = hvs ? (const void*)k_bjoin8<true> : (const void*)k_bjoin8<false>;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, f, kT8, 0) != hipSuccess)
            return fail(PGPU_EDEVICE, "occupancy query failed");
        *grid = cus * std::max(blocks, 1);
        return PGPU_OK;
    }
#define X(KT, HV, V, CAP, T)                                             \
    if ((kb == 8) == (sizeof(KT) == 8) && hvs == HV && variant == V) { \
        PGPU_TRY((occupancy_t<KT, HV, CAP, T>(&blocks)));                \
        *grid = cus * std::max(blocks, 1);                               \
        return PGPU_OK;                                                  \
    }
    PGPU_JOIN_VARIANTS(X)
#undef X
    return fail(PGPU_EINVAL, "unknown join variant %d", variant);
}

int launch_zero(hipStream_t st, unsigned long long* p, uint64_t n) {
    if (!n) return PGPU_OK;
    const unsigned g = (unsigned)std::min<uint64_t>(ceil_div(n, 256), 1024);
    hipLaunchKernelGGL(k_bzero, dim3(g), dim3(256), 0, st, p, n);
    PGPU_LAUNCH_CHECK();
    return PGPU_OK;
}

}  // namespace batch
}  // namespace pgpu
