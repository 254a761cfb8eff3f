// Equi-join with a fused count + sum consumer on gfx950.
//
// Reference path replaced: HashJoinOp build (operators/join.cpp:396-464,
// :641-667) inserting every build tuple under a bucket spinlock into a
// chained-bucket HashTable (util/hashtable.cpp:230-273), then the probe loop
// (operators/join.cpp:503-562) walking the bucket chain and comparing keys
// through Comparator::eval, feeding AggregateCount / AggregateSum
// (operators/aggregatecount.cpp:37-52, operators/aggregatesum.cpp:81-107).
//
// Two device algorithms produce the same (count, sum) result:
//
// RADIX (2-pass radix partitioning, the MSD split of
// ParameterizedModuloValueHasher::generate(2), hash.cpp:237-267: pass 1 takes
// the high B1 = B/2 bits of the B-bit radix value, pass 2 the low B2 = B-B1):
//   k_hist      one read of each key column; per block of `blk` rows an LDS
//               histogram over all 2^B final partitions, written to H[blk][2^B],
//               and per pass-1 tile of the block the rows of each p1 in the
//               block's tiles before it (toff)
//   k_colscan   exclusive scan of H down the blocks; per-bin totals; per
//               (block, p1) the sum of the block's prefixes over p2, which
//               with base[] gives pass 1's write offsets (pass1_off)
//   k_binscan   exclusive scan of the totals -> base[] (partition starts)
//   k_scatter1  one workgroup per tile: reorders the tile by p1 in LDS and
//               writes contiguous digit runs at its exact offsets (no atomics);
//               consecutive tiles run side by side on one XCD
//   k_scatter2  pass-2 block (p1, b) = the rows block b wrote to p1; it
//               reorders them by p2 and writes runs at base(p1,p2)+H[b][p1,p2]
//   k_icount /  probe items: each final partition's probe rows are cut into
//   k_binscan   items of at most M rows, so one hot (Zipf) partition is spread
//               over many workgroups
//   k_join_cs   per item: the partition's build rows go into an LDS table
//               (open addressing, duplicates folded into (count, sum)), then
//               the item's probe rows are streamed from HBM and looked up
// Build partitions larger than the LDS table are handled in LDS-sized pieces
// (each piece re-probes the item's rows), so no build size overflows.
//
// GLOBAL (one open-addressing table in HBM, the "plain" per-query path):
//   k_gclear, k_gbuild (CAS insert), k_gprobe.
#include <algorithm>
#include <mutex>
#include <type_traits>
#include <cstring>

#include "batch_internal.h"
#include "join_common.h"
#include "pgpu_internal.h"

namespace {

using pgpu::dev::nt_load;
using pgpu::dev::nt_store;
using pgpu::dev::xcd_remap;
using pgpu::dev::xcd_remap_chunked;
using pgpu::dev::buf_rsrc;
using pgpu::dev::buf_load;

constexpr uint64_t kFib = 0x9E3779B97F4A7C15ull;
constexpr int kMaxBits = 14;
static_assert((kMaxBits + 1) / 2 <= 7, "one pass takes at most 7 bits (kDigits)");
constexpr int kScatterThreads = 512;
constexpr int kJoinCap = 4096;            // LDS table slots
constexpr int kJoinPiece = 2048;          // build rows per LDS piece (load <= 0.5)
constexpr uint32_t kItemRows = 32768;     // probe rows per join item
constexpr int kTileBytes = 77824;         // LDS staging tile of the pass-2 scatter (two workgroups per CU):
                                          // 9728 8-byte rows hold a whole unit of ~8K rows (with 64 KB,
                                          // half the units spilled a few rows into a second tile)
constexpr int kDigits = 128;              // digits of one pass: at most 7 of the <= 14 radix bits
constexpr uint64_t kForkMinProbe = 1ull << 26;  // probe rows from which the build side runs on a second stream
// One-pass route: one radix pass of at most 10 bits (partitions of ~4096
// build rows, 128 KB scatter tiles) and the batched-join LDS kernels.
constexpr uint32_t kOnePassPartRows = 4096;
constexpr uint64_t kOnePassMaxBuild = 1ull << 22;  // 1024 partitions
constexpr uint32_t kOnePassMaxItems = 1u << 10;
constexpr uint64_t kOpSeg = 0, kOpChunk = 256;     // table offsets in the one-pass device / staging block
constexpr int kTileBytes1 = 65536;        // pass-1 tile (one workgroup each, two per CU); a multiple of the
                                          // histogram's 4096-row step for rows of at most 16 bytes
constexpr int kPartSlots = 2048;          // partial-result slots of the join kernels
static_assert(kJoinPiece < 65536, "LDS match counters are 16 bits wide");
static_assert(2 * kJoinPiece <= kJoinCap, "LDS table load factor above 0.5");

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

template <typename KT, bool HV>
struct Elem;
template <typename KT>
struct Elem<KT, true> {
    KT k;
    KT v;
};
template <typename KT>
struct Elem<KT, false> {
    KT k;
};

// ---------------------------------------------------------------- helpers
// Exclusive scan of in[0..n) (n <= 256) by one wavefront; out[n] = total.
__device__ __forceinline__ void wave_excl_scan256(const uint32_t* in, uint32_t* out, int n) {
    const int lane = threadIdx.x & 63;
    uint32_t a[4];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int idx = lane * 4 + i;
        a[i] = idx < n ? in[idx] : 0u;
        s += a[i];
    }
    uint32_t x = s;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    uint32_t ex = x - s;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int idx = lane * 4 + i;
        if (idx < n) out[idx] = ex;
        ex += a[i];
    }
    if (lane == 63) out[n] = x;
}

// Exclusive scan across a 1024-thread block of one uint64 per thread.
__device__ __forceinline__ uint64_t block_excl_scan1024(uint64_t v, uint64_t* total) {
    __shared__ uint64_t ws[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint64_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint64_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) ws[w] = x;
    __syncthreads();
    if (w == 0) {
        uint64_t t = lane < 16 ? ws[lane] : 0;
        uint64_t inc = t;
#pragma unroll
        for (int d = 1; d < 16; d <<= 1) {
            uint64_t y = __shfl_up(inc, d, 64);
            if (lane >= d) inc += y;
        }
        if (lane < 16) ws[lane] = inc - t;
        if (lane == 15 && total) *total = inc;
    }
    __syncthreads();
    return ws[w] + (x - v);
}

// ---------------------------------------------------------------- radix kernels
// Block b's rows are pass-1 tiles of T1 rows (blk is a multiple of T1).
// Besides H[b], toff[t][p1] receives, for each tile t of the block, the
// block's rows of partition p1 in its tiles before t.  Each wavefront counts
// whole tiles (tiles w, w + 8, ... of the block) in its own LDS row, so the
// sweep needs no barrier; the counts go to toff as the wavefront finishes a
// tile and become in-block prefixes after the sweep.  Dynamic LDS: 2^B words.
template <typename KT>
__global__ __launch_bounds__(512) void k_hist(const KT* __restrict__ keys, uint64_t n, uint64_t blk, uint32_t T1,
                                              int B, int B2, uint32_t* __restrict__ H, uint32_t* __restrict__ toff) {
    extern __shared__ uint32_t lh[];
    __shared__ uint32_t lt[8 * kDigits];  // [wavefront][p1] counts of the tile being swept
    const uint32_t nb = 1u << B, nb1 = 1u << (B - B2);
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    for (uint32_t i = tid; i < nb; i += 512) lh[i] = 0;
    for (uint32_t i = tid; i < 8 * kDigits; i += 512) lt[i] = 0;
    __syncthreads();
    const uint64_t s = (uint64_t)blockIdx.x * blk;
    const uint64_t e = min(n, s + blk);
    constexpr int U = 8;
    uint32_t* lw = lt + w * kDigits;
    for (uint64_t t0 = s + (uint64_t)w * T1; t0 < e; t0 += 8ull * T1) {
        const uint64_t te = min(e, t0 + T1);
        uint64_t i = t0 + lane;
        for (; i + (U - 1) * 64 < te; i += U * 64) {
            KT k[U];
#pragma unroll
            for (int u = 0; u < U; ++u) k[u] = nt_load(&keys[i + u * 64]);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t d = (uint32_t)(hmix(k[u]) >> (64 - B));
                atomicAdd(&lh[d], 1u);
                atomicAdd(&lw[d >> B2], 1u);
            }
        }
        for (; i < te; i += 64) {
            const uint32_t d = (uint32_t)(hmix(keys[i]) >> (64 - B));
            atomicAdd(&lh[d], 1u);
            atomicAdd(&lw[d >> B2], 1u);
        }
        // one wavefront's LDS operations complete in order: its counts are final
        uint32_t* tc = toff + (t0 / T1) * kDigits;
        for (uint32_t d = lane; d < nb1; d += 64) {
            tc[d] = lw[d];
            lw[d] = 0;
        }
    }
    __syncthreads();  // every tile's counts are in toff
    if ((uint32_t)tid < nb1 && e > s) {
        const uint64_t t1 = (e + T1 - 1) / T1;
        uint32_t run = 0;
#pragma unroll 8
        for (uint64_t t = s / T1; t < t1; ++t) {
            const uint32_t c = toff[t * kDigits + tid];
            toff[t * kDigits + tid] = run;
            run += c;
        }
    }
    uint32_t* Hb = H + (uint64_t)blockIdx.x * nb;
    for (uint32_t j = tid; j < nb; j += 512) Hb[j] = lh[j];
}

// One workgroup per 64 bins; its 8 wavefronts scan 8 contiguous segments of
// blocks, then each bin's column becomes its exclusive prefix over blocks.
// Each lane reads its segment's counts once, all loads in flight together
// (at most kMaxBlocks / 8 rows), and writes the prefixes from registers.
// px[sub][b][p1] receives the sum of block b's prefixes over this
// workgroup's bins of partition p1 (sub = which 64-bin half of p1 when p1 has
// 128 bins): the rows of p1 in blocks before b, so pass 1 needs no offset
// kernel of its own.
constexpr uint32_t kMaxBlocks = 512;  // histogram blocks of one side
__global__ __launch_bounds__(512) void k_colscan(uint32_t* __restrict__ H, uint32_t nblk, uint32_t nb, int B2,
                                                 uint32_t* __restrict__ tot, uint32_t* __restrict__ px) {
    constexpr int kSeg = kMaxBlocks / 8;
    __shared__ uint32_t part[8][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t bin = blockIdx.x * 64 + lane;
    const uint32_t seg = (nblk + 7) / 8;
    const uint32_t b0 = min(nblk, w * seg), nseg = min(nblk, b0 + seg) - b0;
    const bool live = bin < nb;
    const uint32_t nb2 = 1u << B2, seg2 = min(nb2, 64u);  // bins of one p1 within these 64
    const uint32_t p1 = bin >> B2, sub = nb2 > 64 ? (bin >> 6) & 1 : 0;
    uint32_t h[kSeg];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kSeg; ++j) {
        h[j] = live && (uint32_t)j < nseg ? H[(uint64_t)(b0 + j) * nb + bin] : 0u;
        s += h[j];
    }
    part[w][lane] = s;
    __syncthreads();
    uint32_t run = 0;
    for (int k = 0; k < w; ++k) run += part[k][lane];
#pragma unroll
    for (int j = 0; j < kSeg; ++j) {
        if ((uint32_t)j < nseg) {  // uniform across the wavefront: the shuffles see every lane
            if (live) H[(uint64_t)(b0 + j) * nb + bin] = run;
            uint32_t ps = live ? run : 0u;
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1)  // unrolled: the rows' shuffles overlap
                if ((uint32_t)d < seg2) ps += __shfl_xor(ps, d, 64);
            if (live && (lane & (seg2 - 1)) == 0) px[((uint64_t)sub * kMaxBlocks + b0 + j) * kDigits + p1] = ps;
        }
        run += h[j];
    }
    if (live && w == 7) tot[bin] = run;
}

// Pass-1 write offset of block b's rows of partition p1 (b == nblk: the
// partition's end): base(p1, 0) plus the rows of p1 in blocks before b.
__device__ __forceinline__ uint64_t pass1_off(const uint64_t* __restrict__ base, const uint32_t* __restrict__ px,
                                              uint32_t b, uint32_t nblk, uint32_t p1, int B2) {
    if (b == nblk) return base[(uint64_t)(p1 + 1) << B2];
    uint64_t o = base[(uint64_t)p1 << B2] + px[(uint64_t)b * kDigits + p1];
    if (B2 > 6) o += px[((uint64_t)kMaxBlocks + b) * kDigits + p1];
    return o;
}

// base[i] = sum of in[0..i) for i <= n (n <= 16384, sums below 2^32), one
// workgroup: the counts are staged in LDS with coalesced loads, each thread
// scans 16 consecutive ones, and the prefixes go out with coalesced stores.
__global__ __launch_bounds__(1024) void k_binscan(const uint32_t* __restrict__ in, uint32_t n,
                                                  uint64_t* __restrict__ base) {
    constexpr int PER = 16, N = 1024 * PER;
    __shared__ uint32_t v[N + N / 32];  // element i at i + i / 32: the 16-element strips read without conflicts
    const int t = threadIdx.x;
    for (int i = t; i < N; i += 1024) v[i + (i >> 5)] = (uint32_t)i < n ? in[i] : 0u;
    __syncthreads();
    uint32_t a[PER];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int i = t * PER + k;
        a[k] = v[i + (i >> 5)];
        s += a[k];
    }
    __shared__ uint64_t tot_sh;
    uint32_t ex = (uint32_t)block_excl_scan1024(s, &tot_sh);
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int i = t * PER + k;
        v[i + (i >> 5)] = ex;
        ex += a[k];
    }
    __syncthreads();
    for (int i = t; (uint32_t)i < n; i += 1024) base[i] = v[i + (i >> 5)];
    if (t == 0) base[n] = tot_sh;
}

// Tile t (rows [t*T, t*T + T), inside block b = t / ntb) goes to
// pass1_off(b, p1) + toff[t][p1] for each of its digit runs.
template <typename KT, bool HV, int IPT>
__global__ __launch_bounds__(kScatterThreads, 4) void k_scatter1(const KT* __restrict__ keys, const KT* __restrict__ vals,
                                                                 uint64_t n, uint32_t ntb, uint32_t nblk, int B, int B2,
                                                                 const uint64_t* __restrict__ base,
                                                                 const uint32_t* __restrict__ px,
                                                                 const uint32_t* __restrict__ toff,
                                                                 Elem<KT, HV>* __restrict__ out) {
    using E = Elem<KT, HV>;
    constexpr int T = kScatterThreads * IPT;
    __shared__ E stage[T];
    __shared__ uint32_t hist[kDigits];
    __shared__ uint32_t lofs[kDigits + 1];
    __shared__ uint64_t run[kDigits];
    const int tid = threadIdx.x;
    const int nb1 = 1 << (B - B2);
    // tiles t and t+1 write adjacent runs of every p1: neighbours share an XCD
    // (in chunks of 32 tiles), so the L2 merges the line they share
    const uint32_t t = xcd_remap_chunked<32>(blockIdx.x, gridDim.x);
    const uint64_t s = (uint64_t)t * T;
    const uint32_t cnt = (uint32_t)min<uint64_t>(T, n - s);
    // the tile's columns through buffer resources (offsets below T rows)
    const __amdgpu_buffer_rsrc_t rk = buf_rsrc(keys + s, cnt * (uint32_t)sizeof(KT));
    const __amdgpu_buffer_rsrc_t rv = buf_rsrc(HV ? vals + s : keys + s, cnt * (uint32_t)sizeof(KT));
    E el[IPT];
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        const uint32_t idx = tid + k * kScatterThreads;
        if (idx < cnt) {
            el[k].k = buf_load<KT>(rk, idx * (uint32_t)sizeof(KT));
            if constexpr (HV) el[k].v = buf_load<KT>(rv, idx * (uint32_t)sizeof(KT));
        }
    }
    for (int d = tid; d < nb1; d += kScatterThreads) {
        run[d] = pass1_off(base, px, t / ntb, nblk, (uint32_t)d, B2) + toff[(uint64_t)t * kDigits + d];
        hist[d] = 0;
    }
    __syncthreads();
    uint32_t dg[IPT], rk_[IPT];
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        const uint32_t idx = tid + k * kScatterThreads;
        dg[k] = ~0u;  // rows past the tile's end
        if (idx < cnt) {
            dg[k] = (uint32_t)(hmix(el[k].k) >> (64 - B)) >> B2;
            rk_[k] = atomicAdd(&hist[dg[k]], 1u);
        }
    }
    __syncthreads();
    if (tid < 64) wave_excl_scan256(hist, lofs, nb1);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < IPT; ++k)
        if (dg[k] != ~0u) stage[lofs[dg[k]] + rk_[k]] = el[k];
    if (tid < nb1) run[tid] -= lofs[tid];  // write base: staged row j of digit d goes to run[d] + j
    __syncthreads();
    for (uint32_t j = tid; j < cnt; j += kScatterThreads) {
        const E x = stage[j];
        const uint32_t d = (uint32_t)(hmix(x.k) >> (64 - B)) >> B2;
        out[run[d] + j] = x;
    }
}

// A pass-2 unit is partition p1's rows from pass-1 blocks [b0, b0+G): they
// are contiguous in the pass-1 output (block b wrote p1's rows at pass1_off(b, p1)),
// and the units' output runs for (p1, p2) are contiguous too, starting at
// base(p1,p2) + H[b0][p1,p2].  G > 1 keeps units near a tile when blocks
// are small.
// Two workgroups per CU where the registers allow it (pass-2 units are short:
// the second workgroup hides one unit's start-up and tail).
template <typename KT, bool HV, int IPT>
__global__ __launch_bounds__(kScatterThreads, IPT <= 20 ? 4 : 1) void k_scatter2(const Elem<KT, HV>* __restrict__ in, uint32_t nblk,
                                                              uint32_t G, int B, int B2,
                                                              const uint32_t* __restrict__ px,
                                                              const uint32_t* __restrict__ Hs,
                                                              const uint64_t* __restrict__ base,
                                                              Elem<KT, HV>* __restrict__ out) {
    using E = Elem<KT, HV>;
    constexpr int T = kScatterThreads * IPT;
    __shared__ E stage[T];
    __shared__ uint32_t hist[kDigits];
    __shared__ uint32_t lofs[kDigits + 1];
    __shared__ uint64_t run[kDigits];
    const int tid = threadIdx.x;
    const uint32_t nb1 = 1u << (B - B2), nb2 = 1u << B2, nb = 1u << B;
    const uint32_t ngrp = (nblk + G - 1) / G;
    // units (p1, b0) and (p1, b0 + G) write adjacent runs of every final
    // partition: neighbours share an XCD, so the L2 merges the line they share
    // (in chunks of 32 units: a hot partition's units still cover every XCD)
    const uint32_t u = xcd_remap_chunked<32>(blockIdx.x, gridDim.x);
    const uint32_t p1 = u / ngrp, b0 = (u % ngrp) * G, b1 = min(nblk, b0 + G);
    const uint64_t s0 = pass1_off(base, px, b0, nblk, p1, B2);
    const uint64_t e = pass1_off(base, px, b1, nblk, p1, B2) - s0;  // the unit's rows: [s0, s0 + e)
    if (e == 0) return;
    const uint32_t mask2 = nb2 - 1;
    static_assert(kScatterThreads >= kDigits, "one thread per pass-2 digit");
    // The run starts and the unit's first tile are read together: the tile
    // does not wait for the offsets.
    uint64_t rbase = 0;
    if ((uint32_t)tid < nb2) {
        const uint64_t pp = ((uint64_t)p1 << B2) + tid;
        rbase = base[pp] + Hs[(uint64_t)b0 * nb + pp];
    }
    // the unit spans at most max(block rows, T * 2^7) rows: far below 4 GB
    const __amdgpu_buffer_rsrc_t src = buf_rsrc(in + s0, (uint32_t)(e * sizeof(E)));
    E el[IPT];
    auto load = [&](uint64_t t0, uint32_t cnt) {
#pragma unroll
        for (int k = 0; k < IPT; ++k) {
            const uint32_t idx = tid + k * kScatterThreads;
            if (idx < cnt) el[k] = buf_load<E>(src, (uint32_t)(t0 + idx) * (uint32_t)sizeof(E));
        }
    };
    uint32_t cnt = (uint32_t)min<uint64_t>(T, e);
    load(0, cnt);
    if ((uint32_t)tid < nb2) {
        run[tid] = rbase;
        hist[tid] = 0;
    }
    __syncthreads();
    for (uint64_t t0 = 0; t0 < e;) {
        uint32_t dg[IPT], rk[IPT];
#pragma unroll
        for (int k = 0; k < IPT; ++k) {
            const uint32_t idx = tid + k * kScatterThreads;
            dg[k] = ~0u;  // rows past the tile's end
            if (idx < cnt) {
                dg[k] = (uint32_t)(hmix(el[k].k) >> (64 - B)) & mask2;
                rk[k] = atomicAdd(&hist[dg[k]], 1u);
            }
        }
        __syncthreads();
        if (tid < 64) wave_excl_scan256(hist, lofs, (int)nb2);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < IPT; ++k)
            if (dg[k] != ~0u) stage[lofs[dg[k]] + rk[k]] = el[k];
        if ((uint32_t)tid < nb2) run[tid] -= lofs[tid];  // write base: staged row j of digit d goes to run[d] + j
        __syncthreads();
        const uint64_t next = t0 + cnt;
        const uint32_t ncnt = next < e ? (uint32_t)min<uint64_t>(T, e - next) : 0u;
        if (ncnt) load(next, ncnt);
        for (uint32_t j = tid; j < cnt; j += kScatterThreads) {
            const E x = stage[j];
            const uint32_t d = (uint32_t)(hmix(x.k) >> (64 - B)) & mask2;
            out[run[d] + j] = x;
        }
        __syncthreads();
        for (uint32_t d = tid; d < nb2; d += kScatterThreads) {
            run[d] += lofs[d + 1];  // = old run + hist
            hist[d] = 0;
        }
        __syncthreads();
        t0 = next;
        cnt = ncnt;
    }
}

__global__ __launch_bounds__(256) void k_icount(const uint32_t* __restrict__ totS, const uint32_t* __restrict__ totR,
                                                uint32_t nb, uint32_t M, uint32_t* __restrict__ icount) {
    const uint32_t pp = blockIdx.x * 256 + threadIdx.x;
    if (pp >= nb) return;
    icount[pp] = totR[pp] ? (totS[pp] + M - 1) / M : 0u;
}

// One join work item: a probe sub-range of one final partition and that
// partition's build range, read by the join workgroup with one 32-byte load.
struct __align__(16) JoinItem {
    uint64_t sB;   // first probe row
    uint64_t rB;   // first build row
    uint32_t sN;   // probe rows
    uint32_t nR;   // build rows
    uint64_t pad;
};

__global__ __launch_bounds__(256) void k_items(const uint32_t* __restrict__ icount, const uint64_t* __restrict__ ioff,
                                               const uint64_t* __restrict__ baseS, const uint32_t* __restrict__ totS,
                                               const uint64_t* __restrict__ baseR, const uint32_t* __restrict__ totR,
                                               uint32_t nb, uint32_t M, JoinItem* __restrict__ items) {
    const uint32_t pp = blockIdx.x * 256 + threadIdx.x;
    if (pp >= nb) return;
    const uint64_t o = ioff[pp];
    const uint64_t s0 = baseS[pp], ns = totS[pp];
    for (uint32_t i = 0; i < icount[pp]; ++i) {
        JoinItem it;
        it.sB = s0 + (uint64_t)i * M;
        it.sN = (uint32_t)min<uint64_t>(M, ns - (uint64_t)i * M);
        it.rB = baseR[pp];
        it.nR = totR[pp];
        it.pad = 0;
        items[o + i] = it;
    }
}

// Folds the per-slot partial results into result[0..3].
__global__ __launch_bounds__(256) void k_reduce_partials(const unsigned long long* __restrict__ part, int nslots,
                                                         unsigned long long* __restrict__ result) {
    unsigned long long c = 0, b = 0, p = 0, m = 0;
    for (int i = threadIdx.x; i < nslots; i += 256) {
        c += part[4 * i];
        b += part[4 * i + 1];
        p += part[4 * i + 2];
        m = max(m, part[4 * i + 3]);
    }
    __shared__ unsigned long long red[4][256];
    red[0][threadIdx.x] = c;
    red[1][threadIdx.x] = b;
    red[2][threadIdx.x] = p;
    red[3][threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
            red[2][threadIdx.x] += red[2][threadIdx.x + w];
            red[3][threadIdx.x] = max(red[3][threadIdx.x], red[3][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        result[0] = red[0][0];
        result[1] = red[1][0];
        result[2] = red[2][0];
        result[3] = red[3][0] > 1 ? red[3][0] : 1;  // LDS passes per partition
    }
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// LDS table slot: key and folded payload sum read with one 16-byte access.
struct __align__(16) LSlot {
    long long key;
    unsigned long long sum;
};

constexpr int log2i(int v) { return v <= 1 ? 0 : 1 + log2i(v / 2); }

// CAP LDS slots, T threads, PIECE build rows per LDS piece, UB probe rows of
// 8 B prefetched per thread, WPE waves per SIMD the register budget allows.
template <typename KT, bool HVS, int CAP, int T, int PIECE, int UB, int WPE>
__global__ __launch_bounds__(T, WPE) void k_join_cs(
    const Elem<KT, true>* __restrict__ R2, const uint64_t* __restrict__ baseR, const uint32_t* __restrict__ totR,
    const Elem<KT, HVS>* __restrict__ S2, const uint64_t* __restrict__ baseS, const uint32_t* __restrict__ totS,
    const uint64_t* __restrict__ ioff, const JoinItem* __restrict__ items, uint32_t nb, int B,
    unsigned long long* __restrict__ part) {
    constexpr int L = log2i(CAP);
    static_assert((1 << L) == CAP, "CAP");
    static_assert(PIECE < 65536 && PIECE <= CAP * 5 / 8, "LDS piece");
    constexpr int kJoinThreads = T;
    constexpr int kJoinPiece = PIECE;
    // Keys are held sign-extended to 64 bits; only an int64 key equal to the
    // empty marker needs the side accumulator.
    constexpr long long EMPTY = INT64_MIN;
    __shared__ LSlot tab[CAP];
    // match counts, two 16-bit counters per word (a piece holds at most
    // kJoinPiece < 2^16 rows, so a counter cannot carry into its neighbour)
    __shared__ uint32_t tc[CAP / 2];
    __shared__ unsigned long long sp_cnt, sp_sum;
    __shared__ unsigned long long red[3][kJoinThreads / 64];
    const int tid = threadIdx.x;
    const uint64_t item = blockIdx.x;
    if (item >= ioff[nb]) return;
    const JoinItem it = items[item];
    const uint32_t sN = it.sN, nR = it.nR;
    const Elem<KT, HVS>* __restrict__ S = S2 + it.sB;
    const Elem<KT, true>* __restrict__ R = R2 + it.rB;
    // LDS slot = bits [64-B-L, 64-B) of hmix(k): they lie in the high word of
    // the 64-bit product (B + L <= 32), computed without its low half.
    const int sh = 32 - B - L;
    auto slot_of = [&](KT key) -> uint32_t {
        const uint64_t kk = (uint64_t)(int64_t)key;
        const uint32_t lo = (uint32_t)kk, hi = (uint32_t)(kk >> 32);
        const uint32_t h = __umulhi(lo, (uint32_t)kFib) + lo * (uint32_t)(kFib >> 32) + hi * (uint32_t)kFib;
        return (h >> sh) & (CAP - 1);
    };
    unsigned long long acc_c = 0, acc_b = 0, acc_p = 0;
    // Probe rows held in registers per thread (one batch = U x T rows).
    constexpr int U = HVS ? UB / 2 : UB;
    constexpr uint32_t BATCH = (uint32_t)U * kJoinThreads;
    Elem<KT, HVS> x[U] = {};
    // rows of the batch at `rb` this thread holds: tid + u*T < sN - rb
    auto rows_in = [&](uint32_t rb) -> uint32_t {
        const uint32_t r0 = rb + tid;
        return r0 < sN ? min<uint32_t>(U, (sN - r0 + kJoinThreads - 1) / kJoinThreads) : 0u;
    };
    // The item's probe rows (at most kItemRows) through a buffer resource:
    // 32-bit offsets, and rows past the item's end read as zero.
    using PE = Elem<KT, HVS>;
    const __amdgpu_buffer_rsrc_t srs = buf_rsrc(S, sN * (uint32_t)sizeof(PE));
    auto fetch = [&](uint32_t rb) {
        const uint32_t cnt = rows_in(rb);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if ((uint32_t)u < cnt) x[u] = buf_load<PE, true>(srs, (rb + tid + u * kJoinThreads) * (uint32_t)sizeof(PE));
    };
    // Unpredicated fetch of the batch at `rb` into z: rows past the item's
    // end read zero and are masked by the probe.
    auto fetch_clamped = [&](PE(&z)[U], uint32_t rb) {
#pragma unroll
        for (int u = 0; u < U; ++u) z[u] = buf_load<PE, true>(srs, (rb + tid + u * kJoinThreads) * (uint32_t)sizeof(PE));
    };
    constexpr int RB = (kJoinPiece + kJoinThreads - 1) / kJoinThreads;  // build rows per thread per piece
    Elem<KT, true> br[RB];
    auto fetch_build = [&](uint32_t p0) {
        const uint32_t pn = min<uint32_t>(kJoinPiece, nR - p0);
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            const uint32_t i = tid + b * kJoinThreads;
            if (i < pn) br[b] = nt_load(&R[p0 + i]);
        }
    };
    // Neither the build rows nor the first probe batch depend on the table:
    // both are in flight while the table is cleared.
    if (nR) {
        fetch_build(0);
        fetch(0);
    }
    const unsigned short* tc16 = (const unsigned short*)tc;  // little-endian halves of tc[]

    for (uint32_t p0 = 0; p0 < nR; p0 += kJoinPiece) {
        if (p0) fetch_build(p0);  // later pieces (skewed build partitions) fetch their rows here
        for (int i = tid; i < CAP; i += kJoinThreads) {
            tab[i].key = EMPTY;
            tab[i].sum = 0;
            if (i < CAP / 2) tc[i] = 0;
        }
        if (tid == 0) sp_cnt = sp_sum = 0;
        __syncthreads();
        const uint32_t pn = min<uint32_t>(kJoinPiece, nR - p0);
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            if ((uint32_t)(tid + b * kJoinThreads) >= pn) continue;
            const Elem<KT, true> r = br[b];
            const long long k = (long long)r.k;
            if (k == EMPTY) {
                atomicAdd(&sp_cnt, 1ull);
                atomicAdd(&sp_sum, (unsigned long long)(int64_t)r.v);
                continue;
            }
            uint32_t slot = slot_of(r.k);
            while (true) {
                const long long prev = (long long)atomicCAS((unsigned long long*)&tab[slot].key,
                                                            (unsigned long long)EMPTY, (unsigned long long)k);
                if (prev == EMPTY || prev == k) {
                    atomicAdd(&tc[slot >> 1], 1u << (16 * (slot & 1)));
                    atomicAdd(&tab[slot].sum, (unsigned long long)(int64_t)r.v);
                    break;
                }
                slot = (slot + 1) & (CAP - 1);
            }
        }
        __syncthreads();
        if (p0) fetch(0);  // later pieces re-read the item's probe rows
        const unsigned long long spc = sp_cnt, sps = sp_sum;
        unsigned long long pe_n = 0, pe_v = 0;  // this piece's probe rows keyed by the empty marker (and payloads)
        // One batch: all U slot reads in flight at once, collision chains
        // walked only by the lanes that need them, selects instead of
        // branches.  FULL batches (every thread holds U rows) skip the
        // row-count masks.
        auto probe_batch = [&](auto full_tag, const Elem<KT, HVS>(&x)[U], uint32_t cnt) {
            constexpr bool FULL = decltype(full_tag)::value;
            uint32_t slot[U];
            LSlot t[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                slot[u] = slot_of(x[u].k);
                t[u] = tab[slot[u]];
            }
            bool any = false;
#pragma unroll
            for (int u = 0; u < U; ++u)
                any |= (FULL || (uint32_t)u < cnt) && t[u].key != (long long)x[u].k && t[u].key != EMPTY;
            if (any) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (!FULL && (uint32_t)u >= cnt) continue;
                    const long long k = (long long)x[u].k;
                    while (t[u].key != k && t[u].key != EMPTY) {
                        slot[u] = (slot[u] + 1) & (CAP - 1);
                        t[u] = tab[slot[u]];
                    }
                }
            }
            // Each walk ended on the key's slot or on an empty one, whose
            // count and sum are 0: no hit / miss select.  A probe key equal to
            // the empty marker also ends on an empty slot; those rows are
            // tallied and take the side accumulator once per piece.
            uint32_t bc = 0;  // fits 32 bits: U rows x PIECE matches
            unsigned long long bs = 0, bp = 0;
            bool em = false;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const bool live = FULL || (uint32_t)u < cnt;
                uint32_t c = tc16[slot[u]];
                unsigned long long sb = t[u].sum;
                if (!live) {
                    c = 0;
                    sb = 0;
                }
                bc += c;
                bs += sb;
                if constexpr (HVS) {
                    if constexpr (sizeof(KT) == 8)
                        bp += (unsigned long long)c * (unsigned long long)(int64_t)x[u].v;
                    else  // c <= PIECE < 2^31: one signed 32 x 32 -> 64-bit multiply-add
                        bp += (unsigned long long)((int64_t)(int32_t)c * (int64_t)x[u].v);
                }
                if constexpr (sizeof(KT) == 8)  // sign-extended int32 keys never equal it
                    em |= live && (long long)x[u].k == EMPTY;
            }
            acc_c += bc;
            acc_b += bs;
            acc_p += bp;
            if (sizeof(KT) == 8 && __any(em)) {  // rare: a wavefront holds a probe key equal to the empty marker
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if ((FULL || (uint32_t)u < cnt) && (long long)x[u].k == EMPTY) {
                        pe_n += 1;
                        if constexpr (HVS) pe_v += (unsigned long long)(int64_t)x[u].v;
                    }
                }
            }
        };
        auto probe = [&](const Elem<KT, HVS>(&z)[U], uint32_t rb) {
            if (rb + BATCH <= sN)
                probe_batch(std::true_type{}, z, (uint32_t)U);
            else
                probe_batch(std::false_type{}, z, rows_in(rb));
        };
        // Two batches in flight: the next batch's rows load while this one
        // probes the LDS table.
        Elem<KT, HVS> y[U];
        for (uint32_t rb = 0; rb < sN; rb += 2 * BATCH) {
            fetch_clamped(y, rb + BATCH);
            probe(x, rb);
            fetch_clamped(x, rb + 2 * BATCH);
            if (rb + BATCH < sN) probe(y, rb + BATCH);
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
        unsigned long long c = 0, b = 0, p = 0;
        for (int k = 0; k < kJoinThreads / 64; ++k) {
            c += red[0][k];
            b += red[1][k];
            p += red[2][k];
        }
        // spread over kPartSlots addresses: no same-address atomic chain
        unsigned long long* slot = part + 4 * (item % kPartSlots);
        if (c) {
            atomicAdd(&slot[0], c);
            atomicAdd(&slot[1], b);
            atomicAdd(&slot[2], p);
        }
        const unsigned long long pieces = (nR + kJoinPiece - 1) / kJoinPiece;
        if (pieces > 1) atomicMax(&slot[3], pieces);
    }
}

// ---------------------------------------------------------------- global table
template <typename KT>
struct GSlot {
    KT k;
    uint32_t c;
    unsigned long long s;
};

template <typename KT>
__global__ __launch_bounds__(256) void k_gclear(GSlot<KT>* __restrict__ t, uint64_t cap,
                                                unsigned long long* __restrict__ special) {
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < cap; i += stride) {
        t[i].k = KeyEmpty<KT>::v;
        t[i].c = 0;
        t[i].s = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x < 2) special[threadIdx.x] = 0;
}

template <typename KT>
__global__ __launch_bounds__(256) void k_gbuild(const KT* __restrict__ rk, const KT* __restrict__ rv, uint64_t n,
                                                GSlot<KT>* __restrict__ t, int logcap,
                                                unsigned long long* __restrict__ special) {
    constexpr KT EMPTY = KeyEmpty<KT>::v;
    const uint64_t mask = (1ull << logcap) - 1;
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
        const KT k = rk[i];
        const unsigned long long v = (unsigned long long)(int64_t)rv[i];
        if (k == EMPTY) {
            atomicAdd(&special[0], 1ull);
            atomicAdd(&special[1], v);
            continue;
        }
        uint64_t slot = (hmix(k) >> (64 - logcap)) & mask;
        while (true) {
            KT prev;
            if constexpr (sizeof(KT) == 8)
                prev = (KT)atomicCAS((unsigned long long*)&t[slot].k, (unsigned long long)EMPTY, (unsigned long long)k);
            else
                prev = (KT)atomicCAS((unsigned int*)&t[slot].k, (unsigned int)EMPTY, (unsigned int)k);
            if (prev == EMPTY || prev == k) {
                atomicAdd(&t[slot].c, 1u);
                atomicAdd(&t[slot].s, v);
                break;
            }
            slot = (slot + 1) & mask;
        }
    }
}

template <typename KT, bool HVS>
__global__ __launch_bounds__(256) void k_gprobe(const KT* __restrict__ sk, const KT* __restrict__ sv, uint64_t n,
                                                const GSlot<KT>* __restrict__ t, int logcap,
                                                const unsigned long long* __restrict__ special,
                                                unsigned long long* __restrict__ result) {
    constexpr KT EMPTY = KeyEmpty<KT>::v;
    const uint64_t mask = (1ull << logcap) - 1;
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    unsigned long long acc_c = 0, acc_b = 0, acc_p = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
        const KT k = sk[i];
        unsigned long long c = 0, sb = 0;
        if (k == EMPTY) {
            c = special[0];
            sb = special[1];
        } else {
            uint64_t slot = (hmix(k) >> (64 - logcap)) & mask;
            while (true) {
                const KT tk = t[slot].k;
                if (tk == k) {
                    c = t[slot].c;
                    sb = t[slot].s;
                    break;
                }
                if (tk == EMPTY) break;
                slot = (slot + 1) & mask;
            }
        }
        acc_c += c;
        acc_b += sb;
        if constexpr (HVS) acc_p += c * (unsigned long long)(int64_t)sv[i];
    }
    __shared__ unsigned long long red[3][4];
    acc_c = wave_sum(acc_c);
    acc_b = wave_sum(acc_b);
    acc_p = wave_sum(acc_p);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][w] = acc_c;
        red[1][w] = acc_b;
        red[2][w] = acc_p;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long c = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        unsigned long long b = red[1][0] + red[1][1] + red[1][2] + red[1][3];
        unsigned long long p = red[2][0] + red[2][1] + red[2][2] + red[2][3];
        unsigned long long* slot = result + 4 * (blockIdx.x % kPartSlots);
        if (c) {
            atomicAdd(&slot[0], c);
            atomicAdd(&slot[1], b);
            atomicAdd(&slot[2], p);
        }
    }
}

__global__ __launch_bounds__(256) void k_fill64(unsigned long long* __restrict__ p, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) p[i] = 0;
}

}  // namespace

// ====================================================================== plan
struct pgpu_join_plan {
    pgpu_ctx* ctx = nullptr;
    pgpu_join_cfg cfg{};
    uint64_t maxR = 0, maxS = 0;
    int bits = 0;           // fixed bits from cfg (0 = choose per call)
    uint32_t nbmax = 0;     // 2^maxbits allocated
    uint32_t nblkR = 0, nblkS = 0;  // allocated block rows
    void *R1 = nullptr, *R2 = nullptr, *S1 = nullptr, *S2 = nullptr;
    uint32_t *HR = nullptr, *HS = nullptr, *totR = nullptr, *totS = nullptr, *icount = nullptr;
    uint64_t *baseR = nullptr, *baseS = nullptr, *ioff = nullptr;
    uint32_t *pxR = nullptr, *pxS = nullptr;  // per-(block, p1) prefixes of k_colscan (2 x kMaxBlocks x kDigits)
    uint32_t *toffR = nullptr, *toffS = nullptr;  // per-(pass-1 tile, p1) offsets within the tile's block
    unsigned long long* result = nullptr;  // [4]: count, sum_build, sum_probe, max pieces
    // one-pass route (PGPU_ALGO_RADIX1, plans of at most kOnePassMaxBuild build
    // rows): the batched-join kernels over one build and one probe segment
    void* opdev = nullptr;                 // segs, chunks, job, item -> job, at kOnePass* offsets
    char* ophost = nullptr;                // pinned staging of the same tables
    uint32_t* opH = nullptr;               // per-chunk histograms
    uint32_t* opbnd = nullptr;             // partition bounds of both sides
    unsigned long long* opres = nullptr;   // [4] result words, [2] queue words
    uint32_t* opredo = nullptr;            // partitions handed back by the unique-key join
    uint64_t opchunks = 0;                 // chunk capacity
    int opgrid = 0;                        // persistent join workgroups
    unsigned long long* part = nullptr;    // [kPartSlots][4] per-workgroup partials
    JoinItem* items = nullptr;             // one record per join work item
    uint64_t max_items = 0;
    unsigned long long* h_result = nullptr; // pinned
    void* gtable = nullptr;
    uint64_t gcap = 0;
    int glog = 0;
    unsigned long long* gspecial = nullptr;
    uint32_t last_bits = 0;
    int last_algo = 0;
    bool pending = false;
    pgpu::KernelTimer* timer = nullptr;
    // The build side is partitioned on a second stream (default priority)
    // beside the probe side: its kernels share CUs with the probe side's
    // histogram, whose LDS counters, not HBM, bound it.
    hipStream_t st2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_build = nullptr;
    pgpu::KernelTimer* timer2 = nullptr;  // kernels on st2
    std::vector<void*> owned;
    int cus = 0;  // compute units of the device (persistent grid sizing)
};

namespace {

using pgpu::ceil_div;

int plan_alloc(pgpu_join_plan* p, uint64_t bytes, void** out) {
    PGPU_TRY(pgpu::ctx_alloc(p->ctx, bytes, out));
    p->owned.push_back(*out);
    return PGPU_OK;
}

// Elements per LDS staging tile for an element of `esz` bytes.
inline uint64_t tile_elems(uint64_t esz) { return kTileBytes / esz; }
// pass-1 rows per thread: the 64 KB tile, at most 16 rows (two workgroups
// per CU hold at most 128 registers per lane)
constexpr int pass1_ipt(int esz) { return kTileBytes1 / esz / kScatterThreads < 16 ? kTileBytes1 / esz / kScatterThreads : 16; }
inline uint64_t tile1_elems(uint64_t esz) { return (uint64_t)pass1_ipt((int)esz) * kScatterThreads; }

// Rows per histogram/scatter block: at most 256 blocks (one per CU), tile-aligned.
inline uint64_t block_rows(uint64_t n, uint64_t esz) {
    const uint64_t T = tile1_elems(esz);
    uint64_t blk = ceil_div(std::max<uint64_t>(n, 1), 256);
    blk = ceil_div(blk, T) * T;
    return std::max<uint64_t>(blk, T);
}

inline int choose_bits(uint64_t nR) {
    // ~1024 build rows per final partition fits the 2048-slot LDS table.
    int b = (int)pgpu::ilog2_ceil(ceil_div(std::max<uint64_t>(nR, 1), 1024));
    return std::max(2, std::min(kMaxBits, b));
}

struct Timed {  // t may be null (untimed launches)
    pgpu::KernelTimer* t;
    int idx;
    Timed(pgpu::KernelTimer* t_, const char* name) : t(t_), idx(t_ ? t_->start(name) : -1) {}
    ~Timed() {
        if (t) t->stop(idx);
    }
};

template <typename KT, bool HV>
int run_partition(pgpu_join_plan* p, hipStream_t st, const KT* keys, const KT* vals, uint64_t n, int B, uint32_t* H,
                  uint32_t* tot, uint64_t* base, uint32_t* px, uint32_t* toff, Elem<KT, HV>* P1, Elem<KT, HV>* P2,
                  const char* tag, pgpu::KernelTimer* timer = nullptr) {
    if (!timer) timer = p->timer;
    using E = Elem<KT, HV>;
    const int B2 = B - B / 2;  // generate(2): pass 1 takes floor(B/2) high bits
    const uint32_t nb = 1u << B, nb1 = 1u << (B - B2);
    const uint64_t blk = block_rows(n, sizeof(E));
    const uint32_t nblk = (uint32_t)ceil_div(std::max<uint64_t>(n, 1), blk);
    if (nblk > kMaxBlocks) return pgpu::fail(PGPU_EINVAL, "radix partition: %u blocks (at most %u)", nblk, kMaxBlocks);
    constexpr int IPT = kTileBytes / (int)sizeof(E) / kScatterThreads;    // pass 2
    constexpr int IPT1 = pass1_ipt((int)sizeof(E));                       // pass 1
    const uint64_t T1 = tile1_elems(sizeof(E));
    const uint64_t ntiles = ceil_div(n, T1);  // blocks are whole tiles
    char name[48];
    {
        snprintf(name, sizeof name, "hist_%s", tag);
        Timed tm(timer, name);
        hipLaunchKernelGGL(k_hist<KT>, dim3(nblk), dim3(512), nb * sizeof(uint32_t), st, keys, n, blk, (uint32_t)T1, B,
                           B2, H, toff);
        PGPU_LAUNCH_CHECK();
    }
    {
        snprintf(name, sizeof name, "scan_%s", tag);
        Timed tm(timer, name);
        hipLaunchKernelGGL(k_colscan, dim3((unsigned)ceil_div(nb, 64)), dim3(512), 0, st, H, nblk, nb, B2, tot, px);
        PGPU_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_binscan, dim3(1), dim3(1024), 0, st, tot, nb, base);
        PGPU_LAUNCH_CHECK();
    }
    {
        snprintf(name, sizeof name, "scatter1_%s", tag);
        Timed tm(timer, name);
        if (ntiles) {
            hipLaunchKernelGGL((k_scatter1<KT, HV, IPT1>), dim3((unsigned)ntiles), dim3(kScatterThreads), 0, st, keys,
                               vals, n, (uint32_t)(blk / T1), nblk, B, B2, base, px, toff, P1);
            PGPU_LAUNCH_CHECK();
        }
    }
    {
        snprintf(name, sizeof name, "scatter2_%s", tag);
        Timed tm(timer, name);
        // group pass-1 blocks so a pass-2 unit averages about one tile
        const uint64_t T = tile_elems(sizeof(E));
        const uint32_t G = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(nblk, T * nb1 / blk));
        const uint32_t ngrp = (nblk + G - 1) / G;
        hipLaunchKernelGGL((k_scatter2<KT, HV, IPT>), dim3(ngrp * nb1), dim3(kScatterThreads), 0, st, P1, nblk, G, B,
                           B2, px, H, base, P2);
        PGPU_LAUNCH_CHECK();
    }
    return PGPU_OK;
}

// Items and the LDS join over partitions already in place: build partition
// d at R2[baseR[d], + totR[d]), probe partitions from p's probe side.
template <typename KT, bool HVS>
int run_lds_join(pgpu_join_plan* p, hipStream_t st, const Elem<KT, true>* R2, const uint64_t* baseR,
                 const uint32_t* totR, uint64_t nS, int B, uint32_t variant) {
    const uint32_t nb = 1u << B;
    {
        Timed tm(p->timer, "join_plan");
        hipLaunchKernelGGL(k_icount, dim3((unsigned)ceil_div(nb, 256)), dim3(256), 0, st, p->totS, totR, nb,
                           kItemRows, p->icount);
        PGPU_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_binscan, dim3(1), dim3(1024), 0, st, p->icount, nb, p->ioff);
        PGPU_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_items, dim3((unsigned)ceil_div(nb, 256)), dim3(256), 0, st, p->icount, p->ioff,
                           p->baseS, p->totS, baseR, totR, nb, kItemRows, p->items);
        PGPU_LAUNCH_CHECK();
    }
    const uint64_t max_items = nb + ceil_div(nS, kItemRows) + 1;
    if (max_items > p->max_items) return pgpu::fail(PGPU_EINVAL, "join items exceed the plan");
    {
        Timed tm(p->timer, "join_lds");
#define PGPU_JOIN(CAP, T, PIECE, UB, WPE)                                                                     \
    hipLaunchKernelGGL((k_join_cs<KT, HVS, CAP, T, PIECE, UB, WPE>), dim3((unsigned)max_items), dim3(T), 0, st, \
                       R2, baseR, totR, (const Elem<KT, HVS>*)p->S2, p->baseS, p->totS, p->ioff, p->items, nb, B, \
                       p->part)
        switch (variant) {
            case 1: PGPU_JOIN(4096, 512, 2048, 8, 4); break;
            case 2: PGPU_JOIN(4096, 256, 2048, 32, 2); break;
            case 3: PGPU_JOIN(2048, 256, 1280, 16, 4); break;
            case 4: PGPU_JOIN(4096, 512, 2048, 16, 4); break;
            case 5: PGPU_JOIN(4096, 512, 2048, 8, 2); break;
            case 7: PGPU_JOIN(4096, 512, 2560, 6, 4); break;   // ~2048 build rows per partition (13 bits at 2^24)
            case 8: PGPU_JOIN(8192, 1024, 5120, 6, 4); break;  // ~4096 build rows per partition (12 bits at 2^24)
            default: PGPU_JOIN(4096, 512, 2048, 6, 4); break;  // no VGPR spills at 4 waves per SIMD
        }
#undef PGPU_JOIN
        PGPU_LAUNCH_CHECK();
    }
    return PGPU_OK;
}

template <typename KT, bool HVS>
int run_radix(pgpu_join_plan* p, const KT* rk, const KT* rv, uint64_t nR, const KT* sk, const KT* sv, uint64_t nS,
              int B) {
    hipStream_t st = p->ctx->stream;
    // Large probe sides: the build side on st2 beside the probe side on st,
    // the join waiting for both.  Below 2^26 probe rows the fork / join
    // between the streams costs a query more than the overlap saves (a C1
    // query through the operators: 0.38 -> 0.49 ms).
    const bool fork = nS >= kForkMinProbe;
    hipStream_t sb = fork ? p->st2 : st;
    if (fork) {
        PGPU_HIP(hipEventRecord(p->ev_fork, st));
        PGPU_HIP(hipStreamWaitEvent(p->st2, p->ev_fork, 0));
    }
    PGPU_TRY((run_partition<KT, true>(p, sb, rk, rv, nR, B, p->HR, p->totR, p->baseR, p->pxR, p->toffR,
                                      (Elem<KT, true>*)p->R1, (Elem<KT, true>*)p->R2, "build",
                                      fork ? p->timer2 : p->timer)));
    if (fork) PGPU_HIP(hipEventRecord(p->ev_build, p->st2));
    PGPU_TRY((run_partition<KT, HVS>(p, st, sk, sv, nS, B, p->HS, p->totS, p->baseS, p->pxS, p->toffS,
                                     (Elem<KT, HVS>*)p->S1, (Elem<KT, HVS>*)p->S2, "probe")));
    if (fork) PGPU_HIP(hipStreamWaitEvent(st, p->ev_build, 0));
    return run_lds_join<KT, HVS>(p, st, (const Elem<KT, true>*)p->R2, p->baseR, p->totR, nS, B, p->cfg.join_variant);
}

// The global-table join over explicit buffers: `table` of at least
// 2^ceil(log2(2 nR)) slots, `special` (2 words), `part` (kPartSlots x 4
// words, zeroed by the caller) receiving the per-workgroup partials.
template <typename KT, bool HVS>
int launch_global(hipStream_t st, pgpu::KernelTimer* timer, void* table, unsigned long long* special,
                  unsigned long long* part, const KT* rk, const KT* rv, uint64_t nR, const KT* sk, const KT* sv,
                  uint64_t nS) {
    const int logcap = (int)pgpu::ilog2_ceil(std::max<uint64_t>(2 * nR, 2));
    const uint64_t cap = 1ull << logcap;
    auto* t = (GSlot<KT>*)table;
    {
        Timed tm(timer, "gclear");
        hipLaunchKernelGGL(k_gclear<KT>, dim3((unsigned)std::min<uint64_t>(ceil_div(cap, 256), 4096)), dim3(256), 0,
                           st, t, cap, special);
        PGPU_LAUNCH_CHECK();
    }
    if (nR) {
        Timed tm(timer, "gbuild");
        hipLaunchKernelGGL(k_gbuild<KT>, dim3((unsigned)std::min<uint64_t>(ceil_div(nR, 256), 8192)), dim3(256), 0,
                           st, rk, rv, nR, t, logcap, special);
        PGPU_LAUNCH_CHECK();
    }
    if (nS) {
        Timed tm(timer, "gprobe");
        hipLaunchKernelGGL((k_gprobe<KT, HVS>), dim3((unsigned)std::min<uint64_t>(ceil_div(nS, 256), 8192)),
                           dim3(256), 0, st, sk, sv, nS, t, logcap, special, part);
        PGPU_LAUNCH_CHECK();
    }
    return PGPU_OK;
}

template <typename KT, bool HVS>
int run_global(pgpu_join_plan* p, const KT* rk, const KT* rv, uint64_t nR, const KT* sk, const KT* sv, uint64_t nS) {
    const uint64_t cap = 1ull << pgpu::ilog2_ceil(std::max<uint64_t>(2 * nR, 2));
    if (cap > p->gcap) return pgpu::fail(PGPU_EINVAL, "global table: %llu slots needed, plan holds %llu",
                                         (unsigned long long)cap, (unsigned long long)p->gcap);
    return launch_global<KT, HVS>(p->ctx->stream, p->timer, p->gtable, p->gspecial, p->part, rk, rv, nR, sk, sv, nS);
}

uint64_t onepass_job_off(uint64_t chunks) { return kOpChunk + ((chunks * sizeof(pgpu::batch::Chunk) + 255) & ~255ull); }

// PGPU_ALGO_RADIX1: both sides as one segment each of the batched plans'
// one-pass partitioning (chunks of one scatter tile, exact per-tile
// offsets), then the persistent LDS join (variant 3: unique-key 8-byte slots
// with the general kernel for the partitions it hands back; variant 2 for
// 8-byte keys).  Results land in opres[0..3].
template <typename KT, bool HVS>
int run_onepass(pgpu_join_plan* p, const KT* rk, const KT* rv, uint64_t nR, const KT* sk, const KT* sv, uint64_t nS) {
    using namespace pgpu::batch;
    hipStream_t st = p->ctx->stream;
    constexpr int kb = (int)sizeof(KT);
    const uint32_t bits = bits_for(nR, kOnePassPartRows);
    if (nR > kOnePassMaxBuild || bits > 10) return pgpu::fail(PGPU_EINVAL, "one-pass join: build side too large");
    Layout lr, ls;
    lr.add(rk, rv, nR, bits, 0, scatter_tile_rows(kb, true, bits));
    const uint32_t nbc = (uint32_t)lr.chunks.size();
    ls.add(sk, HVS ? sv : nullptr, nS, bits, nbc, scatter_tile_rows(kb, HVS, bits));
    const uint32_t nsc = (uint32_t)ls.chunks.size();
    for (auto& ch : ls.chunks) ch.seg = 1;
    ls.segs[0].hist += lr.hist;
    ls.segs[0].bnd += lr.bnd;
    if ((uint64_t)nbc + nsc > p->opchunks) return pgpu::fail(PGPU_EINVAL, "one-pass join: chunks exceed the plan");
    Job jb{};
    jb.bits = bits;  // both sides' rows and bounds start at 0; one item per partition
    const uint32_t nitems = 1u << bits;
    // the staging block may still feed the previous call's copy
    if (p->pending) PGPU_HIP(hipStreamSynchronize(st));
    char* h = p->ophost;
    const uint64_t job_off = onepass_job_off(p->opchunks), item_off = job_off + 256;
    memcpy(h + kOpSeg, &lr.segs[0], sizeof(Seg));
    memcpy(h + kOpSeg + sizeof(Seg), &ls.segs[0], sizeof(Seg));
    memcpy(h + kOpChunk, lr.chunks.data(), nbc * sizeof(Chunk));
    memcpy(h + kOpChunk + nbc * sizeof(Chunk), ls.chunks.data(), nsc * sizeof(Chunk));
    memcpy(h + job_off, &jb, sizeof(Job));
    memset(h + item_off, 0, nitems * sizeof(uint32_t));  // every item belongs to job 0
    char* d = (char*)p->opdev;
    PGPU_HIP(hipMemcpyAsync(d, h, item_off + nitems * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    const Seg* segs = (const Seg*)(d + kOpSeg);
    const Chunk* chunks = (const Chunk*)(d + kOpChunk);
    PGPU_TRY(launch_zero(st, p->opres, 6));
    {
        Timed tm(p->timer, "hist");
        PGPU_TRY(launch_hist(st, kb, segs, chunks, 0, nbc + nsc, p->opH));
    }
    {
        Timed tm(p->timer, "scan");
        PGPU_TRY(launch_scan(st, bits, segs, 0, 2, p->opH, p->opbnd));
    }
    {
        Timed tm(p->timer, "scatter_build");
        PGPU_TRY(launch_scatter(st, kb, true, bits, segs, chunks, 0, nbc, p->opH, p->R2));
    }
    {
        Timed tm(p->timer, "scatter_probe");
        PGPU_TRY(launch_scatter(st, kb, HVS, bits, segs, chunks, nbc, nsc, p->opH, p->S2));
    }
    {
        Timed tm(p->timer, "join_lds");
        PGPU_TRY(launch_join(st, kb, HVS, kb == 4 ? 3 : 2, std::min<int>(p->opgrid, (int)nitems),
                             (const Job*)(d + job_off), (const uint32_t*)(d + item_off), nitems, p->R2, p->S2,
                             p->opbnd, p->opbnd + lr.bnd, (uint32_t*)(p->opres + 4), p->opres, p->opredo));
    }
    return PGPU_OK;
}

}  // namespace

using namespace pgpu;

// A build side partitioned once and probed by many queries (the serving
// engine's large queries): rows / base / tot in the layout of a plan's
// partitioned build side.
int pgpu::radix_build_resident(pgpu_join_plan* p, hipStream_t st, const int64_t* key, const int64_t* pay,
                               uint64_t n, int bits, void* rows, uint64_t* base, uint32_t* tot) {
    if (!p || !p->R1 || p->cfg.key_bytes != 8)
        return fail(PGPU_EINVAL, "radix_build_resident: a plan with 8-byte radix buffers is required");
    if (!n || n > p->maxR)
        return fail(PGPU_EINVAL, "radix_build_resident: %llu build rows (the plan holds %llu)", (unsigned long long)n,
                    (unsigned long long)p->maxR);
    if (bits < 2 || bits > kMaxBits || (1u << bits) > p->nbmax)
        return fail(PGPU_EINVAL, "radix_build_resident: %d radix bits", bits);
    PGPU_TRY(ctx_activate(p->ctx));
    return run_partition<int64_t, true>(p, st, key, pay, n, bits, p->HR, tot, base, p->pxR, p->toffR,
                                        (Elem<int64_t, true>*)p->R1, (Elem<int64_t, true>*)rows, "build");
}

int pgpu::radix_probe_resident(pgpu_join_plan* p, hipStream_t st, const int64_t* probe, uint64_t m, int bits,
                               const void* rows, const uint64_t* base, const uint32_t* tot,
                               unsigned long long* res) {
    if (!p || !p->S1 || p->cfg.key_bytes != 8 || p->cfg.probe_payload)
        return fail(PGPU_EINVAL, "radix_probe_resident: a plan with 8-byte keys and no probe payload is required");
    if (!m || m > p->maxS)
        return fail(PGPU_EINVAL, "radix_probe_resident: %llu probe rows (the plan holds %llu)", (unsigned long long)m,
                    (unsigned long long)p->maxS);
    if (bits < 2 || bits > kMaxBits || (1u << bits) > p->nbmax)
        return fail(PGPU_EINVAL, "radix_probe_resident: %d radix bits", bits);
    PGPU_TRY(ctx_activate(p->ctx));
    hipLaunchKernelGGL(k_fill64, dim3(8), dim3(256), 0, st, p->part, (uint64_t)4 * kPartSlots);
    PGPU_LAUNCH_CHECK();
    PGPU_TRY((run_partition<int64_t, false>(p, st, probe, nullptr, m, bits, p->HS, p->totS, p->baseS, p->pxS, p->toffS,
                                            (Elem<int64_t, false>*)p->S1, (Elem<int64_t, false>*)p->S2, "probe")));
    PGPU_TRY((run_lds_join<int64_t, false>(p, st, (const Elem<int64_t, true>*)rows, base, tot, m, bits, 0)));
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, st, p->part, kPartSlots, res);
    PGPU_LAUNCH_CHECK();
    return PGPU_OK;
}

extern "C" {

int pgpu_join_plan_create(pgpu_ctx* ctx, const pgpu_join_cfg* cfg, uint64_t max_build, uint64_t max_probe,
                          pgpu_join_plan** out) {
    if (!ctx || !cfg || !out) return fail(PGPU_EINVAL, "pgpu_join_plan_create: null argument");
    *out = nullptr;
    if (cfg->key_bytes != 4 && cfg->key_bytes != 8)
        return fail(PGPU_EUNSUPPORTED, "join keys must be 4 or 8 bytes (got %u)", cfg->key_bytes);
    if (cfg->radix_bits != 0 && (cfg->radix_bits < 2 || cfg->radix_bits > kMaxBits))
        return fail(PGPU_EINVAL, "radix_bits must be 0 or in [2, %d] (got %u)", kMaxBits, cfg->radix_bits);
    if (cfg->algo > PGPU_ALGO_RADIX1) return fail(PGPU_EINVAL, "unknown join algorithm %u", cfg->algo);
    if (cfg->algo == PGPU_ALGO_RADIX1 && max_build > kOnePassMaxBuild)
        return fail(PGPU_EINVAL, "the one-pass join takes at most %llu build rows",
                    (unsigned long long)kOnePassMaxBuild);
    if (cfg->join_variant > 8) return fail(PGPU_EINVAL, "unknown LDS join variant %u", cfg->join_variant);
    // Histogram columns, per-bin totals and the per-(block, bin) prefixes of
    // k_colscan are 32-bit: one final partition may hold every row of a side,
    // so each side must stay under 2^32 rows.
    if (max_build >= (1ull << 32) || max_probe >= (1ull << 32))
        return fail(PGPU_EINVAL, "relation too large (build and probe must each hold fewer than 2^32 rows)");
    PGPU_TRY(ctx_activate(ctx));
    auto* p = new pgpu_join_plan();
    p->ctx = ctx;
    p->cfg = *cfg;
    p->maxR = max_build;
    p->maxS = max_probe;
    p->bits = (int)cfg->radix_bits;
    p->timer = new KernelTimer(ctx->stream);
    auto cleanup = [&](int rc) {
        pgpu_join_plan_destroy(p);
        return rc;
    };
    if (hipStreamCreateWithFlags(&p->st2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&p->ev_build, hipEventDisableTiming) != hipSuccess)
        return cleanup(fail(PGPU_EDEVICE, "cannot create the plan's build-side stream"));
    p->timer2 = new KernelTimer(p->st2);
    if (hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess ||
        p->cus < 1)
        return cleanup(fail(PGPU_EDEVICE, "cannot read the device's compute unit count"));
    const uint64_t kb = cfg->key_bytes;
    const uint64_t esR = 2 * kb, esS = cfg->probe_payload ? 2 * kb : kb;
    const bool want_radix = cfg->algo != PGPU_ALGO_GLOBAL;
    const bool want_global = cfg->algo != PGPU_ALGO_RADIX;
    int rc;
    if ((rc = plan_alloc(p, 4 * sizeof(unsigned long long), (void**)&p->result)) != PGPU_OK) return cleanup(rc);
    if ((rc = plan_alloc(p, 4 * kPartSlots * sizeof(unsigned long long), (void**)&p->part)) != PGPU_OK)
        return cleanup(rc);
    if (hipHostMalloc((void**)&p->h_result, 4 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess)
        return cleanup(fail(PGPU_ENOMEM, "hipHostMalloc for the result failed"));
    if (want_radix) {
        const int maxbits = p->bits ? p->bits : kMaxBits;
        p->nbmax = 1u << maxbits;
        p->nblkR = (uint32_t)ceil_div(std::max<uint64_t>(max_build, 1), block_rows(max_build, esR));
        p->nblkS = (uint32_t)ceil_div(std::max<uint64_t>(max_probe, 1), block_rows(max_probe, esS));
        // block_rows() never yields more than 512 blocks; size for that bound
        // so smaller calls (which may pick a different block size) fit too.
        p->nblkR = std::max<uint32_t>(p->nblkR, 512);
        p->nblkS = std::max<uint32_t>(p->nblkS, 512);
        const uint64_t nb = p->nbmax;
        p->max_items = nb + ceil_div(std::max<uint64_t>(max_probe, 1), kItemRows) + 1;
        struct {
            void** ptr;
            uint64_t bytes;
        } a[] = {
            {&p->R1, std::max<uint64_t>(max_build, 1) * esR},
            {&p->R2, std::max<uint64_t>(max_build, 1) * esR},
            {&p->S1, std::max<uint64_t>(max_probe, 1) * esS},
            {&p->S2, std::max<uint64_t>(max_probe, 1) * esS},
            {(void**)&p->HR, (uint64_t)p->nblkR * nb * 4},
            {(void**)&p->HS, (uint64_t)p->nblkS * nb * 4},
            {(void**)&p->totR, nb * 4},
            {(void**)&p->totS, nb * 4},
            {(void**)&p->icount, nb * 4},
            {(void**)&p->baseR, (nb + 1) * 8},
            {(void**)&p->baseS, (nb + 1) * 8},
            {(void**)&p->ioff, (nb + 1) * 8},
            {(void**)&p->items, (nb + ceil_div(std::max<uint64_t>(max_probe, 1), kItemRows) + 1) * sizeof(JoinItem)},
            {(void**)&p->pxR, 2ull * kMaxBlocks * kDigits * 4},
            {(void**)&p->pxS, 2ull * kMaxBlocks * kDigits * 4},
            {(void**)&p->toffR, (ceil_div(std::max<uint64_t>(max_build, 1), tile1_elems(esR)) + 1) * kDigits * 4},
            {(void**)&p->toffS, (ceil_div(std::max<uint64_t>(max_probe, 1), tile1_elems(esS)) + 1) * kDigits * 4},
        };
        for (auto& x : a)
            if ((rc = plan_alloc(p, x.bytes, x.ptr)) != PGPU_OK) return cleanup(rc);
    }
    if (want_radix && max_build <= kOnePassMaxBuild) {
        const uint32_t crR = batch::scatter_tile_rows((int)kb, true, 10);
        const uint32_t crS = batch::scatter_tile_rows((int)kb, cfg->probe_payload != 0, 10);
        p->opchunks = ceil_div(std::max<uint64_t>(max_build, 1), crR) + ceil_div(std::max<uint64_t>(max_probe, 1), crS);
        const uint64_t block = onepass_job_off(p->opchunks) + 256 + kOnePassMaxItems * 4;
        struct {
            void** ptr;
            uint64_t bytes;
        } a[] = {
            {&p->opdev, block},
            {(void**)&p->opH, p->opchunks * kOnePassMaxItems * 4},
            {(void**)&p->opbnd, 2 * (kOnePassMaxItems + 1) * 4},
            {(void**)&p->opres, 6 * 8},
            {(void**)&p->opredo, kOnePassMaxItems * 4},
        };
        for (auto& x : a)
            if ((rc = plan_alloc(p, x.bytes, x.ptr)) != PGPU_OK) return cleanup(rc);
        if (hipHostMalloc((void**)&p->ophost, block, hipHostMallocDefault) != hipSuccess)
            return cleanup(fail(PGPU_ENOMEM, "hipHostMalloc for the one-pass tables failed"));
        if ((rc = batch::join_grid((int)kb, cfg->probe_payload != 0, kb == 4 ? 3 : 2, &p->opgrid)) != PGPU_OK)
            return cleanup(rc);
    }
    if (want_global) {
        p->glog = (int)ilog2_ceil(std::max<uint64_t>(2 * max_build, 2));
        p->gcap = 1ull << p->glog;
        const uint64_t slot = kb == 8 ? sizeof(GSlot<int64_t>) : sizeof(GSlot<int32_t>);
        if ((rc = plan_alloc(p, p->gcap * slot, &p->gtable)) != PGPU_OK) return cleanup(rc);
        if ((rc = plan_alloc(p, 2 * sizeof(unsigned long long), (void**)&p->gspecial)) != PGPU_OK)
            return cleanup(rc);
    }
    *out = p;
    return PGPU_OK;
}

int pgpu_join_plan_destroy(pgpu_join_plan* p) {
    if (!p) return PGPU_OK;
    if (p->ctx) {
        (void)hipSetDevice(p->ctx->device);
        (void)hipStreamSynchronize(p->ctx->stream);
        if (p->st2) (void)hipStreamSynchronize(p->st2);
    }
    delete p->timer;
    delete p->timer2;
    if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
    if (p->ev_build) (void)hipEventDestroy(p->ev_build);
    if (p->st2) (void)hipStreamDestroy(p->st2);
    for (void* x : p->owned) (void)ctx_free(p->ctx, x);
    if (p->h_result) (void)hipHostFree(p->h_result);
    if (p->ophost) (void)hipHostFree(p->ophost);
    delete p;
    return PGPU_OK;
}

int pgpu_join_count_sum_async(pgpu_join_plan* p, const void* d_rkey, const void* d_rpay, uint64_t nR,
                              const void* d_skey, const void* d_spay, uint64_t nS) {
    if (!p) return fail(PGPU_EINVAL, "pgpu_join_count_sum_async: null plan");
    if (nR > p->maxR || nS > p->maxS)
        return fail(PGPU_EINVAL, "join of %llu x %llu rows exceeds the plan's %llu x %llu",
                    (unsigned long long)nR, (unsigned long long)nS, (unsigned long long)p->maxR,
                    (unsigned long long)p->maxS);
    if ((nR && (!d_rkey || !d_rpay)) || (nS && !d_skey) || (nS && p->cfg.probe_payload && !d_spay))
        return fail(PGPU_EINVAL, "pgpu_join_count_sum_async: null column");
    PGPU_TRY(ctx_activate(p->ctx));
    hipStream_t st = p->ctx->stream;
    hipLaunchKernelGGL(k_fill64, dim3(8), dim3(256), 0, st, p->part, (uint64_t)4 * kPartSlots);
    PGPU_LAUNCH_CHECK();
    int algo = (int)p->cfg.algo;
    if (algo == PGPU_ALGO_AUTO)
        algo = (nR <= (1u << 16) || !p->R1)            ? PGPU_ALGO_GLOBAL
               : (p->opdev && !p->bits && nR <= kOnePassMaxBuild) ? PGPU_ALGO_RADIX1
                                                                 : PGPU_ALGO_RADIX;
    if (algo == PGPU_ALGO_GLOBAL && !p->gtable) algo = PGPU_ALGO_RADIX;
    if (algo == PGPU_ALGO_RADIX && !p->R1) algo = PGPU_ALGO_GLOBAL;
    p->last_algo = algo;
    int rc = PGPU_OK;
    const bool hv = p->cfg.probe_payload != 0;
    const bool onepass = algo == PGPU_ALGO_RADIX1 && nR && nS;
    if (onepass) {
        p->last_bits = batch::bits_for(nR, kOnePassPartRows);
        if (p->cfg.key_bytes == 8)
            rc = hv ? run_onepass<int64_t, true>(p, (const int64_t*)d_rkey, (const int64_t*)d_rpay, nR,
                                                 (const int64_t*)d_skey, (const int64_t*)d_spay, nS)
                    : run_onepass<int64_t, false>(p, (const int64_t*)d_rkey, (const int64_t*)d_rpay, nR,
                                                  (const int64_t*)d_skey, nullptr, nS);
        else
            rc = hv ? run_onepass<int32_t, true>(p, (const int32_t*)d_rkey, (const int32_t*)d_rpay, nR,
                                                 (const int32_t*)d_skey, (const int32_t*)d_spay, nS)
                    : run_onepass<int32_t, false>(p, (const int32_t*)d_rkey, (const int32_t*)d_rpay, nR,
                                                  (const int32_t*)d_skey, nullptr, nS);
        if (rc != PGPU_OK) return rc;
        PGPU_HIP(hipMemcpyAsync(p->h_result, p->opres, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        p->pending = true;
        return PGPU_OK;
    }
    if (algo == PGPU_ALGO_RADIX1) algo = PGPU_ALGO_RADIX;  // an empty side: every count is 0
    if (nR && nS) {
        if (algo == PGPU_ALGO_RADIX) {
            const int B = p->bits ? p->bits : std::min(choose_bits(nR), (int)ilog2_ceil(p->nbmax));
            p->last_bits = (uint32_t)B;
            if (p->cfg.key_bytes == 8)
                rc = hv ? run_radix<int64_t, true>(p, (const int64_t*)d_rkey, (const int64_t*)d_rpay, nR,
                                                   (const int64_t*)d_skey, (const int64_t*)d_spay, nS, B)
                        : run_radix<int64_t, false>(p, (const int64_t*)d_rkey, (const int64_t*)d_rpay, nR,
                                                    (const int64_t*)d_skey, nullptr, nS, B);
            else
                rc = hv ? run_radix<int32_t, true>(p, (const int32_t*)d_rkey, (const int32_t*)d_rpay, nR,
                                                   (const int32_t*)d_skey, (const int32_t*)d_spay, nS, B)
                        : run_radix<int32_t, false>(p, (const int32_t*)d_rkey, (const int32_t*)d_rpay, nR,
                                                    (const int32_t*)d_skey, nullptr, nS, B);
        } else {
            p->last_bits = 0;
            if (p->cfg.key_bytes == 8)
                rc = hv ? run_global<int64_t, true>(p, (const int64_t*)d_rkey, (const int64_t*)d_rpay, nR,
                                                    (const int64_t*)d_skey, (const int64_t*)d_spay, nS)
                        : run_global<int64_t, false>(p, (const int64_t*)d_rkey, (const int64_t*)d_rpay, nR,
                                                     (const int64_t*)d_skey, nullptr, nS);
            else
                rc = hv ? run_global<int32_t, true>(p, (const int32_t*)d_rkey, (const int32_t*)d_rpay, nR,
                                                    (const int32_t*)d_skey, (const int32_t*)d_spay, nS)
                        : run_global<int32_t, false>(p, (const int32_t*)d_rkey, (const int32_t*)d_rpay, nR,
                                                     (const int32_t*)d_skey, nullptr, nS);
        }
    }
    if (rc != PGPU_OK) return rc;
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, st, p->part, kPartSlots, p->result);
    PGPU_LAUNCH_CHECK();
    PGPU_HIP(hipMemcpyAsync(p->h_result, p->result, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    p->pending = true;
    return PGPU_OK;
}

int pgpu_join_plan_result(pgpu_join_plan* p, pgpu_join_result* out) {
    if (!p || !out) return fail(PGPU_EINVAL, "pgpu_join_plan_result: null argument");
    if (!p->pending) return fail(PGPU_ESTATE, "pgpu_join_plan_result: no join enqueued");
    PGPU_TRY(ctx_activate(p->ctx));
    PGPU_HIP(hipStreamSynchronize(p->ctx->stream));
    p->pending = false;
    memset(out, 0, sizeof *out);
    out->matches = p->h_result[0];
    out->sum_build = p->h_result[1];
    out->sum_probe = p->h_result[2];
    out->radix_bits = p->last_bits;
    out->lds_passes = std::max<uint32_t>(1, (uint32_t)p->h_result[3]);
    return PGPU_OK;
}

int pgpu_join_count_sum(pgpu_join_plan* p, const void* d_rkey, const void* d_rpay, uint64_t nR, const void* d_skey,
                        const void* d_spay, uint64_t nS, pgpu_join_result* out) {
    PGPU_TRY(pgpu_join_count_sum_async(p, d_rkey, d_rpay, nR, d_skey, d_spay, nS));
    return pgpu_join_plan_result(p, out);
}

int pgpu_join_plan_timing(pgpu_join_plan* p, int enable) {
    if (!p) return fail(PGPU_EINVAL, "pgpu_join_plan_timing: null plan");
    PGPU_TRY(ctx_activate(p->ctx));
    p->timer->enable(enable != 0);
    if (p->timer2) p->timer2->enable(enable != 0);
    return PGPU_OK;
}

int pgpu_join_plan_times(pgpu_join_plan* p, pgpu_kernel_time* out, int cap, int* n) {
    if (!p || !n || (cap > 0 && !out)) return fail(PGPU_EINVAL, "pgpu_join_plan_times: null argument");
    PGPU_TRY(ctx_activate(p->ctx));
    p->timer->harvest();
    std::vector<KernelTimer::Total> t = p->timer->totals();  // the build side's kernels (st2) first
    if (p->timer2) {
        p->timer2->harvest();
        t.insert(t.begin(), p->timer2->totals().begin(), p->timer2->totals().end());
    }
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

// ====================================================================== no-allocation join
// pgpu_join_count_sum_noalloc: the global-table join whose table, partials
// and result words live in `scratch`, a file-scope device array (each device
// holds its own copy until the process ends).  Host code reaches a device's
// copy only through the address hipGetSymbolAddress returns with that device
// current; the array's name is not a device address on the host.
namespace {
constexpr uint64_t kScratchTableBytes = 32ull << 20;                 // global hash table
constexpr uint64_t kScratchPartOffset = kScratchTableBytes;          // per-workgroup partials
constexpr uint64_t kScratchPartBytes = 4ull * kPartSlots * sizeof(unsigned long long);
constexpr uint64_t kScratchSpecialOffset = kScratchPartOffset + kScratchPartBytes;  // 2 words
constexpr uint64_t kScratchResultOffset = kScratchSpecialOffset + 256;              // 4 words
constexpr uint64_t kScratchBytes = kScratchResultOffset + 256;
constexpr int kScratchDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
char* g_scratch_addr[kScratchDevices] = {};  // per device, from hipGetSymbolAddress

// The calling thread's current device must be `device`.
int scratch_address(int device, char** out) {
    if (device < 0 || device >= kScratchDevices) return pgpu::fail(PGPU_EINVAL, "device %d out of range", device);
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_addr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return pgpu::fail(PGPU_EDEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                              hipGetErrorString(e));
        g_scratch_addr[device] = (char*)p;
    }
    *out = g_scratch_addr[device];
    return PGPU_OK;
}

// Build rows the scratch table holds: its 2^ceil(log2(2 n)) slots must fit.
uint64_t noalloc_max_build(uint32_t key_bytes) {
    const uint64_t slot = key_bytes == 8 ? sizeof(GSlot<int64_t>) : sizeof(GSlot<int32_t>);
    uint64_t cap = 1;
    while (cap * 2 * slot <= kScratchTableBytes) cap *= 2;
    return cap / 2;
}
}  // namespace

extern "C" {

int pgpu_join_count_sum_noalloc(pgpu_ctx* ctx, uint32_t key_bytes, uint32_t probe_payload, const void* d_rkey,
                                const void* d_rpay, uint64_t n_build, const void* d_skey, const void* d_spay,
                                uint64_t n_probe, pgpu_join_result* out) {
    if (!ctx || !out) return fail(PGPU_EINVAL, "pgpu_join_count_sum_noalloc: null argument");
    if (key_bytes != 4 && key_bytes != 8)
        return fail(PGPU_EUNSUPPORTED, "join keys must be 4 or 8 bytes (got %u)", key_bytes);
    const uint64_t maxb = noalloc_max_build(key_bytes);
    if (n_build > maxb)
        return fail(PGPU_EINVAL, "pgpu_join_count_sum_noalloc: %llu build rows exceed its %llu-row table (%llu MiB scratch)",
                    (unsigned long long)n_build, (unsigned long long)maxb,
                    (unsigned long long)(kScratchTableBytes >> 20));
    if ((n_build && (!d_rkey || !d_rpay)) || (n_probe && !d_skey) || (n_probe && probe_payload && !d_spay))
        return fail(PGPU_EINVAL, "pgpu_join_count_sum_noalloc: null column");
    PGPU_TRY(ctx_activate(ctx));
    char* base = nullptr;
    PGPU_TRY(scratch_address(ctx->device, &base));
    auto* part = (unsigned long long*)(base + kScratchPartOffset);
    auto* special = (unsigned long long*)(base + kScratchSpecialOffset);
    auto* result = (unsigned long long*)(base + kScratchResultOffset);
    hipStream_t st = ctx->stream;
    // the partials start at zero (fill kernels, not memsets: capturable)
    hipLaunchKernelGGL(k_fill64, dim3(8), dim3(256), 0, st, part, (uint64_t)4 * kPartSlots);
    PGPU_LAUNCH_CHECK();
    int rc = PGPU_OK;
    if (n_build && n_probe) {
        const bool hv = probe_payload != 0;
        if (key_bytes == 8)
            rc = hv ? launch_global<int64_t, true>(st, nullptr, base, special, part, (const int64_t*)d_rkey,
                                                   (const int64_t*)d_rpay, n_build, (const int64_t*)d_skey,
                                                   (const int64_t*)d_spay, n_probe)
                    : launch_global<int64_t, false>(st, nullptr, base, special, part, (const int64_t*)d_rkey,
                                                    (const int64_t*)d_rpay, n_build, (const int64_t*)d_skey, nullptr,
                                                    n_probe);
        else
            rc = hv ? launch_global<int32_t, true>(st, nullptr, base, special, part, (const int32_t*)d_rkey,
                                                   (const int32_t*)d_rpay, n_build, (const int32_t*)d_skey,
                                                   (const int32_t*)d_spay, n_probe)
                    : launch_global<int32_t, false>(st, nullptr, base, special, part, (const int32_t*)d_rkey,
                                                    (const int32_t*)d_rpay, n_build, (const int32_t*)d_skey, nullptr,
                                                    n_probe);
    }
    if (rc != PGPU_OK) return rc;
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, st, part, kPartSlots, result);
    PGPU_LAUNCH_CHECK();
    unsigned long long h[4] = {0, 0, 0, 0};
    PGPU_HIP(hipMemcpyAsync(h, result, sizeof h, hipMemcpyDeviceToHost, st));
    PGPU_HIP(hipStreamSynchronize(st));
    memset(out, 0, sizeof *out);
    out->matches = h[0];
    out->sum_build = h[1];
    out->sum_probe = h[2];
    out->radix_bits = 0;
    out->lds_passes = (uint32_t)h[3];
    return PGPU_OK;
}

}  // extern "C"
