// Row-format join and aggregation behind the gpuhashjoin / gpuaggregate_*
// operators, where tuples arrive and leave as the engine's packed pages.
//
// Join (replaces HashJoinOp build + probe + constructOutputTuple,
// operators/join.cpp:475-667): the build rows (sbuild layout) are uploaded
// once; an open-addressing multimap in HBM (at most a quarter full) holds
// (key, row index) per build row.  A probe batch is joined in two launches
// over per-workgroup chunks: a count pass records each row's match count and
// first match, a write pass places every workgroup's rows after the counts
// of the workgroups before it and writes the projected rows in probe order
// (no global atomics).  When the output buffer is too small the write pass
// runs again into a larger one.  Host rows and pulled output go through
// pooled pinned staging.
//
// Aggregation (replaces GenericAggregate::remember and the AggregateSum /
// AggregateCount folds, operators/genericaggregate.cpp:240-282): an
// open-addressing table in HBM keyed by the packed group-key bytes; a slot is
// claimed with a state CAS (empty -> writing -> ready) and folded with
// hardware atomics in the accumulator's type.  The table grows (rehash) when
// half full.  finish() compacts the ready slots into output rows.  Without a
// GROUP BY the rows are summed per workgroup and then in one workgroup (no
// same-address atomics).  Each aggregate runs on a stream of its own.
#include <atomic>
#include <algorithm>
#include <cstring>

#include "pgpu_internal.h"

namespace {

constexpr uint32_t kEmptyIdx = 0xFFFFFFFFu;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBull;
    z ^= z >> 31;
    return z;
}

// Packed rows (no padding: schema.cpp:82-117) align a column only when every
// width before it is a multiple of 4 (no char(n) column).  Kernels take an
// `a4` flag, true when the layout and the base pointers are 4-byte aligned,
// and then use word accesses; otherwise bytes.
__device__ __forceinline__ uint32_t ld_u32(const char* p, bool a4) {
    if (a4) return *(const uint32_t*)p;
    uint32_t v;
    __builtin_memcpy(&v, p, 4);
    return v;
}

__device__ __forceinline__ uint64_t ld_u64(const char* p, bool a4) {
    if (a4) {
        const uint32_t* q = (const uint32_t*)p;  // 8-byte columns may sit at 4 mod 8
        return (uint64_t)q[0] | ((uint64_t)q[1] << 32);
    }
    uint64_t v;
    __builtin_memcpy(&v, p, 8);
    return v;
}

// bytes % 4 == 0 when a4
__device__ __forceinline__ void copy_cols(char* dst, const char* src, uint32_t bytes, bool a4) {
    if (a4) {
        for (uint32_t w = 0; w < bytes / 4; ++w) ((uint32_t*)dst)[w] = ((const uint32_t*)src)[w];
        return;
    }
    for (uint32_t b = 0; b < bytes; ++b) dst[b] = src[b];
}

__device__ __forceinline__ int64_t load_key(const char* p, uint32_t bytes, bool a4) {
    if (bytes == 4) return (int32_t)ld_u32(p, a4);
    return (int64_t)ld_u64(p, a4);
}

__host__ __device__ inline bool aligned4(uint64_t v) { return (v & 3) == 0; }

struct alignas(16) JSlot {  // one 16-byte load per probe step
    int64_t key;
    uint32_t idx;
    uint32_t pad;
};

// How join keys compare (Comparator::init, comparator.cpp:43-413):
//   kKeyInt  integers of 4 or 8 bytes by value: the key is the sign-extended value
//   kKeyDbl  a dec key on either side: the key is the canonical bit pattern
//            of the value as a double (-0 -> +0); a NaN key matches nothing
//   kKeyChr  char against char: strncmp over `nchar` bytes; the key is a hash
//            of the bytes before the first NUL, equal hashes are confirmed
//            against the build row's bytes
enum : uint32_t { kKeyInt = 0, kKeyDbl = 1, kKeyChr = 2 };

struct JoinDev {
    pgpu_hj_desc d;
    uint32_t out_row_bytes;
    uint32_t dst_off[PGPU_MAX_PROJ];
    uint32_t layout_a4;  // every row width, offset and projected width a multiple of 4
    uint32_t key_mode;   // kKeyInt / kKeyDbl / kKeyChr
    uint32_t nchar;      // kKeyChr: bytes compared (the smaller key width)
};

// The join key of the row field at p (type: enum pgpu_col_type, 0 = signed
// integer); false when the key can match nothing (a NaN).
__device__ __forceinline__ bool join_key(const char* p, uint32_t bytes, uint32_t type, uint32_t mode,
                                         uint32_t nchar, bool a4, int64_t* key) {
    if (mode == kKeyInt) {
        *key = load_key(p, bytes, a4);
        return true;
    }
    if (mode == kKeyDbl) {
        const double v = type == PGPU_COL_F64 ? __longlong_as_double((long long)ld_u64(p, a4))
                                              : (double)load_key(p, bytes, a4);
        if (v != v) return false;
        *key = v == 0.0 ? 0 : (int64_t)__double_as_longlong(v);
        return true;
    }
    uint64_t h = 14695981039346656037ull;  // FNV-1a over the bytes before the first NUL
    for (uint32_t i = 0; i < nchar; ++i) {
        const unsigned char c = (unsigned char)p[i];
        if (!c) break;
        h = (h ^ c) * 1099511628211ull;
    }
    *key = (int64_t)h;
    return true;
}

// strncmp(a, b, n) == 0
__device__ __forceinline__ bool chars_equal(const char* a, const char* b, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) {
        if (a[i] != b[i]) return false;
        if (!a[i]) return true;
    }
    return true;
}

__global__ __launch_bounds__(256) void k_jclear(JSlot* t, uint64_t cap) {
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < cap; i += stride) {
        t[i].key = 0;
        t[i].idx = kEmptyIdx;
    }
}

__global__ __launch_bounds__(256) void k_jbuild(JoinDev jd, const char* __restrict__ rows, uint64_t n,
                                                JSlot* __restrict__ t, int logcap) {
    const pgpu_hj_desc& d = jd.d;
    const uint64_t mask = (1ull << logcap) - 1;
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    const bool a4 = aligned4((uintptr_t)rows | d.build_row_bytes | d.build_key_offset);
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
        int64_t k;
        if (!join_key(rows + i * d.build_row_bytes + d.build_key_offset, d.build_key_bytes, d.build_key_type,
                      jd.key_mode, jd.nchar, a4, &k))
            continue;  // a NaN key joins nothing
        uint64_t s = mix64((uint64_t)k) & mask;
        while (atomicCAS(&t[s].idx, kEmptyIdx, (uint32_t)i) != kEmptyIdx) s = (s + 1) & mask;
        t[s].key = k;
    }
}

// Exclusive scan of one value per thread over a 256-thread block.
__device__ __forceinline__ uint32_t block_scan256(uint32_t v, uint32_t* total) {
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

// Sum of one value per thread over a 256-thread block, returned to every thread.
__device__ __forceinline__ unsigned long long block_sum256(unsigned long long v) {
    __shared__ unsigned long long ws[4];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    const unsigned long long t = ws[0] + ws[1] + ws[2] + ws[3];
    __syncthreads();
    return t;
}

// The probe runs in two launches over the same split of the rows into
// contiguous per-block chunks.  k_jcount probes every row once and records
// (match count << 32 | first match's build row) per row, plus each block's
// total; k_jwrite starts each block at the sum of the totals before it and
// writes its rows in probe order: a row with one match (the primary-key
// case) is written from its record without touching the table again, only
// rows with several matches walk it a second time.  No block takes a global
// atomic (one cursor shared by thousands of blocks serialises them).  Each
// thread holds kProbeU consecutive rows, so their loads are in flight
// together and a wave reads one contiguous span.
constexpr unsigned kProbeBlocks = 2048;
constexpr int kProbeU = 4;
constexpr unsigned kProbeTile = 256 * kProbeU;

__global__ __launch_bounds__(256) void k_jcount(JoinDev jd, const char* __restrict__ build,
                                                const JSlot* __restrict__ t, int logcap,
                                                const char* __restrict__ probe, uint64_t n, uint64_t chunk,
                                                unsigned long long* __restrict__ counts,
                                                unsigned long long* __restrict__ rec) {
    const pgpu_hj_desc& d = jd.d;
    const uint64_t mask = (1ull << logcap) - 1;
    const bool a4 = jd.layout_a4 && aligned4((uintptr_t)probe);
    const uint64_t lo = (uint64_t)blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
    unsigned long long c = 0;
    for (uint64_t b0 = lo; b0 < hi; b0 += kProbeTile) {
        const uint64_t i0 = b0 + (uint64_t)threadIdx.x * kProbeU;
        int64_t k[kProbeU];
        uint64_t sl[kProbeU];
        JSlot e[kProbeU];
        bool valid[kProbeU];
#pragma unroll
        for (int u = 0; u < kProbeU; ++u) {
            k[u] = 0;
            valid[u] = i0 + u < hi &&
                       join_key(probe + (i0 + u) * d.probe_row_bytes + d.probe_key_offset, d.probe_key_bytes,
                                d.probe_key_type, jd.key_mode, jd.nchar, a4, &k[u]);
        }
#pragma unroll
        for (int u = 0; u < kProbeU; ++u) {
            sl[u] = mix64((uint64_t)k[u]) & mask;
            e[u] = t[sl[u]];
        }
        // the kProbeU collision chains are walked in lock step: one round of
        // dependent loads per step instead of one per row and step
        uint32_t cnt[kProbeU], first[kProbeU];
        bool act[kProbeU];
#pragma unroll
        for (int u = 0; u < kProbeU; ++u) {
            cnt[u] = first[u] = 0;
            act[u] = valid[u];
        }
        while (true) {
            bool any = false;
#pragma unroll
            for (int u = 0; u < kProbeU; ++u) {
                if (!act[u]) continue;
                if (e[u].idx == kEmptyIdx) {
                    act[u] = false;
                    continue;
                }
                if (e[u].key == k[u] &&
                    (jd.key_mode != kKeyChr ||
                     chars_equal(build + (uint64_t)e[u].idx * d.build_row_bytes + d.build_key_offset,
                                 probe + (i0 + u) * d.probe_row_bytes + d.probe_key_offset, jd.nchar))) {
                    if (!cnt[u]) first[u] = e[u].idx;
                    ++cnt[u];
                }
                sl[u] = (sl[u] + 1) & mask;
                any = true;
            }
            if (!any) break;
#pragma unroll
            for (int u = 0; u < kProbeU; ++u)
                if (act[u]) e[u] = t[sl[u]];
        }
#pragma unroll
        for (int u = 0; u < kProbeU; ++u) {
            if (i0 + u >= hi) break;
            rec[i0 + u] = ((unsigned long long)cnt[u] << 32) | first[u];
            c += cnt[u];
        }
    }
    c = block_sum256(c);
    if (threadIdx.x == 0) counts[blockIdx.x] = c;
}

// The last block writes the total match count to pinned host memory
// (*host_total, read after the stream synchronizes: no copy on the probe's
// path); rows at or beyond cap_rows are counted but not written.
__global__ __launch_bounds__(256) void k_jwrite(JoinDev jd, const char* __restrict__ build,
                                                const JSlot* __restrict__ t, int logcap,
                                                const char* __restrict__ probe, uint64_t n, uint64_t chunk,
                                                const unsigned long long* __restrict__ counts,
                                                const unsigned long long* __restrict__ rec, char* __restrict__ out,
                                                uint64_t cap_rows, unsigned long long* host_total) {
    const pgpu_hj_desc& d = jd.d;
    const uint64_t mask = (1ull << logcap) - 1;
    const bool a4 = jd.layout_a4 && aligned4((uintptr_t)build | (uintptr_t)probe | (uintptr_t)out);
    unsigned long long pre = 0;
    for (unsigned j = threadIdx.x; j < blockIdx.x; j += 256) pre += counts[j];
    uint64_t base = block_sum256(pre);
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *host_total = base + counts[blockIdx.x];
    const uint64_t lo = (uint64_t)blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
    auto emit = [&](uint64_t pos, const char* prow, uint32_t bidx) {
        if (pos >= cap_rows) return;
        char* o = out + pos * jd.out_row_bytes;
        const char* brow = build + (uint64_t)bidx * d.build_row_bytes;
        for (uint32_t c = 0; c < d.nproj; ++c) {
            const char* src = (d.proj[c].side ? prow : brow) + d.proj[c].src_offset;
            copy_cols(o + jd.dst_off[c], src, d.proj[c].bytes, a4);
        }
    };
    for (uint64_t b0 = lo; b0 < hi; b0 += kProbeTile) {  // whole tiles, so every thread reaches the scans
        const uint64_t i0 = b0 + (uint64_t)threadIdx.x * kProbeU;
        unsigned long long r[kProbeU];
        uint32_t mine = 0;
#pragma unroll
        for (int u = 0; u < kProbeU; ++u) {
            r[u] = i0 + u < hi ? rec[i0 + u] : 0ull;
            mine += (uint32_t)(r[u] >> 32);
        }
        uint32_t total;
        uint64_t pos = base + block_scan256(mine, &total);
        base += total;
        if (!mine) continue;
#pragma unroll
        for (int u = 0; u < kProbeU; ++u) {
            const uint32_t cnt = (uint32_t)(r[u] >> 32);
            if (!cnt) continue;
            const char* prow = probe + (i0 + u) * d.probe_row_bytes;
            if (cnt == 1) {
                emit(pos++, prow, (uint32_t)r[u]);
                continue;
            }
            int64_t k = 0;
            join_key(prow + d.probe_key_offset, d.probe_key_bytes, d.probe_key_type, jd.key_mode, jd.nchar, a4, &k);
            for (uint64_t s = mix64((uint64_t)k) & mask; t[s].idx != kEmptyIdx; s = (s + 1) & mask)
                if (t[s].key == k &&
                    (jd.key_mode != kKeyChr ||
                     chars_equal(build + (uint64_t)t[s].idx * d.build_row_bytes + d.build_key_offset,
                                 prow + d.probe_key_offset, jd.nchar)))
                    emit(pos++, prow, t[s].idx);
        }
    }
}

// ---------------------------------------------------------------- aggregation
constexpr int kMaxKeyWords = 8;  // 64 key bytes

struct AggDev {
    uint32_t row_bytes, nkeys, kind, val_offset;
    uint32_t key_offset[8], key_bytes[8];
    uint32_t key_words;  // ceil(total key bytes / 8)
    uint32_t key_total;
    uint32_t slot_words; // 1 (state) + key_words + 1 (acc)
    uint32_t layout_a4;  // row width, value and key offsets and key widths multiples of 4
};

__device__ __forceinline__ void gather_key(const AggDev& a, const char* row, uint64_t* kw, bool a4) {
#pragma unroll
    for (int w = 0; w < kMaxKeyWords; ++w) kw[w] = 0;
    if (a4) {  // whole 4-byte words, packed in key order
        uint32_t o = 0;
        for (uint32_t c = 0; c < a.nkeys; ++c)
            for (uint32_t b = 0; b < a.key_bytes[c]; b += 4, o += 4)
                kw[o / 8] |= (uint64_t)ld_u32(row + a.key_offset[c] + b, true) << (8 * (o % 8));
        return;
    }
    char buf[kMaxKeyWords * 8];
    uint32_t o = 0;
    for (uint32_t c = 0; c < a.nkeys; ++c) {
        for (uint32_t b = 0; b < a.key_bytes[c]; ++b) buf[o + b] = row[a.key_offset[c] + b];
        o += a.key_bytes[c];
    }
    for (uint32_t b = o; b < a.key_words * 8; ++b) buf[b] = 0;
    for (uint32_t w = 0; w < a.key_words; ++w) __builtin_memcpy(&kw[w], buf + 8 * w, 8);
}

__device__ __forceinline__ void fold(const AggDev& a, unsigned long long* acc, const char* row, bool a4) {
    switch (a.kind) {
        case PGPU_AGG_SUM_I32:
            atomicAdd((unsigned int*)acc, ld_u32(row + a.val_offset, a4));  // low word of the little-endian slot
            break;
        case PGPU_AGG_SUM_I64:
            atomicAdd(acc, (unsigned long long)ld_u64(row + a.val_offset, a4));
            break;
        case PGPU_AGG_SUM_F64:
            atomicAdd((double*)acc, __longlong_as_double((long long)ld_u64(row + a.val_offset, a4)));
            break;
        default:
            atomicAdd(acc, 1ull);
    }
}

// Insert-or-fold of `row` (or of an existing slot's key + partial when
// `partial` is set, used by rehash).  Slot layout in 64-bit words:
// [state][key words...][acc]; state 0 empty, 1 being written, 2 ready.
__device__ void agg_insert(const AggDev& a, unsigned long long* t, uint64_t mask, const uint64_t* kw,
                           const char* row, const unsigned long long* partial,
                           unsigned long long* __restrict__ ngroups, bool a4 = false) {
    uint64_t h = 0x9E3779B97F4A7C15ull;
    for (uint32_t w = 0; w < a.key_words; ++w) h = mix64(h ^ kw[w]);
    uint64_t s = h & mask;
    while (true) {
        unsigned long long* slot = t + s * a.slot_words;
        unsigned long long st = __hip_atomic_load(slot, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if (st == 0) {
            if (atomicCAS(slot, 0ull, 1ull) == 0ull) {
                for (uint32_t w = 0; w < a.key_words; ++w) slot[1 + w] = kw[w];
                __threadfence();
                __hip_atomic_store(slot, 2ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                atomicAdd(ngroups, 1ull);
                st = 2;
            } else {
                continue;  // lost the race: re-read the same slot
            }
        } else if (st == 1) {
            continue;  // the winner stores the key in the same step
        }
        bool eq = true;
        for (uint32_t w = 0; w < a.key_words && eq; ++w) eq = slot[1 + w] == kw[w];
        if (eq) {
            unsigned long long* acc = slot + 1 + a.key_words;
            if (partial) {
                if (a.kind == PGPU_AGG_SUM_F64)
                    atomicAdd((double*)acc, __longlong_as_double((long long)*partial));
                else if (a.kind == PGPU_AGG_SUM_I32)
                    atomicAdd((unsigned int*)acc, (unsigned int)*partial);
                else
                    atomicAdd(acc, *partial);
            } else {
                fold(a, acc, row, a4);
            }
            return;
        }
        s = (s + 1) & mask;
    }
}

__global__ __launch_bounds__(256) void k_agg_append(AggDev a, const char* __restrict__ rows, uint64_t n,
                                                    unsigned long long* __restrict__ t, uint64_t mask,
                                                    unsigned long long* __restrict__ ngroups) {
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    const bool a4 = a.layout_a4 && aligned4((uintptr_t)rows);
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
        const char* row = rows + i * a.row_bytes;
        uint64_t kw[kMaxKeyWords];
        gather_key(a, row, kw, a4);
        agg_insert(a, t, mask, kw, row, nullptr, ngroups, a4);
    }
}

// GROUP BY with few groups: a workgroup first folds its kLocalRows rows in
// an LDS table (kLocalSlots slots, linear probing, the same state protocol
// as agg_insert), then adds each of its groups' partials to the HBM table
// once.  Without it every row is an HBM atomic on its group's accumulator, and
// with a handful of groups (GROUP BY flags, as in TPC-H Q1) those atomics hit a
// few addresses and serialise in L2.  A row whose key finds no LDS slot within
// kLocalProbes steps is folded into the HBM table directly, so any key count is
// correct; the host picks this kernel only while the groups found so far fit
// the LDS table.  Keys of at most KW 64-bit words.
constexpr int kLocalSlots = 1024;
constexpr int kLocalRows = 1024;
constexpr int kLocalProbes = 16;
constexpr uint32_t kLocalMaxKeyWords = 4;

__device__ __forceinline__ void fold_local(const AggDev& a, unsigned long long* acc, const char* row, bool a4) {
    switch (a.kind) {
        case PGPU_AGG_SUM_I32:
            atomicAdd((unsigned int*)acc, ld_u32(row + a.val_offset, a4));  // low word: the sum wraps mod 2^32
            break;
        case PGPU_AGG_SUM_I64:
            atomicAdd(acc, (unsigned long long)ld_u64(row + a.val_offset, a4));
            break;
        case PGPU_AGG_SUM_F64:
            atomicAdd((double*)acc, __longlong_as_double((long long)ld_u64(row + a.val_offset, a4)));
            break;
        default:
            atomicAdd(acc, 1ull);
    }
}

template <int KW>
__global__ __launch_bounds__(256) void k_agg_local(AggDev a, const char* __restrict__ rows, uint64_t n,
                                                   unsigned long long* __restrict__ t, uint64_t mask,
                                                   unsigned long long* __restrict__ ngroups) {
    __shared__ uint64_t lkey[KW][kLocalSlots];
    __shared__ unsigned long long lacc[kLocalSlots];
    __shared__ uint32_t lst[kLocalSlots];
    for (int i = threadIdx.x; i < kLocalSlots; i += 256) {
        lst[i] = 0;
        lacc[i] = 0;
    }
    __syncthreads();
    const bool a4 = a.layout_a4 && aligned4((uintptr_t)rows);
    const uint64_t r0 = (uint64_t)blockIdx.x * kLocalRows;
    const uint64_t r1 = min(n, r0 + kLocalRows);
    for (uint64_t i = r0 + threadIdx.x; i < r1; i += 256) {
        const char* row = rows + i * a.row_bytes;
        uint64_t kw[kMaxKeyWords];
        gather_key(a, row, kw, a4);  // words past key_words are zero
        uint64_t h = 0x9E3779B97F4A7C15ull;
        for (uint32_t w = 0; w < a.key_words; ++w) h = mix64(h ^ kw[w]);
        uint32_t sl = (uint32_t)(h >> 40) & (kLocalSlots - 1);
        bool done = false;
        for (int probe = 0; probe < kLocalProbes && !done;) {
            uint32_t st = __hip_atomic_load(&lst[sl], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (st == 0) {
                if (atomicCAS(&lst[sl], 0u, 1u) != 0u) continue;  // lost the race: re-read the slot
#pragma unroll
                for (int w = 0; w < KW; ++w) lkey[w][sl] = kw[w];
                __hip_atomic_store(&lst[sl], 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            } else if (st == 1) {
                continue;  // the winner stores the key in the same step
            }
            bool eq = true;
#pragma unroll
            for (int w = 0; w < KW; ++w) eq = eq && lkey[w][sl] == kw[w];
            if (eq) {
                fold_local(a, &lacc[sl], row, a4);
                done = true;
            } else {
                sl = (sl + 1) & (kLocalSlots - 1);
                ++probe;
            }
        }
        if (!done) agg_insert(a, t, mask, kw, row, nullptr, ngroups, a4);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kLocalSlots; i += 256) {
        if (lst[i] != 2) continue;
        uint64_t kw[kMaxKeyWords] = {};
#pragma unroll
        for (int w = 0; w < KW; ++w) kw[w] = lkey[w][i];
        const unsigned long long part = lacc[i];
        agg_insert(a, t, mask, kw, nullptr, &part, ngroups);
    }
}

// No GROUP BY: one group, folded in two launches with no same-address
// atomics.  k_agg_partial: each workgroup sums its rows (in the accumulator's
// type, so an int32 sum still wraps mod 2^32; several independent loads per
// thread in flight) into partial[blockIdx.x].  k_agg_single: one workgroup
// sums the partials in a fixed order and folds the total into the single
// slot.
constexpr unsigned kAggPartials = 1024;
constexpr int kAggUnroll = 4;

__device__ __forceinline__ void acc_row(const AggDev& a, const char* row, bool a4, unsigned long long& acc,
                                        double& accf) {
    switch (a.kind) {
        case PGPU_AGG_SUM_I32:
            acc += ld_u32(row + a.val_offset, a4);  // low 32 bits are the wrapped int32 sum
            break;
        case PGPU_AGG_SUM_I64:
            acc += ld_u64(row + a.val_offset, a4);
            break;
        case PGPU_AGG_SUM_F64:
            accf += __longlong_as_double((long long)ld_u64(row + a.val_offset, a4));
            break;
        default:
            acc += 1;
    }
}

// Sums of (acc, accf) over a 256-thread block, valid in thread 0.
__device__ __forceinline__ void block_sum2(unsigned long long& acc, double& accf) {
    __shared__ unsigned long long red[256 / 64];
    __shared__ double redf[256 / 64];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        acc += __shfl_xor(acc, d, 64);
        accf += __shfl_xor(accf, d, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = acc;
        redf[threadIdx.x >> 6] = accf;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        acc = red[0] + red[1] + red[2] + red[3];
        accf = ((redf[0] + redf[1]) + redf[2]) + redf[3];
    }
}

__global__ __launch_bounds__(256) void k_agg_partial(AggDev a, const char* __restrict__ rows, uint64_t n,
                                                     unsigned long long* __restrict__ partial) {
    const bool a4 = a.layout_a4 && aligned4((uintptr_t)rows);
    unsigned long long acc = 0;
    double accf = 0;
    const uint64_t step = (uint64_t)gridDim.x * 256 * kAggUnroll;
    for (uint64_t i0 = (uint64_t)blockIdx.x * 256 * kAggUnroll + threadIdx.x; i0 < n; i0 += step) {
        if (i0 + 256 * (kAggUnroll - 1) < n) {
            // all loads issued before the first add
            uint64_t v[kAggUnroll];
#pragma unroll
            for (int u = 0; u < kAggUnroll; ++u) {
                const char* p = rows + (i0 + 256 * u) * a.row_bytes + a.val_offset;
                v[u] = a.kind == PGPU_AGG_COUNT ? 1u : a.kind == PGPU_AGG_SUM_I32 ? ld_u32(p, a4) : ld_u64(p, a4);
            }
#pragma unroll
            for (int u = 0; u < kAggUnroll; ++u) {
                if (a.kind == PGPU_AGG_SUM_F64)
                    accf += __longlong_as_double((long long)v[u]);
                else
                    acc += v[u];
            }
        } else {
            for (uint64_t i = i0; i < n; i += 256) acc_row(a, rows + i * a.row_bytes, a4, acc, accf);
        }
    }
    block_sum2(acc, accf);
    if (threadIdx.x == 0)
        partial[blockIdx.x] = a.kind == PGPU_AGG_SUM_F64 ? (unsigned long long)__double_as_longlong(accf) : acc;
}

__global__ __launch_bounds__(256) void k_agg_single(AggDev a, const unsigned long long* __restrict__ partial,
                                                    unsigned np, unsigned long long* __restrict__ t, uint64_t mask,
                                                    unsigned long long* __restrict__ ngroups) {
    unsigned long long acc = 0;
    double accf = 0;
    for (unsigned j = threadIdx.x; j < np; j += 256) {
        if (a.kind == PGPU_AGG_SUM_F64)
            accf += __longlong_as_double((long long)partial[j]);
        else
            acc += partial[j];
    }
    block_sum2(acc, accf);
    if (threadIdx.x == 0) {
        unsigned long long p = a.kind == PGPU_AGG_SUM_F64 ? (unsigned long long)__double_as_longlong(accf) : acc;
        uint64_t kw[kMaxKeyWords] = {};
        agg_insert(a, t, mask, kw, nullptr, &p, ngroups);
    }
}

__global__ __launch_bounds__(256) void k_agg_rehash(AggDev a, const unsigned long long* __restrict__ old,
                                                    uint64_t oldcap, unsigned long long* __restrict__ t,
                                                    uint64_t mask, unsigned long long* __restrict__ ngroups) {
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < oldcap; i += stride) {
        const unsigned long long* slot = old + i * a.slot_words;
        if (slot[0] != 2) continue;
        uint64_t kw[kMaxKeyWords];
        for (int w = 0; w < kMaxKeyWords; ++w) kw[w] = w < (int)a.key_words ? slot[1 + w] : 0;
        agg_insert(a, t, mask, kw, nullptr, slot + 1 + a.key_words, ngroups);
    }
}

__global__ __launch_bounds__(256) void k_fill_u64(unsigned long long* p, uint64_t n, unsigned long long v) {
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) p[i] = v;
}

__global__ __launch_bounds__(256) void k_agg_compact(AggDev a, const unsigned long long* __restrict__ t,
                                                     uint64_t cap, char* __restrict__ out, uint32_t out_row_bytes,
                                                     uint32_t acc_bytes, unsigned long long* __restrict__ cursor) {
    __shared__ unsigned long long s_base;
    for (uint64_t b0 = (uint64_t)blockIdx.x * 256; b0 < cap; b0 += (uint64_t)gridDim.x * 256) {
        const uint64_t i = b0 + threadIdx.x;
        const unsigned long long* slot = t + (i < cap ? i : 0) * a.slot_words;
        const uint32_t live = (i < cap && slot[0] == 2) ? 1u : 0u;
        uint32_t total;
        const uint32_t off = block_scan256(live, &total);
        if (threadIdx.x == 0) s_base = total ? atomicAdd(cursor, (unsigned long long)total) : 0ull;
        __syncthreads();
        if (live) {
            char* o = out + (s_base + off) * out_row_bytes;
            const char* kb = (const char*)(slot + 1);
            for (uint32_t b = 0; b < a.key_total; ++b) o[b] = kb[b];
            const char* ab = (const char*)(slot + 1 + a.key_words);
            for (uint32_t b = 0; b < acc_bytes; ++b) o[a.key_total + b] = ab[b];
        }
        __syncthreads();
    }
}

inline unsigned grid_for(uint64_t n) { return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(pgpu::ceil_div(n, 256), 8192)); }

}  // namespace

// ====================================================================== join
namespace {
struct Stage {  // a pinned host staging block from the context's pool
    char* p = nullptr;
    uint64_t cap = 0;
};
}  // namespace

struct pgpu_hj {
    pgpu_ctx* ctx = nullptr;
    JoinDev jd{};
    std::mutex mu;
    std::vector<char> staged;
    uint64_t nbuild = 0;
    bool built = false;
    char* d_build = nullptr;
    bool build_external = false;  // d_build is the caller's device buffer
    JSlot* d_table = nullptr;
    int logcap = 0;
    struct Slot {
        hipStream_t st = nullptr;
        char* d_in = nullptr;
        uint64_t in_cap = 0;
        char* d_out = nullptr;
        uint64_t out_cap = 0;  // rows
        Stage hin, hout;        // pinned staging of host probe rows / pulled output rows
        char* d_rec = nullptr;  // per probe row: match count << 32 | first match
        uint64_t rec_cap = 0;
        unsigned long long* d_cursor = nullptr;  // per-block match counts
        unsigned long long* h_cursor = nullptr;
        uint64_t nout = 0, pulled = 0;
    };
    std::vector<Slot> slots;
    // Updated by every worker probing its own slot, hence atomic.
    struct {
        std::atomic<uint64_t> build_rows{0}, probe_rows{0}, output_rows{0}, table_slots{0};
        std::atomic<uint64_t> h2d_bytes{0}, d2h_bytes{0}, probe_reruns{0};
    } stats;
};

using namespace pgpu;

namespace {
int slot_reserve(pgpu_ctx* ctx, char** p, uint64_t* cap, uint64_t need) {
    if (*cap >= need && *p) return PGPU_OK;
    if (*p) PGPU_TRY(ctx_free_reuse(ctx, *p));
    *p = nullptr;
    *cap = 0;
    uint64_t c = std::max<uint64_t>(need, 4096);
    PGPU_TRY(ctx_alloc_reuse(ctx, c, (void**)p));
    *cap = c;
    return PGPU_OK;
}
int hj_build_table(pgpu_hj* h);

// Pinned staging of host rows: pageable copies go through the runtime's own
// bounce buffer one at a time, so many workers' copies queue behind each
// other.  Rows are copied into a pinned block in 4 MiB pieces, each piece's
// DMA queued on `st` as soon as it is copied (the CPU copy of the next piece
// overlaps it).  The stream must have finished with the previous contents.
int stage_reserve(pgpu_ctx* ctx, Stage& sg, uint64_t bytes) {
    if (sg.cap >= bytes && sg.p) return PGPU_OK;
    ctx_host_free_reuse(ctx, sg.p);
    sg.p = nullptr;
    sg.cap = 0;
    return ctx_host_alloc_reuse(ctx, bytes, (void**)&sg.p, &sg.cap);
}

constexpr uint64_t kStageMax = 32ull << 20;  // pinned staging per user: larger copies go in rounds

int stage_h2d(pgpu_ctx* ctx, Stage& sg, hipStream_t st, char* d_dst, const char* src, uint64_t bytes) {
    if (!bytes) return PGPU_OK;
    PGPU_TRY(stage_reserve(ctx, sg, std::min(bytes, kStageMax)));
    constexpr uint64_t kPiece = 4ull << 20;
    for (uint64_t r = 0; r < bytes; r += sg.cap) {
        if (r) PGPU_HIP(hipStreamSynchronize(st));  // the previous round's DMAs have read the staging
        const uint64_t rb = std::min(sg.cap, bytes - r);
        for (uint64_t o = 0; o < rb; o += kPiece) {
            const uint64_t b = std::min(kPiece, rb - o);
            memcpy(sg.p + o, src + r + o, b);
            PGPU_HIP(hipMemcpyAsync(d_dst + r + o, sg.p + o, b, hipMemcpyHostToDevice, st));
        }
    }
    return PGPU_OK;
}

// Device rows to host memory through the pinned staging (a DMA, then a CPU copy, per round).
int stage_d2h(pgpu_ctx* ctx, Stage& sg, hipStream_t st, char* dst, const char* d_src, uint64_t bytes) {
    if (!bytes) return PGPU_OK;
    PGPU_TRY(stage_reserve(ctx, sg, std::min(bytes, kStageMax)));
    for (uint64_t r = 0; r < bytes; r += sg.cap) {
        const uint64_t rb = std::min(sg.cap, bytes - r);
        PGPU_HIP(hipMemcpyAsync(sg.p, d_src + r, rb, hipMemcpyDeviceToHost, st));
        PGPU_HIP(hipStreamSynchronize(st));
        memcpy(dst + r, sg.p, rb);
    }
    return PGPU_OK;
}
}  // namespace

extern "C" {

int pgpu_hj_create(pgpu_ctx* ctx, const pgpu_hj_desc* d, pgpu_hj** out) {
    if (!ctx || !d || !out) return fail(PGPU_EINVAL, "pgpu_hj_create: null argument");
    *out = nullptr;
    // key pairs Comparator::init accepts (comparator.cpp:43-413)
    const uint32_t bt = d->build_key_type, pt = d->probe_key_type;
    auto is_int = [](uint32_t t, uint32_t b) {
        return (t == 0 || t == PGPU_COL_INT32 || t == PGPU_COL_INT64) && (b == 4 || b == 8) &&
               (t != PGPU_COL_INT32 || b == 4) && (t != PGPU_COL_INT64 || b == 8);
    };
    auto is_dbl = [](uint32_t t, uint32_t b) { return t == PGPU_COL_F64 && b == 8; };
    auto is_chr = [](uint32_t t, uint32_t b) { return t == PGPU_COL_CHAR && b >= 1; };
    uint32_t mode;
    if (is_int(bt, d->build_key_bytes) && is_int(pt, d->probe_key_bytes))
        mode = kKeyInt;
    else if ((is_int(bt, d->build_key_bytes) || is_dbl(bt, d->build_key_bytes)) &&
             (is_int(pt, d->probe_key_bytes) || is_dbl(pt, d->probe_key_bytes)))
        mode = kKeyDbl;
    else if (is_chr(bt, d->build_key_bytes) && is_chr(pt, d->probe_key_bytes))
        mode = kKeyChr;
    else
        return fail(PGPU_EUNSUPPORTED,
                    "join keys: integers (4 or 8 bytes), dec against dec or integer, or char against char");
    if (d->build_key_offset + d->build_key_bytes > d->build_row_bytes ||
        d->probe_key_offset + d->probe_key_bytes > d->probe_row_bytes)
        return fail(PGPU_EINVAL, "join key outside its row");
    if (d->nproj > PGPU_MAX_PROJ || d->slots == 0 || d->slots > 1024)
        return fail(PGPU_EINVAL, "bad projection count or slot count");
    auto* h = new pgpu_hj();
    h->ctx = ctx;
    h->jd.d = *d;
    uint32_t o = 0;
    for (uint32_t c = 0; c < d->nproj; ++c) {
        const auto& p = d->proj[c];
        const uint32_t lim = p.side ? d->probe_row_bytes : d->build_row_bytes;
        if (p.side > 1 || p.src_offset + p.bytes > lim) {
            delete h;
            return fail(PGPU_EINVAL, "projected column %u outside its row", c);
        }
        h->jd.dst_off[c] = o;
        o += p.bytes;
    }
    h->jd.out_row_bytes = o;
    h->jd.key_mode = mode;
    h->jd.nchar = std::min(d->build_key_bytes, d->probe_key_bytes);
    uint32_t lay = d->build_row_bytes | d->build_key_offset | d->probe_row_bytes | d->probe_key_offset | o;
    for (uint32_t c = 0; c < d->nproj; ++c) lay |= d->proj[c].src_offset | d->proj[c].bytes | h->jd.dst_off[c];
    h->jd.layout_a4 = aligned4(lay) ? 1u : 0u;
    h->slots.resize(d->slots);
    *out = h;
    return PGPU_OK;
}

int pgpu_hj_destroy(pgpu_hj* h) {
    if (!h) return PGPU_OK;
    (void)ctx_activate(h->ctx);
    // every stream is synchronized before its buffers go back to the pools
    bool synced = true;
    for (auto& s : h->slots)
        if (s.st) synced = hipStreamSynchronize(s.st) == hipSuccess && synced;
    synced = hipStreamSynchronize(h->ctx->stream) == hipSuccess && synced;
    auto drop = [&](void* p) { (void)(synced ? ctx_free_reuse(h->ctx, p) : ctx_free(h->ctx, p)); };
    for (auto& s : h->slots) {
        if (s.st && synced) ctx_stream_put(h->ctx, s.st);
        else if (s.st) (void)hipStreamDestroy(s.st);
        if (s.d_in) drop(s.d_in);
        if (s.d_out) drop(s.d_out);
        if (s.d_cursor) drop(s.d_cursor);
        if (s.d_rec) drop(s.d_rec);
        if (synced) {
            ctx_host_free_reuse(h->ctx, s.hin.p);
            ctx_host_free_reuse(h->ctx, s.hout.p);
        }
        if (s.h_cursor) ctx_host_word_put(h->ctx, s.h_cursor);
    }
    if (h->d_build && !h->build_external) drop(h->d_build);
    if (h->d_table) drop(h->d_table);
    delete h;
    return PGPU_OK;
}

int pgpu_hj_build_append(pgpu_hj* h, const void* rows, uint64_t n) {
    if (!h || (n && !rows)) return fail(PGPU_EINVAL, "pgpu_hj_build_append: null argument");
    std::lock_guard<std::mutex> g(h->mu);
    if (h->built) return fail(PGPU_ESTATE, "pgpu_hj_build_append after build_finish");
    const uint64_t b = n * h->jd.d.build_row_bytes;
    h->staged.insert(h->staged.end(), (const char*)rows, (const char*)rows + b);
    h->nbuild += n;
    return PGPU_OK;
}

int pgpu_hj_build_finish(pgpu_hj* h) {
    if (!h) return fail(PGPU_EINVAL, "pgpu_hj_build_finish: null handle");
    std::lock_guard<std::mutex> g(h->mu);
    if (h->built) return PGPU_OK;
    PGPU_TRY(ctx_activate(h->ctx));
    if (h->nbuild >= kEmptyIdx) return fail(PGPU_EINVAL, "build side exceeds 2^32-1 rows");
    hipStream_t st = h->ctx->stream;
    const uint64_t bytes = h->staged.size();
    PGPU_TRY(ctx_alloc_reuse(h->ctx, std::max<uint64_t>(bytes, 8), (void**)&h->d_build));
    if (bytes) PGPU_HIP(hipMemcpyAsync(h->d_build, h->staged.data(), bytes, hipMemcpyHostToDevice, st));
    PGPU_TRY(hj_build_table(h));
    h->stats.h2d_bytes += bytes;
    std::vector<char>().swap(h->staged);
    h->built = true;
    return PGPU_OK;
}

int pgpu_hj_build_device(pgpu_hj* h, const void* d_rows, uint64_t n) {
    if (!h || (n && !d_rows)) return fail(PGPU_EINVAL, "pgpu_hj_build_device: null argument");
    std::lock_guard<std::mutex> g(h->mu);
    if (h->built || h->nbuild) return fail(PGPU_ESTATE, "pgpu_hj_build_device: the table already has build rows");
    if (n >= kEmptyIdx) return fail(PGPU_EINVAL, "build side exceeds 2^32-1 rows");
    PGPU_TRY(ctx_activate(h->ctx));
    h->d_build = (char*)d_rows;
    h->build_external = true;
    h->nbuild = n;
    PGPU_TRY(hj_build_table(h));
    h->built = true;
    return PGPU_OK;
}

}  // extern "C"

namespace {
// Claims and fills the open-addressing table over h->d_build's h->nbuild rows.
int hj_build_table(pgpu_hj* h) {
    hipStream_t st = h->ctx->stream;
    // load factor at most 1/4 (1/2 beyond 2^26 build rows): short collision chains
    h->logcap = (int)ilog2_ceil(std::max<uint64_t>((h->nbuild <= (1ull << 26) ? 4 : 2) * h->nbuild, 16));
    const uint64_t cap = 1ull << h->logcap;
    PGPU_TRY(ctx_alloc_reuse(h->ctx, cap * sizeof(JSlot), (void**)&h->d_table));
    hipLaunchKernelGGL(k_jclear, dim3(grid_for(cap)), dim3(256), 0, st, h->d_table, cap);
    PGPU_LAUNCH_CHECK();
    if (h->nbuild) {
        hipLaunchKernelGGL(k_jbuild, dim3(grid_for(h->nbuild)), dim3(256), 0, st, h->jd, h->d_build, h->nbuild,
                           h->d_table, h->logcap);
        PGPU_LAUNCH_CHECK();
    }
    PGPU_HIP(hipStreamSynchronize(st));
    h->stats.build_rows = h->nbuild;
    h->stats.table_slots = cap;
    return PGPU_OK;
}

// Probes n rows that are already in device memory (d_in) on the slot's
// stream; the output rows stay in the slot's device buffer.
int hj_probe_rows(pgpu_hj* h, pgpu_hj::Slot& s, const char* d_in, uint64_t n, uint64_t* n_out) {
    const uint32_t orb = std::max<uint32_t>(h->jd.out_row_bytes, 1);
    uint64_t cap_rows = s.out_cap / orb;
    if (cap_rows < n) {
        PGPU_TRY(slot_reserve(h->ctx, &s.d_out, &s.out_cap, n * orb));
        cap_rows = s.out_cap / orb;
    }
    // contiguous chunks of whole tiles, at most kProbeBlocks of them
    const uint64_t chunk = ceil_div(ceil_div(n, kProbeBlocks), kProbeTile) * kProbeTile;
    const unsigned nb = (unsigned)ceil_div(n, chunk);
    PGPU_TRY(slot_reserve(h->ctx, &s.d_rec, &s.rec_cap, n * 8));
    const auto* rec = (const unsigned long long*)s.d_rec;
    hipLaunchKernelGGL(k_jcount, dim3(nb), dim3(256), 0, s.st, h->jd, h->d_build, h->d_table, h->logcap, d_in, n,
                       chunk,
                       s.d_cursor, (unsigned long long*)s.d_rec);
    PGPU_LAUNCH_CHECK();
    for (int attempt = 0; attempt < 2; ++attempt) {
        hipLaunchKernelGGL(k_jwrite, dim3(nb), dim3(256), 0, s.st, h->jd, h->d_build, h->d_table, h->logcap, d_in,
                           n, chunk, s.d_cursor, rec, s.d_out, cap_rows, s.h_cursor);
        PGPU_LAUNCH_CHECK();
        PGPU_HIP(hipStreamSynchronize(s.st));
        const uint64_t total = *s.h_cursor;
        if (total <= cap_rows) {
            s.nout = total;
            break;
        }
        // Output larger than the buffer: grow to the exact size and write again
        // (the per-block counts stay valid).
        PGPU_TRY(slot_reserve(h->ctx, &s.d_out, &s.out_cap, total * orb));
        cap_rows = s.out_cap / orb;
        h->stats.probe_reruns++;
        if (attempt == 1) return fail(PGPU_EOVERFLOW, "probe output did not fit after regrowth");
    }
    h->stats.probe_rows += n;
    h->stats.output_rows += s.nout;
    *n_out = s.nout;
    return PGPU_OK;
}

// Checks a probe call and prepares the slot; *skip when there is nothing to join.
int hj_probe_begin(pgpu_hj* h, uint32_t slot, uint64_t n, uint64_t* n_out, bool* skip) {
    if (slot >= h->slots.size()) return fail(PGPU_EINVAL, "pgpu_hj_probe: slot %u out of range", slot);
    if (!h->built) return fail(PGPU_ESTATE, "pgpu_hj_probe before build_finish");
    PGPU_TRY(ctx_activate(h->ctx));
    auto& s = h->slots[slot];
    if (!s.st) PGPU_TRY(ctx_stream_get(h->ctx, &s.st));
    if (!s.d_cursor) PGPU_TRY(ctx_alloc_reuse(h->ctx, 8 * kProbeBlocks, (void**)&s.d_cursor));
    if (!s.h_cursor) PGPU_TRY(ctx_host_word(h->ctx, &s.h_cursor));
    s.nout = s.pulled = 0;
    *n_out = 0;
    *skip = n == 0 || h->nbuild == 0;
    if (*skip) h->stats.probe_rows += n;
    return PGPU_OK;
}
}  // namespace

extern "C" {

int pgpu_hj_probe_device(pgpu_hj* h, uint32_t slot, const void* d_rows, uint64_t n, uint64_t* n_out) {
    if (!h || !n_out || (n && !d_rows)) return fail(PGPU_EINVAL, "pgpu_hj_probe_device: null argument");
    bool skip = false;
    PGPU_TRY(hj_probe_begin(h, slot, n, n_out, &skip));
    if (skip) return PGPU_OK;
    return hj_probe_rows(h, h->slots[slot], (const char*)d_rows, n, n_out);
}

int pgpu_hj_output(pgpu_hj* h, uint32_t slot, const void** d_rows, uint64_t* n_rows) {
    if (!h || !d_rows || !n_rows) return fail(PGPU_EINVAL, "pgpu_hj_output: null argument");
    if (slot >= h->slots.size()) return fail(PGPU_EINVAL, "pgpu_hj_output: slot out of range");
    auto& s = h->slots[slot];
    *d_rows = s.d_out ? s.d_out + s.pulled * (uint64_t)h->jd.out_row_bytes : nullptr;
    *n_rows = s.nout - s.pulled;
    s.pulled = s.nout;
    return PGPU_OK;
}

int pgpu_hj_probe(pgpu_hj* h, uint32_t slot, const void* rows, uint64_t n, uint64_t* n_out) {
    if (!h || !n_out || (n && !rows)) return fail(PGPU_EINVAL, "pgpu_hj_probe: null argument");
    bool skip = false;
    PGPU_TRY(hj_probe_begin(h, slot, n, n_out, &skip));
    if (skip) return PGPU_OK;
    auto& s = h->slots[slot];
    const uint32_t prb = h->jd.d.probe_row_bytes;
    PGPU_TRY(slot_reserve(h->ctx, &s.d_in, &s.in_cap, n * prb));
    PGPU_TRY(stage_h2d(h->ctx, s.hin, s.st, s.d_in, (const char*)rows, n * prb));
    h->stats.h2d_bytes += n * prb;
    return hj_probe_rows(h, s, s.d_in, n, n_out);
}

int pgpu_hj_pull(pgpu_hj* h, uint32_t slot, void* dst, uint64_t cap_rows, uint64_t* n_rows) {
    if (!h || !n_rows || (cap_rows && !dst)) return fail(PGPU_EINVAL, "pgpu_hj_pull: null argument");
    if (slot >= h->slots.size()) return fail(PGPU_EINVAL, "pgpu_hj_pull: slot out of range");
    auto& s = h->slots[slot];
    const uint64_t k = std::min(cap_rows, s.nout - s.pulled);
    *n_rows = k;
    if (!k) return PGPU_OK;
    PGPU_TRY(ctx_activate(h->ctx));
    const uint64_t orb = h->jd.out_row_bytes;
    PGPU_TRY(stage_d2h(h->ctx, s.hout, s.st, (char*)dst, s.d_out + s.pulled * orb, k * orb));
    s.pulled += k;
    h->stats.d2h_bytes += k * orb;
    return PGPU_OK;
}

int pgpu_hj_stats_get(pgpu_hj* h, pgpu_hj_stats* out) {
    if (!h || !out) return fail(PGPU_EINVAL, "pgpu_hj_stats_get: null argument");
    out->build_rows = h->stats.build_rows.load();
    out->probe_rows = h->stats.probe_rows.load();
    out->output_rows = h->stats.output_rows.load();
    out->table_slots = h->stats.table_slots.load();
    out->h2d_bytes = h->stats.h2d_bytes.load();
    out->d2h_bytes = h->stats.d2h_bytes.load();
    out->probe_reruns = h->stats.probe_reruns.load();
    return PGPU_OK;
}

}  // extern "C"

// ====================================================================== aggregation
struct pgpu_agg {
    pgpu_ctx* ctx = nullptr;
    AggDev a{};
    std::mutex mu;
    unsigned long long* d_table = nullptr;
    uint64_t cap = 0;
    unsigned long long* d_ngroups = nullptr;
    unsigned long long* h_ngroups = nullptr;
    char* d_in = nullptr;
    uint64_t in_cap = 0;
    char* d_out = nullptr;
    uint64_t out_cap = 0;
    unsigned long long* d_partial = nullptr;  // per-workgroup sums of the single-group fold
    Stage hin;                                // pinned staging of appended host rows
    uint64_t ngroups_out = 0;
    uint32_t out_row_bytes = 0, acc_bytes = 0;
    bool finished = false;
    // Each aggregate runs on a stream of its own, so that the per-thread
    // aggregates of one query fold concurrently instead of queueing on the
    // context's stream.
    hipStream_t st = nullptr;
    hipEvent_t ctx_done = nullptr;  // orders device-row folds after the context's stream
};

namespace {
int agg_alloc_table(pgpu_agg* g, uint64_t cap, unsigned long long** t) {
    const uint64_t words = cap * g->a.slot_words;
    PGPU_TRY(ctx_alloc_reuse(g->ctx, words * 8, (void**)t));
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(words)), dim3(256), 0, g->st, *t, words, 0ull);
    PGPU_LAUNCH_CHECK();
    return PGPU_OK;
}

int agg_grow(pgpu_agg* g, uint64_t newcap) {
    unsigned long long* nt = nullptr;
    PGPU_TRY(agg_alloc_table(g, newcap, &nt));
    hipStream_t st = g->st;
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(256), 0, st, g->d_ngroups, 1, 0ull);
    PGPU_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_agg_rehash, dim3(grid_for(g->cap)), dim3(256), 0, st, g->a, g->d_table, g->cap, nt,
                       newcap - 1, g->d_ngroups);
    PGPU_LAUNCH_CHECK();
    PGPU_HIP(hipStreamSynchronize(st));
    PGPU_TRY(ctx_free_reuse(g->ctx, g->d_table));
    g->d_table = nt;
    g->cap = newcap;
    return PGPU_OK;
}
}  // namespace

extern "C" {

int pgpu_agg_create(pgpu_ctx* ctx, const pgpu_agg_desc* d, pgpu_agg** out) {
    if (!ctx || !d || !out) return fail(PGPU_EINVAL, "pgpu_agg_create: null argument");
    *out = nullptr;
    if (d->nkeys > 8) return fail(PGPU_EINVAL, "at most 8 group columns");
    uint32_t kt = 0;
    for (uint32_t c = 0; c < d->nkeys; ++c) {
        if (d->key_offset[c] + d->key_bytes[c] > d->row_bytes) return fail(PGPU_EINVAL, "group column outside the row");
        kt += d->key_bytes[c];
    }
    if (kt > kMaxKeyWords * 8) return fail(PGPU_EUNSUPPORTED, "group key wider than 64 bytes");
    if (d->kind < PGPU_AGG_SUM_I32 || d->kind > PGPU_AGG_COUNT) return fail(PGPU_EINVAL, "unknown aggregate kind");
    const uint32_t vb = d->kind == PGPU_AGG_SUM_I32 ? 4 : 8;
    if (d->kind != PGPU_AGG_COUNT && d->val_offset + vb > d->row_bytes)
        return fail(PGPU_EINVAL, "summed column outside the row");
    PGPU_TRY(ctx_activate(ctx));
    auto* g = new pgpu_agg();
    g->ctx = ctx;
    g->a.row_bytes = d->row_bytes;
    g->a.nkeys = d->nkeys;
    g->a.kind = d->kind;
    g->a.val_offset = d->val_offset;
    for (uint32_t c = 0; c < d->nkeys; ++c) {
        g->a.key_offset[c] = d->key_offset[c];
        g->a.key_bytes[c] = d->key_bytes[c];
    }
    g->a.key_total = kt;
    g->a.key_words = (kt + 7) / 8;
    g->a.slot_words = 2 + g->a.key_words;
    uint32_t lay = d->row_bytes | (d->kind != PGPU_AGG_COUNT ? d->val_offset : 0u);
    for (uint32_t c = 0; c < d->nkeys; ++c) lay |= d->key_offset[c] | d->key_bytes[c];
    g->a.layout_a4 = aligned4(lay) ? 1u : 0u;
    g->acc_bytes = vb;
    g->out_row_bytes = kt + vb;
    g->cap = 1ull << ilog2_ceil(std::max<uint64_t>(2 * (d->groups_hint ? d->groups_hint : 32768), 64));
    int rc = ctx_stream_get(ctx, &g->st);
    if (rc == PGPU_OK && hipEventCreateWithFlags(&g->ctx_done, hipEventDisableTiming) != hipSuccess) {
        g->ctx_done = nullptr;
        rc = fail(PGPU_EDEVICE, "hipEventCreateWithFlags failed");
    }
    if (rc == PGPU_OK) rc = agg_alloc_table(g, g->cap, &g->d_table);
    // the compacted output for as many groups as the table holds before it
    // grows, so that finish allocates nothing unless the table grew
    if (rc == PGPU_OK) rc = slot_reserve(ctx, &g->d_out, &g->out_cap, (g->cap / 2) * g->out_row_bytes);
    if (rc == PGPU_OK) rc = ctx_alloc_reuse(ctx, 8, (void**)&g->d_ngroups);
    if (rc == PGPU_OK && d->nkeys == 0) rc = ctx_alloc_reuse(ctx, 8 * kAggPartials, (void**)&g->d_partial);
    if (rc == PGPU_OK) rc = ctx_host_word(ctx, &g->h_ngroups);
    if (rc == PGPU_OK) {
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(256), 0, g->st, g->d_ngroups, 1, 0ull);
        if (hipGetLastError() != hipSuccess) rc = fail(PGPU_EDEVICE, "launch failed");
    }
    if (rc != PGPU_OK) {
        pgpu_agg_destroy(g);
        return rc;
    }
    *out = g;
    return PGPU_OK;
}

int pgpu_agg_destroy(pgpu_agg* g) {
    if (!g) return PGPU_OK;
    (void)ctx_activate(g->ctx);
    // the stream is synchronized before its buffers go back to the pools
    const bool synced = !g->st || hipStreamSynchronize(g->st) == hipSuccess;
    auto drop = [&](void* p) { (void)(synced ? ctx_free_reuse(g->ctx, p) : ctx_free(g->ctx, p)); };
    if (g->d_table) drop(g->d_table);
    if (g->d_ngroups) drop(g->d_ngroups);
    if (g->h_ngroups) ctx_host_word_put(g->ctx, g->h_ngroups);
    if (g->d_in) drop(g->d_in);
    if (g->d_out) drop(g->d_out);
    if (g->d_partial) drop(g->d_partial);
    if (synced) ctx_host_free_reuse(g->ctx, g->hin.p);
    if (g->st && synced) ctx_stream_put(g->ctx, g->st);
    else if (g->st) (void)hipStreamDestroy(g->st);
    if (g->ctx_done) (void)hipEventDestroy(g->ctx_done);
    delete g;
    return PGPU_OK;
}

}  // extern "C"

namespace {
// Folds n device-resident rows into the table (caller holds g->mu; the
// context is current).  `upload`: rows are on the host and go through d_in.
//
// With a GROUP BY the rows are folded in pieces of at most cap/2 rows, and
// before each piece the table doubles while the groups already seen fill
// more than a quarter of it: even if every row of a piece were a new group
// the table stays at most three quarters full (an insert always finds a
// free slot), while its size follows the distinct groups, not the rows
// appended.  Pieces are at least kAggMinPiece rows (the table is allowed
// that much headroom) so that few groups over many rows take few launches.
constexpr uint64_t kAggMinPiece = 1ull << 18;

// ng: groups in the HBM table before this piece.
int agg_fold_piece(pgpu_agg* g, const char* src, uint64_t n, uint64_t ng) {
    hipStream_t st = g->st;
    if (g->a.nkeys && g->a.key_words <= kLocalMaxKeyWords && ng <= (uint64_t)kLocalSlots / 2) {
        const dim3 grid((unsigned)ceil_div(n, kLocalRows));
        const uint32_t kw = g->a.key_words;
        if (kw == 1)
            hipLaunchKernelGGL(k_agg_local<1>, grid, dim3(256), 0, st, g->a, src, n, g->d_table, g->cap - 1,
                               g->d_ngroups);
        else if (kw == 2)
            hipLaunchKernelGGL(k_agg_local<2>, grid, dim3(256), 0, st, g->a, src, n, g->d_table, g->cap - 1,
                               g->d_ngroups);
        else
            hipLaunchKernelGGL(k_agg_local<4>, grid, dim3(256), 0, st, g->a, src, n, g->d_table, g->cap - 1,
                               g->d_ngroups);
        PGPU_LAUNCH_CHECK();
        return PGPU_OK;
    }
    if (g->a.nkeys == 0) {
        const unsigned np = (unsigned)std::min<uint64_t>(ceil_div(n, 256 * kAggUnroll), kAggPartials);
        hipLaunchKernelGGL(k_agg_partial, dim3(np), dim3(256), 0, st, g->a, src, n, g->d_partial);
        PGPU_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_agg_single, dim3(1), dim3(256), 0, st, g->a, g->d_partial, np, g->d_table, g->cap - 1,
                           g->d_ngroups);
    } else
        hipLaunchKernelGGL(k_agg_append, dim3(grid_for(n)), dim3(256), 0, st, g->a, src, n, g->d_table,
                           g->cap - 1, g->d_ngroups);
    PGPU_LAUNCH_CHECK();
    return PGPU_OK;
}

int agg_fold(pgpu_agg* g, const void* rows, uint64_t n, bool upload) {
    hipStream_t st = g->st;
    const char* src = (const char*)rows;
    if (upload) {
        PGPU_TRY(slot_reserve(g->ctx, &g->d_in, &g->in_cap, n * g->a.row_bytes));
        PGPU_TRY(stage_h2d(g->ctx, g->hin, st, g->d_in, (const char*)rows, n * g->a.row_bytes));
        src = g->d_in;
    } else {
        // rows a kernel on the context's stream may still be writing
        PGPU_HIP(hipEventRecord(g->ctx_done, g->ctx->stream));
        PGPU_HIP(hipStreamWaitEvent(st, g->ctx_done, 0));
    }
    if (g->a.nkeys == 0) {  // one group: the slot exists after the first fold
        PGPU_TRY(agg_fold_piece(g, src, n, 1));
        PGPU_HIP(hipStreamSynchronize(st));
        return PGPU_OK;
    }
    for (uint64_t done = 0; done < n;) {
        PGPU_HIP(hipMemcpyAsync(g->h_ngroups, g->d_ngroups, 8, hipMemcpyDeviceToHost, st));
        PGPU_HIP(hipStreamSynchronize(st));
        const uint64_t ng = *g->h_ngroups;
        const uint64_t want = std::min(n - done, kAggMinPiece);
        uint64_t nc = g->cap;
        while (ng > nc / 4 || nc / 2 < want) nc *= 2;
        if (nc != g->cap) PGPU_TRY(agg_grow(g, nc));
        const uint64_t piece = std::min(n - done, g->cap / 2);
        PGPU_TRY(agg_fold_piece(g, src + done * g->a.row_bytes, piece, ng));
        done += piece;
    }
    PGPU_HIP(hipStreamSynchronize(st));
    return PGPU_OK;
}
}  // namespace

extern "C" {

int pgpu_agg_append(pgpu_agg* g, const void* rows, uint64_t n) {
    if (!g || (n && !rows)) return fail(PGPU_EINVAL, "pgpu_agg_append: null argument");
    std::lock_guard<std::mutex> lk(g->mu);
    if (g->finished) return fail(PGPU_ESTATE, "pgpu_agg_append after finish");
    if (!n) return PGPU_OK;
    PGPU_TRY(ctx_activate(g->ctx));
    return agg_fold(g, rows, n, true);
}

int pgpu_agg_append_device(pgpu_agg* g, const void* d_rows, uint64_t n) {
    if (!g || (n && !d_rows)) return fail(PGPU_EINVAL, "pgpu_agg_append_device: null argument");
    std::lock_guard<std::mutex> lk(g->mu);
    if (g->finished) return fail(PGPU_ESTATE, "pgpu_agg_append_device after finish");
    if (!n) return PGPU_OK;
    PGPU_TRY(ctx_activate(g->ctx));
    return agg_fold(g, d_rows, n, false);
}

int pgpu_agg_finish(pgpu_agg* g, uint64_t* n_groups, uint32_t* out_row_bytes) {
    if (!g || !n_groups) return fail(PGPU_EINVAL, "pgpu_agg_finish: null argument");
    std::lock_guard<std::mutex> lk(g->mu);
    PGPU_TRY(ctx_activate(g->ctx));
    hipStream_t st = g->st;
    if (!g->finished) {
        PGPU_HIP(hipMemcpyAsync(g->h_ngroups, g->d_ngroups, 8, hipMemcpyDeviceToHost, st));
        PGPU_HIP(hipStreamSynchronize(st));
        const uint64_t ng = *g->h_ngroups;
        PGPU_TRY(slot_reserve(g->ctx, &g->d_out, &g->out_cap, ng * g->out_row_bytes));
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(256), 0, st, g->d_ngroups, 1, 0ull);
        PGPU_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_agg_compact, dim3(grid_for(g->cap)), dim3(256), 0, st, g->a, g->d_table, g->cap,
                           g->d_out, g->out_row_bytes, g->acc_bytes, g->d_ngroups);
        PGPU_LAUNCH_CHECK();
        PGPU_HIP(hipMemcpyAsync(g->h_ngroups, g->d_ngroups, 8, hipMemcpyDeviceToHost, st));
        PGPU_HIP(hipStreamSynchronize(st));
        if (*g->h_ngroups != ng) return fail(PGPU_EOVERFLOW, "group count changed during compaction");
        g->ngroups_out = ng;
        g->finished = true;
    }
    *n_groups = g->ngroups_out;
    if (out_row_bytes) *out_row_bytes = g->out_row_bytes;
    return PGPU_OK;
}

int pgpu_agg_pull(pgpu_agg* g, uint64_t first, uint64_t n, void* dst) {
    if (!g || (n && !dst)) return fail(PGPU_EINVAL, "pgpu_agg_pull: null argument");
    if (!g->finished) return fail(PGPU_ESTATE, "pgpu_agg_pull before finish");
    if (first + n > g->ngroups_out) return fail(PGPU_EINVAL, "pgpu_agg_pull: range beyond the groups");
    if (!n) return PGPU_OK;
    PGPU_TRY(ctx_activate(g->ctx));
    PGPU_HIP(hipMemcpyAsync(dst, g->d_out + first * g->out_row_bytes, n * g->out_row_bytes, hipMemcpyDeviceToHost,
                            g->st));
    PGPU_HIP(hipStreamSynchronize(g->st));
    return PGPU_OK;
}

}  // extern "C"

// ====================================================================== columns
namespace {
// One 4- or 8-byte field of every packed row into a dense column.
__global__ __launch_bounds__(256) void k_column(const char* __restrict__ rows, uint64_t n, uint32_t row_bytes,
                                                uint32_t off, uint32_t bytes, char* __restrict__ out) {
    const bool a4 = aligned4((uintptr_t)rows | row_bytes | off);
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
        const char* p = rows + i * row_bytes + off;
        if (bytes == 4)
            ((uint32_t*)out)[i] = ld_u32(p, a4);
        else
            ((uint64_t*)out)[i] = ld_u64(p, a4);
    }
}
}  // namespace

extern "C" {

int pgpu_rows_column(pgpu_ctx* ctx, const void* d_rows, uint64_t n, uint32_t row_bytes, uint32_t offset,
                     uint32_t bytes, void* d_out) {
    if (!ctx || (n && (!d_rows || !d_out))) return fail(PGPU_EINVAL, "pgpu_rows_column: null argument");
    if (bytes != 4 && bytes != 8) return fail(PGPU_EUNSUPPORTED, "pgpu_rows_column: fields of 4 or 8 bytes");
    if ((uint64_t)offset + bytes > row_bytes) return fail(PGPU_EINVAL, "pgpu_rows_column: field outside the row");
    if ((uintptr_t)d_out % bytes) return fail(PGPU_EINVAL, "pgpu_rows_column: output not aligned to the field");
    if (!n) return PGPU_OK;
    PGPU_TRY(ctx_activate(ctx));
    hipLaunchKernelGGL(k_column, dim3(grid_for(n)), dim3(256), 0, ctx->stream, (const char*)d_rows, n, row_bytes,
                       offset, bytes, (char*)d_out);
    PGPU_LAUNCH_CHECK();
    PGPU_HIP(hipStreamSynchronize(ctx->stream));
    return PGPU_OK;
}

}  // extern "C"
