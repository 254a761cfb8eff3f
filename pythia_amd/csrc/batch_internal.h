// Segmented radix partitioning and the persistent work-queue join shared by
// the batched plans (many independent joins in one launch sequence) and the
// serving engine (probe batches of many queries against resident tables).
#pragma once

#include <cstdint>
#include <vector>

#include "pgpu_internal.h"

namespace pgpu {
namespace batch {

constexpr uint32_t kChunkRows = 131072;  // default rows per histogram / scatter block
constexpr int kMaxBits = 12;             // partition bits of one segment (one pass)
constexpr int kScatterTileBytes = 65536; // LDS staging per scatter tile

// One relation side of one join: input columns, where its partitioned rows
// go, and where its histogram and partition bounds live.
struct Seg {
    const void* key;
    const void* pay;   // may be null: payload 0
    uint64_t n;        // rows (< 2^32)
    uint64_t dst;      // first row in the side's partition buffer
    uint64_t hist;     // first histogram word (nchunks x 2^bits, chunk-major)
    uint64_t bnd;      // first bound word (2^bits + 1 entries)
    uint32_t chunk0;   // first chunk (global index)
    uint32_t nchunks;
    uint32_t bits;
    uint32_t chunk_rows;  // rows per chunk (a multiple of 16)
};

struct Chunk {
    uint32_t seg;
    uint32_t pad;
    uint64_t row0;
};

// One join: partition d of job j is work item item0 + d.
struct Job {
    uint64_t rdst, sdst;  // segment bases in the build / probe partition buffers
    uint64_t rbnd, sbnd;  // first bound word of each side
    uint32_t item0;
    uint32_t bits;
    uint32_t split;       // work items per partition (0 or 1: one); item item0 + d * split + k
                          // probes the k-th of `split` equal slices of partition d's probe rows
    uint32_t pad;
};

// Host-side builder of a side's segment and chunk tables.
struct Layout {
    std::vector<Seg> segs;
    std::vector<Chunk> chunks;
    uint64_t rows = 0, hist = 0, bnd = 0;
    // chunk_base: global index of this layout's first chunk.
    void add(const void* key, const void* pay, uint64_t n, uint32_t bits, uint32_t chunk_base,
             uint32_t chunk_rows = kChunkRows);
};

// Rows per partition the join kernel variant holds in LDS in one piece.
int variant_piece(int variant);
// Variants 1 and 2 take any keys; variant 3 (unique-key 8-byte slots, with
// variant 2 for the partitions it hands back) takes 4-byte keys.
bool variant_ok(int variant, int key_bytes);
// Default target build rows per partition for a variant.
uint32_t variant_target(int variant);
// Partition bits for n build rows at `target` rows per partition.
uint32_t bits_for(uint64_t n, uint32_t target);

int launch_hist(hipStream_t st, int key_bytes, const Seg* segs, const Chunk* chunks, uint32_t chunk_base,
                uint32_t nchunks, uint32_t* H);
// max_bits: the largest partition bits of the launched segments.
int launch_scan(hipStream_t st, uint32_t max_bits, const Seg* segs, uint32_t seg_base, uint32_t nsegs, uint32_t* H,
                uint32_t* bnd);
// Rows of one scatter tile for rows of key_bytes (x 2 with a payload) when
// the launch's largest partition bits are max_bits: segments whose chunks
// are one tile long get exact offsets for every tile.
uint32_t scatter_tile_rows(int key_bytes, bool hv, uint32_t max_bits);
// max_bits: the largest partition bits of the launched chunks' segments.
int launch_scatter(hipStream_t st, int key_bytes, bool hv, uint32_t max_bits, const Seg* segs, const Chunk* chunks,
                   uint32_t chunk_base, uint32_t nchunks, const uint32_t* H, void* out);
// Persistent join over n_items work items; `grid` from join_grid().  queue:
// zeroed words (one; three for variant 3); redo: n_items words of scratch
// (variant 3 only: the partitions handed to its second kernel).
int launch_join(hipStream_t st, int key_bytes, bool hvs, int variant, int grid, const Job* jobs,
                const uint32_t* item_job, uint32_t n_items, const void* R, const void* S, const uint32_t* bndR,
                const uint32_t* bndS, uint32_t* queue, unsigned long long* res, uint32_t* redo = nullptr);
int join_grid(int key_bytes, bool hvs, int variant, int* grid);
// Zeroes n 64-bit words with a kernel (usable inside graph capture).
int launch_zero(hipStream_t st, unsigned long long* p, uint64_t n);

}  // namespace batch
}  // namespace pgpu
