// Internal declarations shared by the libpgpu translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/pgpu.h"

namespace pgpu {

// Records a message as the calling thread's last error and returns `code`.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

#define PGPU_HIP(expr)                                                                   \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return ::pgpu::fail(PGPU_EDEVICE, "%s failed: %s (%s:%d)", #expr,            \
                                hipGetErrorString(e_), __FILE__, __LINE__);              \
    } while (0)

#define PGPU_TRY(expr)                                                                   \
    do {                                                                                 \
        int rc_ = (expr);                                                                \
        if (rc_ != PGPU_OK)                                                              \
            return rc_;                                                                  \
    } while (0)

#define PGPU_LAUNCH_CHECK()                                                              \
    do {                                                                                 \
        hipError_t e_ = hipGetLastError();                                               \
        if (e_ != hipSuccess)                                                            \
            return ::pgpu::fail(PGPU_EDEVICE, "kernel launch failed: %s (%s:%d)",        \
                                hipGetErrorString(e_), __FILE__, __LINE__);              \
    } while (0)

inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
inline uint32_t ilog2_ceil(uint64_t n) {
    uint32_t k = 0;
    while (k < 63 && (1ULL << k) < n) ++k;
    return k;
}

// Per-kernel event timing on one stream.  Event pairs are recorded around
// launches while enabled and folded into totals when harvested.
class KernelTimer {
  public:
    explicit KernelTimer(hipStream_t s) : stream_(s) {}
    ~KernelTimer();
    void enable(bool on);
    bool enabled() const { return on_; }
    // Returns an index to pass to stop().  -1 when disabled.
    int start(const char* name);
    void stop(int idx);
    int harvest();  // synchronises pending events into totals
    struct Total {
        std::string name;
        uint64_t launches = 0;
        double ms = 0;
    };
    const std::vector<Total>& totals() const { return totals_; }

  private:
    struct Pending {
        int slot;
        hipEvent_t a, b;
    };
    hipStream_t stream_;
    bool on_ = false;
    std::vector<Total> totals_;
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> pool_;
    hipEvent_t get_event();
};

}  // namespace pgpu

struct pgpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;
    uint64_t used = 0;  // live allocations (cached blocks excluded)
    uint64_t peak = 0;
    std::mutex mu;
    std::unordered_map<void*, uint64_t> allocs;
    // Reuse pools for the resources the operators of one query create and
    // drop (a query's operators live for that query only): streams, pinned
    // host words and device blocks, handed out again by exact size.
    std::vector<hipStream_t> stream_pool;
    std::vector<unsigned long long*> host_words;
    std::vector<void*> host_chunks;
    std::multimap<uint64_t, void*> cache;  // bytes -> block
    uint64_t cached = 0;                   // bytes held in `cache` (count against the budget)
    std::multimap<uint64_t, void*> host_cache;  // pinned host staging blocks: bytes -> block
    std::unordered_map<void*, uint64_t> host_blocks;  // every pinned staging block and its size
};

namespace pgpu {
// Budgeted device allocation on behalf of library objects.
int ctx_alloc(pgpu_ctx* ctx, uint64_t bytes, void** p);
int ctx_free(pgpu_ctx* ctx, void* p);
// Makes ctx->device current on the calling thread.
int ctx_activate(pgpu_ctx* ctx);
// As ctx_alloc, but first takes a cached block of the same rounded size.
int ctx_alloc_reuse(pgpu_ctx* ctx, uint64_t bytes, void** p);
// Keeps a block from ctx_alloc / ctx_alloc_reuse for reuse instead of
// freeing it (hipFree would wait for the whole device).  Every stream that
// used the block must have finished with it.
int ctx_free_reuse(pgpu_ctx* ctx, void* p);
// A pooled non-blocking stream at the default priority; put back only
// after it has been synchronized.
int ctx_stream_get(pgpu_ctx* ctx, hipStream_t* st);
void ctx_stream_put(pgpu_ctx* ctx, hipStream_t st);
// Pinned host staging of at least `bytes` (sizes rounded up to 1 MiB and
// reused by size); host_free_reuse keeps it for the next caller.
int ctx_host_alloc_reuse(pgpu_ctx* ctx, uint64_t bytes, void** p, uint64_t* cap);
void ctx_host_free_reuse(pgpu_ctx* ctx, void* p);
// One pinned host word (8 bytes) from the context's pool.
int ctx_host_word(pgpu_ctx* ctx, unsigned long long** w);
void ctx_host_word_put(pgpu_ctx* ctx, unsigned long long* w);
// True when [p, p + bytes) lies in one block from pgpu_host_alloc (pinned
// host memory that kernels may write directly).
bool host_alloc_holds(const void* p, uint64_t bytes);

// A build side radix-partitioned once (int64 keys and payloads, 2-pass split
// of `bits` bits, as pgpu_join_count_sum partitions it) and joined by many
// probes, all on the caller's stream.  p: a radix plan with 8-byte keys (no
// probe payload) whose build / probe capacities hold n / m rows; a plan's
// build-side and probe-side scratch are separate, so one build and one probe
// may use p at the same time, but not two of either.
//   radix_build_resident  rows (n 16-byte rows), base (2^bits + 1 words),
//                         tot (2^bits words) receive the partitioned side
//   radix_probe_resident  count and build-payload sum of the join of m probe
//                         keys with it into res[0], res[1] (res[2] = 0,
//                         res[3] = LDS passes)
int radix_build_resident(pgpu_join_plan* p, hipStream_t st, const int64_t* key, const int64_t* pay, uint64_t n,
                         int bits, void* rows, uint64_t* base, uint32_t* tot);
int radix_probe_resident(pgpu_join_plan* p, hipStream_t st, const int64_t* probe, uint64_t m, int bits,
                         const void* rows, const uint64_t* base, const uint32_t* tot, unsigned long long* res);
}  // namespace pgpu
