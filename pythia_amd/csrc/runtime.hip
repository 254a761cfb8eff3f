// Context, budgeted HBM allocation, copies, errors and kernel timing.
//
// The reference allocates every operator buffer through
// numaallocate_onnode/numaallocate_local with per-tag byte accounting
// (util/numaallocate.cpp:60-181, :402-457).  Here device memory is the
// resource: every allocation is charged against the context's HBM budget,
// and an allocation that would exceed it fails with PGPU_EBUDGET instead of
// blocking, so admission control can act on it.
#include <algorithm>
#include <mutex>
#include <map>
#include <cstdio>
#include <cstring>

#include "pgpu_internal.h"

namespace {
thread_local std::string g_last_error;
}

namespace pgpu {

int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int ctx_activate(pgpu_ctx* ctx) {
    PGPU_HIP(hipSetDevice(ctx->device));
    return PGPU_OK;
}

namespace {
uint64_t round_bytes(uint64_t bytes) { return bytes ? (bytes + 255) & ~uint64_t(255) : 256; }

// Frees every cached block (ctx->mu held).
void release_cache(pgpu_ctx* ctx) {
    if (ctx->cache.empty()) return;
    (void)hipSetDevice(ctx->device);
    for (auto& kv : ctx->cache) (void)hipFree(kv.second);
    ctx->cache.clear();
    ctx->cached = 0;
}
}  // namespace

int ctx_alloc(pgpu_ctx* ctx, uint64_t bytes, void** p) {
    if (!ctx || !p) return fail(PGPU_EINVAL, "ctx_alloc: null argument");
    *p = nullptr;
    bytes = round_bytes(bytes);
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        if (ctx->used + ctx->cached + bytes > ctx->budget) release_cache(ctx);
        if (ctx->used + bytes > ctx->budget)
            return fail(PGPU_EBUDGET,
                        "HBM budget exceeded: %llu B in use + %llu B requested > %llu B budget",
                        (unsigned long long)ctx->used, (unsigned long long)bytes,
                        (unsigned long long)ctx->budget);
        ctx->used += bytes;  // reserve before the (slow) allocation
        if (ctx->used > ctx->peak) ctx->peak = ctx->used;
    }
    PGPU_TRY(ctx_activate(ctx));
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipErrorOutOfMemory) {  // memory taken outside the library: drop the kept blocks, try again
        (void)hipGetLastError();
        {
            std::lock_guard<std::mutex> g(ctx->mu);
            release_cache(ctx);
        }
        e = hipMalloc(p, bytes);
    }
    if (e != hipSuccess) {
        std::lock_guard<std::mutex> g(ctx->mu);
        ctx->used -= bytes;
        *p = nullptr;
        (void)hipGetLastError();
        return fail(PGPU_ENOMEM, "hipMalloc(%llu) failed: %s", (unsigned long long)bytes,
                    hipGetErrorString(e));
    }
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->allocs[*p] = bytes;
    return PGPU_OK;
}

int ctx_free(pgpu_ctx* ctx, void* p) {
    if (!p) return PGPU_OK;
    uint64_t bytes = 0;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        auto it = ctx->allocs.find(p);
        if (it == ctx->allocs.end())
            return fail(PGPU_EINVAL, "pgpu_free: %p was not allocated by this context", p);
        bytes = it->second;
        ctx->allocs.erase(it);
        ctx->used -= bytes;
    }
    PGPU_TRY(ctx_activate(ctx));
    PGPU_HIP(hipFree(p));
    return PGPU_OK;
}

int ctx_alloc_reuse(pgpu_ctx* ctx, uint64_t bytes, void** p) {
    if (!ctx || !p) return fail(PGPU_EINVAL, "ctx_alloc_reuse: null argument");
    bytes = round_bytes(bytes);
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        auto it = ctx->cache.find(bytes);
        if (it != ctx->cache.end() && ctx->used + bytes <= ctx->budget) {
            *p = it->second;
            ctx->cache.erase(it);
            ctx->cached -= bytes;
            ctx->used += bytes;
            if (ctx->used > ctx->peak) ctx->peak = ctx->used;
            ctx->allocs[*p] = bytes;
            return PGPU_OK;
        }
    }
    return ctx_alloc(ctx, bytes, p);
}

int ctx_free_reuse(pgpu_ctx* ctx, void* p) {
    if (!p) return PGPU_OK;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        auto it = ctx->allocs.find(p);
        if (it == ctx->allocs.end()) return fail(PGPU_EINVAL, "ctx_free_reuse: %p was not allocated by this context", p);
        const uint64_t bytes = it->second;
        // keep at most a quarter of the budget cached
        if (ctx->cached + bytes <= ctx->budget / 4) {
            ctx->allocs.erase(it);
            ctx->used -= bytes;
            ctx->cache.emplace(bytes, p);
            ctx->cached += bytes;
            return PGPU_OK;
        }
    }
    return ctx_free(ctx, p);
}

int ctx_stream_get(pgpu_ctx* ctx, hipStream_t* st) {
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        if (!ctx->stream_pool.empty()) {
            *st = ctx->stream_pool.back();
            ctx->stream_pool.pop_back();
            return PGPU_OK;
        }
    }
    PGPU_TRY(ctx_activate(ctx));
    PGPU_HIP(hipStreamCreateWithFlags(st, hipStreamNonBlocking));
    return PGPU_OK;
}

void ctx_stream_put(pgpu_ctx* ctx, hipStream_t st) {
    if (!st) return;
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->stream_pool.push_back(st);
}

int ctx_host_alloc_reuse(pgpu_ctx* ctx, uint64_t bytes, void** p, uint64_t* cap) {
    const uint64_t b = std::max<uint64_t>(1, (bytes + (1ull << 20) - 1) >> 20) << 20;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        auto it = ctx->host_cache.find(b);
        if (it != ctx->host_cache.end()) {
            *p = it->second;
            ctx->host_cache.erase(it);
            *cap = b;
            return PGPU_OK;
        }
    }
    PGPU_TRY(ctx_activate(ctx));
    PGPU_HIP(hipHostMalloc(p, b, hipHostMallocDefault));
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->host_blocks[*p] = b;
    *cap = b;
    return PGPU_OK;
}

void ctx_host_free_reuse(pgpu_ctx* ctx, void* p) {
    if (!p) return;
    constexpr uint64_t kHostCacheMax = 1ull << 30;  // pinned staging kept for reuse, at most
    std::lock_guard<std::mutex> g(ctx->mu);
    auto it = ctx->host_blocks.find(p);
    if (it == ctx->host_blocks.end()) return;
    uint64_t kept = 0;
    for (auto& kv : ctx->host_cache) kept += kv.first;
    if (kept + it->second <= kHostCacheMax) {
        ctx->host_cache.emplace(it->second, p);
        return;
    }
    (void)hipHostFree(p);
    ctx->host_blocks.erase(it);
}

int ctx_host_word(pgpu_ctx* ctx, unsigned long long** w) {
    std::lock_guard<std::mutex> g(ctx->mu);
    if (ctx->host_words.empty()) {
        constexpr int kWords = 512;
        void* chunk = nullptr;
        PGPU_HIP(hipHostMalloc(&chunk, kWords * 8, hipHostMallocDefault));
        ctx->host_chunks.push_back(chunk);
        for (int i = kWords - 1; i >= 0; --i) ctx->host_words.push_back((unsigned long long*)chunk + i);
    }
    *w = ctx->host_words.back();
    ctx->host_words.pop_back();
    return PGPU_OK;
}

void ctx_host_word_put(pgpu_ctx* ctx, unsigned long long* w) {
    if (!w) return;
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->host_words.push_back(w);
}

// ---------------------------------------------------------------- timer
KernelTimer::~KernelTimer() {
    for (auto& p : pending_) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : pool_) (void)hipEventDestroy(e);
}

hipEvent_t KernelTimer::get_event() {
    if (!pool_.empty()) {
        hipEvent_t e = pool_.back();
        pool_.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void KernelTimer::enable(bool on) {
    harvest();
    on_ = on;
    if (on) totals_.clear();
}

int KernelTimer::start(const char* name) {
    if (!on_) return -1;
    int slot = -1;
    for (size_t i = 0; i < totals_.size(); ++i)
        if (totals_[i].name == name) slot = (int)i;
    if (slot < 0) {
        totals_.push_back(Total{name, 0, 0.0});
        slot = (int)totals_.size() - 1;
    }
    Pending p{slot, get_event(), get_event()};
    if (!p.a || !p.b) return -1;
    (void)hipEventRecord(p.a, stream_);
    pending_.push_back(p);
    return (int)pending_.size() - 1;
}

void KernelTimer::stop(int idx) {
    if (idx < 0 || idx >= (int)pending_.size()) return;
    (void)hipEventRecord(pending_[idx].b, stream_);
}

int KernelTimer::harvest() {
    for (auto& p : pending_) {
        float ms = 0.f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            totals_[p.slot].launches += 1;
            totals_[p.slot].ms += ms;
        }
        pool_.push_back(p.a);
        pool_.push_back(p.b);
    }
    pending_.clear();
    return PGPU_OK;
}

}  // namespace pgpu

using namespace pgpu;

namespace {
// Blocks from pgpu_host_alloc: start -> bytes.
std::mutex g_host_allocs_mu;
std::map<uintptr_t, uint64_t> g_host_allocs;
}  // namespace

bool pgpu::host_alloc_holds(const void* p, uint64_t bytes) {
    const uintptr_t a = (uintptr_t)p;
    std::lock_guard<std::mutex> lk(g_host_allocs_mu);
    auto it = g_host_allocs.upper_bound(a);
    if (it == g_host_allocs.begin()) return false;
    --it;
    return a >= it->first && a - it->first <= it->second && bytes <= it->second - (a - it->first);
}

extern "C" {

const char* pgpu_version(void) { return "pgpu 0.1 gfx950"; }

const char* pgpu_last_error(void) { return g_last_error.c_str(); }

int pgpu_device_count(int* n) {
    if (!n) return fail(PGPU_EINVAL, "pgpu_device_count: null argument");
    *n = 0;
    hipError_t e = hipGetDeviceCount(n);
    if (e != hipSuccess) {
        *n = 0;
        (void)hipGetLastError();
        return fail(PGPU_EDEVICE, "hipGetDeviceCount: %s", hipGetErrorString(e));
    }
    return PGPU_OK;
}

int pgpu_ctx_create(int device, uint64_t hbm_budget_bytes, pgpu_ctx** out) {
    if (!out) return fail(PGPU_EINVAL, "pgpu_ctx_create: null out");
    *out = nullptr;
    int n = 0;
    PGPU_TRY(pgpu_device_count(&n));
    if (device < 0 || device >= n)
        return fail(PGPU_EINVAL, "pgpu_ctx_create: device %d out of range (%d devices)", device, n);
    PGPU_HIP(hipSetDevice(device));
    size_t freeb = 0, totalb = 0;
    PGPU_HIP(hipMemGetInfo(&freeb, &totalb));
    auto* ctx = new pgpu_ctx();
    ctx->device = device;
    ctx->budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)freeb;
    // Default priority stream; never blocks on the legacy null stream.
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return fail(PGPU_EDEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    *out = ctx;
    return PGPU_OK;
}

int pgpu_ctx_destroy(pgpu_ctx* ctx) {
    if (!ctx) return PGPU_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->allocs) (void)hipFree(kv.first);
    ctx->allocs.clear();
    release_cache(ctx);
    for (auto st : ctx->stream_pool) (void)hipStreamDestroy(st);
    for (auto c : ctx->host_chunks) (void)hipHostFree(c);
    for (auto& kv : ctx->host_blocks) (void)hipHostFree(kv.first);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PGPU_OK;
}

int pgpu_ctx_stream(pgpu_ctx* ctx, void** s) {
    if (!ctx || !s) return fail(PGPU_EINVAL, "pgpu_ctx_stream: null argument");
    *s = (void*)ctx->stream;
    return PGPU_OK;
}

int pgpu_alloc(pgpu_ctx* ctx, uint64_t bytes, void** d_ptr) { return ctx_alloc(ctx, bytes, d_ptr); }

int pgpu_free(pgpu_ctx* ctx, void* d_ptr) {
    if (!ctx) return fail(PGPU_EINVAL, "pgpu_free: null ctx");
    return ctx_free(ctx, d_ptr);
}

int pgpu_hbm_usage(pgpu_ctx* ctx, uint64_t* used, uint64_t* budget, uint64_t* peak) {
    if (!ctx) return fail(PGPU_EINVAL, "pgpu_hbm_usage: null ctx");
    std::lock_guard<std::mutex> g(ctx->mu);
    if (used) *used = ctx->used;
    if (budget) *budget = ctx->budget;
    if (peak) *peak = ctx->peak;
    return PGPU_OK;
}

int pgpu_h2d(pgpu_ctx* ctx, void* d_dst, const void* src, uint64_t bytes) {
    if (!ctx || (!d_dst && bytes) || (!src && bytes)) return fail(PGPU_EINVAL, "pgpu_h2d: null argument");
    if (!bytes) return PGPU_OK;
    PGPU_TRY(ctx_activate(ctx));
    PGPU_HIP(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    PGPU_HIP(hipStreamSynchronize(ctx->stream));
    return PGPU_OK;
}

int pgpu_d2h(pgpu_ctx* ctx, void* dst, const void* d_src, uint64_t bytes) {
    if (!ctx || (!dst && bytes) || (!d_src && bytes)) return fail(PGPU_EINVAL, "pgpu_d2h: null argument");
    if (!bytes) return PGPU_OK;
    PGPU_TRY(ctx_activate(ctx));
    PGPU_HIP(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    PGPU_HIP(hipStreamSynchronize(ctx->stream));
    return PGPU_OK;
}

int pgpu_host_alloc(pgpu_ctx* ctx, uint64_t bytes, void** p) {
    if (!ctx || !p) return fail(PGPU_EINVAL, "pgpu_host_alloc: null argument");
    *p = nullptr;
    PGPU_TRY(ctx_activate(ctx));
    bytes = std::max<uint64_t>(bytes, 1);
    PGPU_HIP(hipHostMalloc(p, bytes, hipHostMallocDefault));
    std::lock_guard<std::mutex> lk(g_host_allocs_mu);
    g_host_allocs[(uintptr_t)*p] = bytes;
    return PGPU_OK;
}

int pgpu_host_free(pgpu_ctx* ctx, void* p) {
    if (!ctx) return fail(PGPU_EINVAL, "pgpu_host_free: null ctx");
    if (!p) return PGPU_OK;
    PGPU_TRY(ctx_activate(ctx));
    {
        std::lock_guard<std::mutex> lk(g_host_allocs_mu);
        g_host_allocs.erase((uintptr_t)p);
    }
    PGPU_HIP(hipHostFree(p));
    return PGPU_OK;
}

int pgpu_sync(pgpu_ctx* ctx) {
    if (!ctx) return fail(PGPU_EINVAL, "pgpu_sync: null ctx");
    PGPU_TRY(ctx_activate(ctx));
    PGPU_HIP(hipStreamSynchronize(ctx->stream));
    return PGPU_OK;
}

}  // extern "C"
