// Join + sum serving engine: build tables resident in HBM across queries,
// admission by an HBM budget, continuous batching onto overlapping streams.
//
// Reference path replaced: executing the queryhashjoin plan
// (drivers/executequery.cpp:38-59 over HashJoinOp -> AggregateSum,
// operators/join.cpp:396-562, operators/aggregatesum.cpp:81-107) once per
// query, which rebuilds the build side's HashTable (util/hashtable.cpp:
// 57-97, :230-273) every time.  The reference has no server; its closest
// analogue of a resident table is HashTable::serialize / deserialize
// (util/hashtable.cpp:99-201).
//
// Served mode: each registered table is hashed once into an open-addressing
// table in HBM (16-byte {key, folded payload sum} slots at load <= 0.5; a
// per-slot count array only when the table holds duplicate keys).  A
// scheduler thread owns a few in-flight slots, each with its own stream:
// whenever a slot is free it admits the queued queries that fit the HBM
// budget (FIFO), cuts their probe rows into work items and launches one
// probe kernel over all of them, so queries that arrive while others run
// join the next launch and overlap on the device.  Plain mode runs each
// query alone through the per-query join plan (pgpu_join_count_sum), which
// builds the table from its columns every time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <thread>

#include "join_common.h"
#include "pgpu_internal.h"

using namespace pgpu;
using namespace pgpu::dev;

namespace {

constexpr uint64_t kItemRows = 65536;  // probe rows per work item
constexpr int kProbeThreads = 256;
constexpr int kProbeU = 8;  // independent probes in flight per thread

struct __align__(16) SSlot {
    long long key;
    unsigned long long sum;
};

struct TDesc {
    SSlot* slots;
    uint32_t* cnt;                 // per-slot match counts (duplicates only)
    unsigned long long* special;   // [count, sum] of rows whose key is the empty marker
    int shift;                     // slot = hash >> shift
    uint32_t dups;
    uint64_t mask;
};

struct QItem {
    const long long* probe;
    uint64_t r0, r1;
    uint32_t q;      // query index within the batch
    uint32_t table;
};

__global__ __launch_bounds__(256) void k_tclear(SSlot* __restrict__ t, uint32_t* __restrict__ cnt, uint64_t cap,
                                                unsigned long long* __restrict__ special) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < cap; i += (uint64_t)gridDim.x * 256) {
        t[i].key = INT64_MIN;
        t[i].sum = 0;
        cnt[i] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x < 3) special[threadIdx.x] = 0;
}

// special[0..1]: count / sum of empty-marker keys; special[2]: inserts that
// found their key already present (duplicates).
__global__ __launch_bounds__(256) void k_tbuild(const long long* __restrict__ keys, const long long* __restrict__ pays,
                                                uint64_t n, SSlot* __restrict__ t, uint32_t* __restrict__ cnt,
                                                int shift, uint64_t mask, unsigned long long* __restrict__ special) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
        const long long k = keys[i];
        const unsigned long long v = (unsigned long long)pays[i];
        if (k == INT64_MIN) {
            atomicAdd(&special[0], 1ull);
            atomicAdd(&special[1], v);
            continue;
        }
        uint64_t s = hmix(k) >> shift;
        while (true) {
            const long long prev = (long long)atomicCAS((unsigned long long*)&t[s].key, (unsigned long long)INT64_MIN,
                                                        (unsigned long long)k);
            if (prev == INT64_MIN || prev == k) {
                if (prev == k) atomicAdd(&special[2], 1ull);
                atomicAdd(&cnt[s], 1u);
                atomicAdd(&t[s].sum, v);
                break;
            }
            s = (s + 1) & mask;
        }
    }
}

__global__ __launch_bounds__(kProbeThreads) void k_sprobe(const QItem* __restrict__ items,
                                                          const TDesc* __restrict__ tabs,
                                                          unsigned long long* __restrict__ res) {
    __shared__ unsigned long long red[2][kProbeThreads / 64];
    const QItem it = items[blockIdx.x];
    const TDesc td = tabs[it.table];
    unsigned long long acc_c = 0, acc_b = 0;
    constexpr int U = kProbeU;
    for (uint64_t r = it.r0 + threadIdx.x; r < it.r1; r += (uint64_t)U * kProbeThreads) {
        long long k[U];
        uint64_t s[U];
        SSlot t[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t row = r + (uint64_t)u * kProbeThreads;
            k[u] = row < it.r1 ? it.probe[row] : INT64_MIN;
            s[u] = hmix(k[u]) >> td.shift;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) t[u] = td.slots[s[u]];  // independent loads in flight
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t row = r + (uint64_t)u * kProbeThreads;
            if (row >= it.r1) continue;
            if (k[u] == INT64_MIN) {
                acc_c += td.special[0];
                acc_b += td.special[1];
                continue;
            }
            uint64_t slot = s[u];
            SSlot x = t[u];
            while (x.key != k[u] && x.key != INT64_MIN) {
                slot = (slot + 1) & td.mask;
                x = td.slots[slot];
            }
            if (x.key == k[u]) {
                acc_c += td.dups ? td.cnt[slot] : 1u;
                acc_b += x.sum;
            }
        }
    }
    acc_c = wave_sum(acc_c);
    acc_b = wave_sum(acc_b);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][w] = acc_c;
        red[1][w] = acc_b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long c = 0, b = 0;
        for (int i = 0; i < kProbeThreads / 64; ++i) {
            c += red[0][i];
            b += red[1][i];
        }
        if (c) {
            atomicAdd(&res[2 * (uint64_t)it.q], c);
            atomicAdd(&res[2 * (uint64_t)it.q + 1], b);
        }
    }
}

__global__ __launch_bounds__(256) void k_szero(unsigned long long* __restrict__ p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) p[i] = 0;
}

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Query {
    uint64_t tag = 0;
    uint32_t table = 0;
    const long long* probe = nullptr;
    uint64_t m = 0;
    uint64_t reserve = 0;
    double t_arrival = 0, t_admit = 0, t_done = 0;
    uint64_t matches = 0, sum = 0;
    uint32_t batch = 0, batch_size = 0;
};

struct Table {
    const void* key = nullptr;
    const void* pay = nullptr;
    uint64_t n = 0;
    SSlot* slots = nullptr;
    uint32_t* cnt = nullptr;
    unsigned long long* special = nullptr;
    uint64_t cap = 0;
    uint64_t bytes = 0;
};

struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    QItem* h_items = nullptr;           // pinned
    QItem* d_items = nullptr;
    unsigned long long* h_res = nullptr;  // pinned
    unsigned long long* d_res = nullptr;
    std::vector<Query*> qs;
    bool busy = false;
};

}  // namespace

struct pgpu_server {
    pgpu_ctx* ctx = nullptr;
    pgpu_server_cfg cfg{};
    std::vector<Table> tables;
    TDesc* d_tabs = nullptr;
    bool tabs_dirty = true;
    uint64_t resident = 0;
    uint64_t reserved = 0;  // HBM held by admitted, unfinished queries
    std::vector<Slot> slots;
    pgpu_join_plan* plain = nullptr;
    uint64_t plain_maxR = 0, plain_maxS = 0;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Query*> pending;
    std::deque<Query*> finished;
    uint64_t submitted = 0, completed = 0, batches = 0;
    std::thread th;
    bool stop = false;
    int error = PGPU_OK;
    std::string error_msg;
    double t0 = 0;
};

namespace {

uint32_t max_items_per_slot(const pgpu_server* s) { return std::max<uint32_t>(s->cfg.max_items, 1); }

void set_error(pgpu_server* s, int rc) {
    std::lock_guard<std::mutex> g(s->mu);
    if (s->error == PGPU_OK) {
        s->error = rc;
        s->error_msg = pgpu_last_error();
    }
}

// Completes the queries of a slot whose event has fired.
void reap(pgpu_server* s, Slot& sl) {
    const double t = now_s();
    std::lock_guard<std::mutex> g(s->mu);
    for (size_t i = 0; i < sl.qs.size(); ++i) {
        Query* q = sl.qs[i];
        q->matches = sl.h_res[2 * i];
        q->sum = sl.h_res[2 * i + 1];
        q->t_done = t;
        s->reserved -= q->reserve;
        s->finished.push_back(q);
        ++s->completed;
    }
    sl.qs.clear();
    sl.busy = false;
    s->cv.notify_all();
}

// Admits queued queries into a free slot and launches them.  Called with
// no lock held; returns true if it launched something.
bool launch_batch(pgpu_server* s, Slot& sl) {
    std::vector<Query*> qs;
    uint64_t items = 0;
    const double t = now_s();
    {
        std::lock_guard<std::mutex> g(s->mu);
        while (!s->pending.empty() && qs.size() < s->cfg.max_batch) {
            Query* q = s->pending.front();
            const uint64_t qi = std::max<uint64_t>(1, ceil_div(q->m, kItemRows));
            if (items + qi > max_items_per_slot(s)) break;
            if (s->resident + s->reserved + q->reserve > s->cfg.hbm_budget) break;  // FIFO: the head waits
            s->pending.pop_front();
            s->reserved += q->reserve;
            q->t_admit = t;
            qs.push_back(q);
            items += qi;
        }
    }
    if (qs.empty()) return false;
    ctx_activate(s->ctx);
    const uint32_t bid = (uint32_t)s->batches++;
    uint32_t ni = 0;
    for (uint32_t i = 0; i < qs.size(); ++i) {
        Query* q = qs[i];
        q->batch = bid;
        q->batch_size = (uint32_t)qs.size();
        for (uint64_t r = 0; r < q->m; r += kItemRows)
            sl.h_items[ni++] = QItem{q->probe, r, std::min(q->m, r + kItemRows), i, q->table};
    }
    int rc = PGPU_OK;
    const uint64_t nres = 2 * qs.size();
    hipLaunchKernelGGL(k_szero, dim3((unsigned)std::min<uint64_t>(ceil_div(nres, 256), 64)), dim3(256), 0, sl.st,
                       sl.d_res, nres);
    if (ni) {
        if (hipMemcpyAsync(sl.d_items, sl.h_items, ni * sizeof(QItem), hipMemcpyHostToDevice, sl.st) != hipSuccess)
            rc = fail(PGPU_EDEVICE, "item upload failed");
        hipLaunchKernelGGL(k_sprobe, dim3(ni), dim3(kProbeThreads), 0, sl.st, sl.d_items, s->d_tabs, sl.d_res);
    }
    if (rc == PGPU_OK && hipGetLastError() != hipSuccess) rc = fail(PGPU_EDEVICE, "probe launch failed");
    if (rc == PGPU_OK &&
        hipMemcpyAsync(sl.h_res, sl.d_res, nres * 8, hipMemcpyDeviceToHost, sl.st) != hipSuccess)
        rc = fail(PGPU_EDEVICE, "result copy failed");
    if (rc == PGPU_OK && hipEventRecord(sl.ev, sl.st) != hipSuccess) rc = fail(PGPU_EDEVICE, "event record failed");
    sl.qs = std::move(qs);
    sl.busy = true;
    if (rc != PGPU_OK) set_error(s, rc);
    return true;
}

// Plain mode: the oldest query alone through the per-query join plan.
bool run_plain(pgpu_server* s) {
    Query* q = nullptr;
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->pending.empty()) return false;
        q = s->pending.front();
        s->pending.pop_front();
        s->reserved += q->reserve;
        q->t_admit = now_s();
    }
    ctx_activate(s->ctx);
    const Table& tb = s->tables[q->table];
    pgpu_join_result r{};
    int rc = pgpu_join_count_sum(s->plain, tb.key, tb.pay, tb.n, q->probe, nullptr, q->m, &r);
    if (rc != PGPU_OK) set_error(s, rc);
    std::lock_guard<std::mutex> g(s->mu);
    q->matches = r.matches;
    q->sum = r.sum_build;
    q->t_done = now_s();
    q->batch = (uint32_t)s->batches++;
    q->batch_size = 1;
    s->reserved -= q->reserve;
    s->finished.push_back(q);
    ++s->completed;
    s->cv.notify_all();
    return true;
}

void scheduler(pgpu_server* s) {
    ctx_activate(s->ctx);
    while (true) {
        bool did = false;
        if (s->cfg.mode == PGPU_SERVE_PLAIN) {
            did = run_plain(s);
        } else {
            for (Slot& sl : s->slots) {
                if (sl.busy) {
                    const hipError_t e = hipEventQuery(sl.ev);
                    if (e == hipSuccess) {
                        reap(s, sl);
                        did = true;
                    } else if (e != hipErrorNotReady) {
                        set_error(s, fail(PGPU_EDEVICE, "batch failed: %s", hipGetErrorString(e)));
                        reap(s, sl);
                    }
                }
            }
            for (Slot& sl : s->slots)
                if (!sl.busy && launch_batch(s, sl)) did = true;
        }
        bool busy = false;
        for (Slot& sl : s->slots) busy |= sl.busy;
        std::unique_lock<std::mutex> lk(s->mu);
        if (s->stop && s->pending.empty() && !busy) break;
        if (!did) {
            if (busy)
                lk.unlock(), std::this_thread::yield();  // a batch is on the device: poll its event
            else
                s->cv.wait_for(lk, std::chrono::milliseconds(2), [&] { return s->stop || !s->pending.empty(); });
        }
    }
}

int upload_tabs(pgpu_server* s) {
    std::vector<TDesc> h(s->tables.size());
    for (size_t i = 0; i < s->tables.size(); ++i) {
        const Table& t = s->tables[i];
        const int logcap = (int)ilog2_ceil(t.cap);
        unsigned long long sp[3];
        PGPU_HIP(hipMemcpy(sp, t.special, sizeof sp, hipMemcpyDeviceToHost));
        h[i] = TDesc{t.slots, t.cnt, t.special, 64 - logcap, sp[2] ? 1u : 0u, t.cap - 1};
    }
    if (s->d_tabs) PGPU_TRY(ctx_free(s->ctx, s->d_tabs));
    s->d_tabs = nullptr;
    PGPU_TRY(ctx_alloc(s->ctx, std::max<size_t>(h.size(), 1) * sizeof(TDesc), (void**)&s->d_tabs));
    PGPU_HIP(hipMemcpy(s->d_tabs, h.data(), h.size() * sizeof(TDesc), hipMemcpyHostToDevice));
    s->tabs_dirty = false;
    return PGPU_OK;
}

}  // namespace

extern "C" {

int pgpu_server_create(pgpu_ctx* ctx, const pgpu_server_cfg* cfg, pgpu_server** out) {
    if (!ctx || !cfg || !out) return fail(PGPU_EINVAL, "pgpu_server_create: null argument");
    *out = nullptr;
    if (cfg->mode != PGPU_SERVE_BATCHED && cfg->mode != PGPU_SERVE_PLAIN)
        return fail(PGPU_EINVAL, "unknown serving mode %u", cfg->mode);
    PGPU_TRY(ctx_activate(ctx));
    auto* s = new pgpu_server();
    s->ctx = ctx;
    s->cfg = *cfg;
    if (!s->cfg.hbm_budget) s->cfg.hbm_budget = ctx->budget;
    if (!s->cfg.slots) s->cfg.slots = 3;
    if (!s->cfg.max_batch) s->cfg.max_batch = 256;
    if (!s->cfg.max_items) s->cfg.max_items = 32768;
    if (s->cfg.mode == PGPU_SERVE_PLAIN) s->cfg.slots = 0;
    s->t0 = now_s();
    auto cleanup = [&](int rc) {
        pgpu_server_destroy(s);
        return rc;
    };
    s->slots.resize(s->cfg.slots);
    for (Slot& sl : s->slots) {
        if (hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess)
            return cleanup(fail(PGPU_EDEVICE, "stream / event creation failed"));
        if (hipHostMalloc((void**)&sl.h_items, s->cfg.max_items * sizeof(QItem), hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void**)&sl.h_res, 2ull * s->cfg.max_batch * 8, hipHostMallocDefault) != hipSuccess)
            return cleanup(fail(PGPU_ENOMEM, "pinned staging allocation failed"));
        int rc = ctx_alloc(ctx, s->cfg.max_items * sizeof(QItem), (void**)&sl.d_items);
        if (rc == PGPU_OK) rc = ctx_alloc(ctx, 2ull * s->cfg.max_batch * 8, (void**)&sl.d_res);
        if (rc != PGPU_OK) return cleanup(rc);
    }
    s->th = std::thread(scheduler, s);
    *out = s;
    return PGPU_OK;
}

int pgpu_server_destroy(pgpu_server* s) {
    if (!s) return PGPU_OK;
    if (s->th.joinable()) {
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->stop = true;
            s->cv.notify_all();
        }
        s->th.join();
    }
    (void)hipSetDevice(s->ctx->device);
    for (Slot& sl : s->slots) {
        if (sl.st) (void)hipStreamSynchronize(sl.st);
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        if (sl.st) (void)hipStreamDestroy(sl.st);
        if (sl.h_items) (void)hipHostFree(sl.h_items);
        if (sl.h_res) (void)hipHostFree(sl.h_res);
        if (sl.d_items) (void)ctx_free(s->ctx, sl.d_items);
        if (sl.d_res) (void)ctx_free(s->ctx, sl.d_res);
    }
    for (Table& t : s->tables) {
        if (t.slots) (void)ctx_free(s->ctx, t.slots);
        if (t.cnt) (void)ctx_free(s->ctx, t.cnt);
        if (t.special) (void)ctx_free(s->ctx, t.special);
    }
    if (s->d_tabs) (void)ctx_free(s->ctx, s->d_tabs);
    if (s->plain) pgpu_join_plan_destroy(s->plain);
    for (Query* q : s->pending) delete q;
    for (Query* q : s->finished) delete q;
    delete s;
    return PGPU_OK;
}

int pgpu_server_add_table(pgpu_server* s, const void* d_key, const void* d_pay, uint64_t n, uint32_t* table_id) {
    if (!s || !table_id || (n && (!d_key || !d_pay))) return fail(PGPU_EINVAL, "pgpu_server_add_table: null argument");
    if (s->cfg.key_bytes != 8) return fail(PGPU_EUNSUPPORTED, "served tables have int64 keys");
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->submitted != s->completed || !s->pending.empty())
            return fail(PGPU_ESTATE, "tables are registered before queries are submitted");
    }
    PGPU_TRY(ctx_activate(s->ctx));
    Table t;
    t.key = d_key;
    t.pay = d_pay;
    t.n = n;
    if (s->cfg.mode == PGPU_SERVE_BATCHED) {
        t.cap = 1ull << std::max<uint32_t>(4, ilog2_ceil(2 * std::max<uint64_t>(n, 1)));
        t.bytes = t.cap * (sizeof(SSlot) + 4) + 32;
        if (s->resident + t.bytes > s->cfg.hbm_budget)
            return fail(PGPU_EBUDGET, "resident table of %llu rows exceeds the server's HBM budget",
                        (unsigned long long)n);
        PGPU_TRY(ctx_alloc(s->ctx, t.cap * sizeof(SSlot), (void**)&t.slots));
        int rc = ctx_alloc(s->ctx, t.cap * 4, (void**)&t.cnt);
        if (rc == PGPU_OK) rc = ctx_alloc(s->ctx, 32, (void**)&t.special);
        if (rc != PGPU_OK) {
            if (t.slots) ctx_free(s->ctx, t.slots);
            if (t.cnt) ctx_free(s->ctx, t.cnt);
            return rc;
        }
        hipStream_t st = s->ctx->stream;
        const int logcap = (int)ilog2_ceil(t.cap);
        hipLaunchKernelGGL(k_tclear, dim3((unsigned)std::min<uint64_t>(ceil_div(t.cap, 256), 4096)), dim3(256), 0, st,
                           t.slots, t.cnt, t.cap, t.special);
        if (n)
            hipLaunchKernelGGL(k_tbuild, dim3((unsigned)std::min<uint64_t>(ceil_div(n, 256), 4096)), dim3(256), 0, st,
                               (const long long*)d_key, (const long long*)d_pay, n, t.slots, t.cnt, 64 - logcap,
                               t.cap - 1, t.special);
        PGPU_LAUNCH_CHECK();
        PGPU_HIP(hipStreamSynchronize(st));
    } else {
        t.bytes = 0;  // plain mode keeps only the columns (rebuilt per query)
        s->plain_maxR = std::max(s->plain_maxR, n);
    }
    s->resident += t.bytes;
    s->tables.push_back(t);
    *table_id = (uint32_t)(s->tables.size() - 1);
    if (s->cfg.mode == PGPU_SERVE_BATCHED) PGPU_TRY(upload_tabs(s));
    return PGPU_OK;
}

int pgpu_server_prepare_plain(pgpu_server* s, uint64_t max_probe) {
    if (!s) return fail(PGPU_EINVAL, "pgpu_server_prepare_plain: null server");
    if (s->cfg.mode != PGPU_SERVE_PLAIN) return PGPU_OK;
    PGPU_TRY(ctx_activate(s->ctx));
    if (s->plain) pgpu_join_plan_destroy(s->plain);
    s->plain = nullptr;
    pgpu_join_cfg c{8, 0, 0, PGPU_ALGO_AUTO};
    s->plain_maxS = max_probe;
    return pgpu_join_plan_create(s->ctx, &c, std::max<uint64_t>(s->plain_maxR, 1), std::max<uint64_t>(max_probe, 1),
                                 &s->plain);
}

int pgpu_server_submit(pgpu_server* s, uint32_t table, const void* d_probe, uint64_t m, uint64_t tag,
                       double t_arrival) {
    if (!s || (m && !d_probe)) return fail(PGPU_EINVAL, "pgpu_server_submit: null argument");
    if (table >= s->tables.size()) return fail(PGPU_EINVAL, "unknown table %u", table);
    if (s->cfg.mode == PGPU_SERVE_PLAIN && (!s->plain || m > s->plain_maxS))
        return fail(PGPU_ESTATE, "plain mode needs pgpu_server_prepare_plain for probes of this size");
    if (s->cfg.mode == PGPU_SERVE_BATCHED && ceil_div(m, kItemRows) > max_items_per_slot(s))
        return fail(PGPU_EINVAL, "probe batch of %llu rows exceeds one launch's %u work items",
                    (unsigned long long)m, max_items_per_slot(s));
    auto* q = new Query();
    q->tag = tag;
    q->table = table;
    q->probe = (const long long*)d_probe;
    q->m = m;
    q->reserve = m * 8 + 64;  // the probe batch is held in HBM until its result is out
    q->t_arrival = t_arrival > 0 ? s->t0 + t_arrival : now_s();
    std::lock_guard<std::mutex> g(s->mu);
    if (s->resident + q->reserve > s->cfg.hbm_budget) {
        delete q;
        return fail(PGPU_EBUDGET, "a probe batch of %llu rows can never fit the server's HBM budget",
                    (unsigned long long)m);
    }
    if (s->error != PGPU_OK) {
        delete q;
        return fail(s->error, "server failed earlier: %s", s->error_msg.c_str());
    }
    s->pending.push_back(q);
    ++s->submitted;
    s->cv.notify_all();
    return PGPU_OK;
}

int pgpu_server_drain(pgpu_server* s) {
    if (!s) return fail(PGPU_EINVAL, "pgpu_server_drain: null server");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv.wait(lk, [&] { return s->completed == s->submitted || s->error != PGPU_OK; });
    if (s->error != PGPU_OK) return fail(s->error, "server failed: %s", s->error_msg.c_str());
    return PGPU_OK;
}

int pgpu_server_poll(pgpu_server* s, pgpu_served_result* out, int cap, int* n) {
    if (!s || !n || (cap > 0 && !out)) return fail(PGPU_EINVAL, "pgpu_server_poll: null argument");
    std::lock_guard<std::mutex> g(s->mu);
    int k = 0;
    while (k < cap && !s->finished.empty()) {
        Query* q = s->finished.front();
        s->finished.pop_front();
        out[k++] = pgpu_served_result{q->tag,          q->table, q->batch_size, q->m,
                                      q->matches,      q->sum,   q->t_arrival - s->t0,
                                      q->t_admit - s->t0, q->t_done - s->t0, q->batch};
        delete q;
    }
    *n = k;
    return PGPU_OK;
}

int pgpu_server_now(pgpu_server* s, double* t) {
    if (!s || !t) return fail(PGPU_EINVAL, "pgpu_server_now: null argument");
    *t = now_s() - s->t0;
    return PGPU_OK;
}

int pgpu_server_replay(pgpu_server* s, const uint32_t* tables, const void* const* probes, const uint64_t* ms,
                       const double* offsets, uint64_t n, pgpu_served_result* out) {
    if (!s || (n && (!tables || !probes || !ms || !offsets || !out)))
        return fail(PGPU_EINVAL, "pgpu_server_replay: null argument");
    const double start = now_s() + 1e-3;
    for (uint64_t i = 0; i < n; ++i) {
        const double due = start + offsets[i];
        while (now_s() < due) {
        }
        PGPU_TRY(pgpu_server_submit(s, tables[i], probes[i], ms[i], i, due - s->t0));
    }
    PGPU_TRY(pgpu_server_drain(s));
    uint64_t got = 0;
    std::vector<pgpu_served_result> tmp(1024);
    while (got < n) {
        int k = 0;
        PGPU_TRY(pgpu_server_poll(s, tmp.data(), (int)tmp.size(), &k));
        if (k == 0) return fail(PGPU_ESTATE, "pgpu_server_replay: results of other submissions are pending");
        for (int i = 0; i < k; ++i) {
            if (tmp[i].tag < n) out[tmp[i].tag] = tmp[i];
            ++got;
        }
    }
    return PGPU_OK;
}

int pgpu_server_stats_get(pgpu_server* s, pgpu_server_stats* st) {
    if (!s || !st) return fail(PGPU_EINVAL, "pgpu_server_stats_get: null argument");
    std::lock_guard<std::mutex> g(s->mu);
    memset(st, 0, sizeof *st);
    st->tables = s->tables.size();
    st->resident_bytes = s->resident;
    st->reserved_bytes = s->reserved;
    st->budget = s->cfg.hbm_budget;
    st->submitted = s->submitted;
    st->completed = s->completed;
    st->batches = s->batches;
    st->queued = s->pending.size();
    return PGPU_OK;
}

}  // extern "C"
