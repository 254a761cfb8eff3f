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
// Served mode keeps every registered table resident in two forms:
//   - an open-addressing hash table in HBM (16-byte {key, folded payload
//     sum} slots at load <= 0.5; a per-slot count array is read only when
//     the table holds duplicate keys), probed directly by small queries;
//   - its rows radix-partitioned into LDS-sized partitions (the batched
//     join's layout), joined by large queries after partitioning their
//     probe rows the same way: streaming instead of random HBM access.
// A scheduler thread owns a few in-flight slots, each with its own stream
// and staging.  Whenever a slot is free it admits the queued queries that
// fit the HBM budget (FIFO) and launches them together: one random-probe
// kernel over the small queries' row ranges, one segmented partition pass
// plus one persistent LDS join over the large ones, and a finish kernel
// that writes the results to pinned host memory.  Queries that arrive
// while others run join the next launch and overlap on the device.  Work
// descriptors are read by the kernels from pinned host memory, so no DMA
// copy sits on a query's path.
//
// Plain mode runs each query alone through the per-query join plan
// (pgpu_join_count_sum), which builds the table from its columns each time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <string>
#include <thread>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "batch_internal.h"
#include "join_common.h"

using namespace pgpu;
using namespace pgpu::dev;
using pgpu::batch::Chunk;
using pgpu::batch::Job;
using pgpu::batch::kChunkRows;
using pgpu::batch::Seg;

namespace {

constexpr uint64_t kItemRows = 16384;  // largest random-probe work item (rows)
constexpr uint64_t kMinItemRows = 1024;  // smallest one: half a pass of the probe loop
constexpr uint64_t kBatchItems = 512;    // work items a batch aims for (two per CU)
constexpr int kProbeThreads = 256;
constexpr int kProbeU = 8;             // independent probes in flight per thread
constexpr int kVariant = 2;            // LDS join shape of the partitioned path
constexpr uint32_t kPartTarget = 3584; // build rows per resident partition (one LDS piece
                                       // holds 4096, so a partition rarely needs two)
constexpr uint64_t kSliceRows = 32768; // probe rows per join work item, about (partitions
                                       // with more are split into slices)
constexpr uint32_t kMaxSplit = 64;

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
    uint32_t q;      // query index within the batch (result slot)
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
// Row i's key is keys[i * stride], its payload pays[i * stride] (stride 1:
// columns; 2: interleaved (key, payload) rows).
__global__ __launch_bounds__(256) void k_tbuild(const long long* __restrict__ keys, const long long* __restrict__ pays,
                                                uint64_t n, uint32_t stride, SSlot* __restrict__ t,
                                                uint32_t* __restrict__ cnt, int shift, uint64_t mask,
                                                unsigned long long* __restrict__ special) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
        const long long k = keys[i * stride];
        const unsigned long long v = (unsigned long long)pays[i * stride];
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
            k[u] = row < it.r1 ? nt_load(&it.probe[row]) : INT64_MIN;
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
            atomicAdd(&res[4 * (uint64_t)it.q], c);
            atomicAdd(&res[4 * (uint64_t)it.q + 1], b);
        }
    }
}

// Copies the batch's (count, sum) pairs to pinned host memory and leaves the
// device accumulators and the join queue at zero for the slot's next batch.
__global__ __launch_bounds__(256) void k_sfinish(unsigned long long* __restrict__ res, uint32_t nq,
                                                 unsigned long long* __restrict__ host, uint32_t* __restrict__ queue) {
    for (uint32_t i = threadIdx.x; i < nq; i += 256) {
        host[2 * i] = res[4 * i];
        host[2 * i + 1] = res[4 * i + 1];
        res[4 * i] = res[4 * i + 1] = res[4 * i + 2] = res[4 * i + 3] = 0;
    }
    if (threadIdx.x == 0) *queue = 0;
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
    uint32_t bits = 0;      // partition bits of the resident partitioned rows
    uint64_t rdst = 0;      // first row in the resident partitioned-row arena
    uint64_t rbnd = 0;      // first bound word in the resident bounds arena
    // The same rows in the 2-pass radix layout (radix_build_resident): a lone
    // large query probes them with the 2-pass radix partitioning and the LDS
    // join.  Absent for tables attached from segments or when the budget has
    // no room.
    int rbits = 0;
    void* rrows = nullptr;
    uint64_t* rbase = nullptr;
    uint32_t* rtot = nullptr;
    uint64_t rbytes = 0;
};

// Per-slot capacities of the partitioned path.
struct Caps {
    uint64_t rows = 1ull << 26;      // probe rows partitioned per launch
    uint64_t hist = 1ull << 25;      // histogram words
    uint64_t bnd = 1ull << 20;       // bound words
    uint64_t chunks = 16384;
    uint64_t pitems = 1ull << 20;    // join work items
};

// Pinned, device-readable staging of one slot (one allocation).
struct Staging {
    QItem* items;
    Seg* segs;
    Chunk* chunks;
    Job* jobs;
    uint32_t* item_job;
    unsigned long long* res;  // 2 words per query, written by k_sfinish
};

struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    void* host = nullptr;   // the pinned block
    Staging h{}, d{};       // host / device views of it
    unsigned long long* d_res = nullptr;  // [max_batch][4] accumulators + the queue word
    uint32_t* H = nullptr;
    uint32_t* bnd = nullptr;
    void* SP = nullptr;
    std::vector<Query*> qs;
    bool busy = false;
    bool radix = false;     // its batch holds the server's radix plan
};

}  // namespace

struct pgpu_server {
    pgpu_ctx* ctx = nullptr;
    pgpu_server_cfg cfg{};
    Caps caps;
    std::vector<Table> tables;
    TDesc* d_tabs = nullptr;
    void* R = nullptr;        // resident partitioned rows (Row<int64, true>)
    uint64_t R_used = 0, R_cap = 0;
    uint32_t* RB = nullptr;   // resident partition bounds
    uint64_t RB_used = 0, RB_cap = 0;
    int join_grid = 0;
    uint64_t resident = 0;
    uint64_t reserved = 0;  // HBM held by admitted, unfinished queries
    std::vector<Slot> slots;
    pgpu_join_plan* plain = nullptr;
    uint64_t plain_maxR = 0, plain_maxS = 0;
    // The radix route of lone large queries: one plan (its build-side scratch
    // sized for the largest table, its probe side for caps.rows), used by one
    // slot at a time.
    pgpu_join_plan* rplan = nullptr;
    uint64_t rplan_maxR = 0, rplan_bytes = 0;
    bool rplan_busy = false;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Query*> pending;
    std::deque<Query*> finished;
    uint64_t submitted = 0, completed = 0, batches = 0, part_queries = 0;
    std::thread th;
    bool stop = false;
    int error = PGPU_OK;
    std::string error_msg;
    double t0 = 0;
};

namespace {

void set_error(pgpu_server* s, int rc) {
    std::lock_guard<std::mutex> g(s->mu);
    if (s->error == PGPU_OK) {
        s->error = rc;
        s->error_msg = pgpu_last_error();
    }
}

// Chunk size of a probe batch on the partitioned path: about 256 blocks per
// batch (one per CU), so one query alone still spreads over the device (a
// 2^24-row probe of an 8M-row table: 0.52 ms at 128 blocks, 0.44 ms at 256,
// 0.455 ms at 512).
uint32_t chunk_rows_for(uint64_t m) {
    const uint64_t c = ceil_div(std::max<uint64_t>(m, 1), 256);
    return (uint32_t)std::min<uint64_t>(kChunkRows, std::max<uint64_t>(4096, ceil_div(c, 4096) * 4096));
}

// Rows per random-probe work item of a batch whose random-probe queries hold
// `rows` rows in all: about kBatchItems items, so a lone small query spreads
// over many CUs instead of walking its probes on one (a 2^14-row probe of a
// 2^22-row table takes ~180 us on one workgroup, its random slot reads
// latency-bound), while saturated batches keep 16K-row items.
uint64_t item_rows_for(uint64_t rows) {
    const uint64_t r = ceil_div(ceil_div(std::max<uint64_t>(rows, 1), kBatchItems), kMinItemRows) * kMinItemRows;
    return std::min(kItemRows, r);
}

// Work items per partition of a partitioned query: probes much larger than
// their table (few partitions) are cut into slices of about kSliceRows rows.
uint32_t split_for(uint64_t m, uint32_t bits) {
    const uint64_t per = ceil_div(std::max<uint64_t>(m, 1), 1ull << bits);
    return (uint32_t)std::min<uint64_t>(kMaxSplit, std::max<uint64_t>(1, ceil_div(per, kSliceRows)));
}

bool partitioned(const pgpu_server* s, const Query* q) {
    const uint64_t pmin = s->cfg.part_min_rows ? s->cfg.part_min_rows : (1u << 20);
    if (pmin == UINT32_MAX || q->m == 0) return false;
    const Table& t = s->tables[q->table];
    return q->m >= pmin && (pmin == 1 || 4 * q->m >= t.n) && q->m <= s->caps.rows;
}

// Completes the queries of a slot whose event has fired.
void reap(pgpu_server* s, Slot& sl) {
    const double t = now_s();
    std::lock_guard<std::mutex> g(s->mu);
    for (size_t i = 0; i < sl.qs.size(); ++i) {
        Query* q = sl.qs[i];
        q->matches = sl.h.res[2 * i];
        q->sum = sl.h.res[2 * i + 1];
        q->t_done = t;
        s->reserved -= q->reserve;
        s->finished.push_back(q);
        ++s->completed;
    }
    sl.qs.clear();
    sl.busy = false;
    if (sl.radix) s->rplan_busy = false;
    sl.radix = false;
    s->cv.notify_all();
}

// Admits queued queries into a free slot and launches them.  Called with
// no lock held; returns true if it launched something.
bool launch_batch(pgpu_server* s, Slot& sl) {
    std::vector<Query*> qs;
    uint64_t items = 0, rows = 0, hist = 0, bnd = 0, chunks = 0, pitems = 0, rrows = 0;
    const double t = now_s();
    {
        std::lock_guard<std::mutex> g(s->mu);
        while (!s->pending.empty() && qs.size() < s->cfg.max_batch) {
            Query* q = s->pending.front();
            if (s->resident + s->reserved + q->reserve > s->cfg.hbm_budget) break;  // FIFO: the head waits
            if (partitioned(s, q)) {
                const uint32_t bits = s->tables[q->table].bits;
                const uint64_t nc = ceil_div(q->m, chunk_rows_for(q->m));
                const uint64_t ni = (1ull << bits) * split_for(q->m, bits);
                if (rows + q->m > s->caps.rows || hist + (nc << bits) > s->caps.hist ||
                    bnd + (1ull << bits) + 1 > s->caps.bnd || chunks + nc > s->caps.chunks ||
                    pitems + ni > s->caps.pitems)
                    break;
                rows += q->m;
                hist += nc << bits;
                bnd += (1ull << bits) + 1;
                chunks += nc;
                pitems += ni;
            } else {
                const uint64_t qi = ceil_div(q->m, kItemRows);
                if (items + qi > s->cfg.max_items) break;
                items += qi;
                rrows += q->m;
            }
            s->pending.pop_front();
            s->reserved += q->reserve;
            q->t_admit = t;
            qs.push_back(q);
        }
    }
    if (qs.empty()) return false;
    ctx_activate(s->ctx);
    const uint32_t bid = (uint32_t)s->batches++;
    uint32_t ni = 0, nseg = 0, nchunk = 0, nitem = 0, maxbits = 0;
    uint64_t prow = 0, phist = 0, pbnd = 0;
    // The one large query of a batch (if it holds exactly one) on a table with
    // the radix layout goes through the 2-pass radix partitioning of its probe
    // keys and the LDS join against that layout: one pass over 2048- to
    // 4096-way runs costs more than two 32- to 128-way passes, and the LDS
    // join runs at ~1000 build rows per partition instead of ~3000.  Several
    // large queries share the segmented pass instead (one launch for all).
    int radix_q = -1;
    {
        int nlarge = 0, last = -1;
        for (uint32_t i = 0; i < qs.size(); ++i)
            if (partitioned(s, qs[i])) ++nlarge, last = (int)i;
        if (nlarge == 1 && s->rplan && !s->rplan_busy && s->tables[qs[last]->table].rrows &&
            qs[last]->m <= s->caps.rows)
            radix_q = last;
    }
    // Admission counted kItemRows items; smaller ones are used when they fit.
    uint64_t irows = item_rows_for(rrows), nsmall = 0;
    for (Query* q : qs)
        if (!partitioned(s, q)) nsmall += ceil_div(q->m, irows);
    if (nsmall > s->cfg.max_items) irows = kItemRows;
    for (uint32_t i = 0; i < qs.size(); ++i) {
        Query* q = qs[i];
        q->batch = bid;
        q->batch_size = (uint32_t)qs.size();
        if (!partitioned(s, q)) {
            for (uint64_t r = 0; r < q->m; r += irows)
                sl.h.items[ni++] = QItem{q->probe, r, std::min(q->m, r + irows), i, q->table};
            continue;
        }
        if ((int)i == radix_q) continue;
        const Table& tb = s->tables[q->table];
        Seg sg{};
        sg.key = q->probe;
        sg.pay = nullptr;
        sg.n = q->m;
        sg.dst = prow;
        sg.hist = phist;
        sg.bnd = pbnd;
        sg.chunk0 = nchunk;
        sg.chunk_rows = chunk_rows_for(q->m);
        sg.nchunks = (uint32_t)ceil_div(q->m, sg.chunk_rows);
        sg.bits = tb.bits;
        maxbits = std::max(maxbits, tb.bits);
        for (uint32_t c = 0; c < sg.nchunks; ++c)
            sl.h.chunks[nchunk++] = Chunk{nseg, 0, (uint64_t)c * sg.chunk_rows};
        Job jb{};
        jb.rdst = tb.rdst;
        jb.rbnd = tb.rbnd;
        jb.sdst = prow;
        jb.sbnd = pbnd;
        jb.bits = tb.bits;
        jb.item0 = nitem;
        jb.split = split_for(q->m, tb.bits);
        for (uint64_t d = 0; d < ((uint64_t)jb.split << tb.bits); ++d) sl.h.item_job[nitem++] = i;
        sl.h.jobs[i] = jb;
        sl.h.segs[nseg++] = sg;
        prow += q->m;
        phist += (uint64_t)sg.nchunks << tb.bits;
        pbnd += (1ull << tb.bits) + 1;
    }
    std::atomic_thread_fence(std::memory_order_seq_cst);  // staging written before the launches read it
    int rc = PGPU_OK;
    uint32_t* queue = (uint32_t*)(sl.d_res + 4ull * s->cfg.max_batch);
    if (ni) hipLaunchKernelGGL(k_sprobe, dim3(ni), dim3(kProbeThreads), 0, sl.st, sl.d.items, s->d_tabs, sl.d_res);
    if (rc == PGPU_OK && hipGetLastError() != hipSuccess) rc = fail(PGPU_EDEVICE, "probe launch failed");
    if (rc == PGPU_OK && nseg) {
        rc = batch::launch_hist(sl.st, 8, sl.d.segs, sl.d.chunks, 0, nchunk, sl.H);
        if (rc == PGPU_OK) rc = batch::launch_scan(sl.st, maxbits, sl.d.segs, 0, nseg, sl.H, sl.bnd);
        if (rc == PGPU_OK)
            rc = batch::launch_scatter(sl.st, 8, false, maxbits, sl.d.segs, sl.d.chunks, 0, nchunk, sl.H, sl.SP);
        if (rc == PGPU_OK)
            rc = batch::launch_join(sl.st, 8, false, kVariant, std::min<int>(s->join_grid, (int)nitem), sl.d.jobs,
                                    sl.d.item_job, nitem, s->R, sl.SP, s->RB, sl.bnd, queue, sl.d_res);
    }
    if (rc == PGPU_OK && radix_q >= 0) {
        const Query* q = qs[radix_q];
        const Table& tb = s->tables[q->table];
        s->rplan_busy = true;
        sl.radix = true;
        ++nseg;
        rc = radix_probe_resident(s->rplan, sl.st, (const int64_t*)q->probe, q->m, tb.rbits, tb.rrows, tb.rbase,
                                  tb.rtot, sl.d_res + 4ull * radix_q);
    }
    if (rc == PGPU_OK) {
        hipLaunchKernelGGL(k_sfinish, dim3(1), dim3(256), 0, sl.st, sl.d_res, (uint32_t)qs.size(), sl.d.res, queue);
        if (hipGetLastError() != hipSuccess) rc = fail(PGPU_EDEVICE, "finish launch failed");
    }
    if (rc == PGPU_OK && hipEventRecord(sl.ev, sl.st) != hipSuccess) rc = fail(PGPU_EDEVICE, "event record failed");
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->part_queries += nseg;
    }
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
    return PGPU_OK;
}

// Grows a resident arena to hold `need` elements of `es` bytes, keeping
// its contents.
int grow(pgpu_server* s, void** p, uint64_t* cap, uint64_t used, uint64_t need, uint64_t es) {
    if (need <= *cap) return PGPU_OK;
    const uint64_t ncap = std::max(need, 2 * *cap);
    void* np = nullptr;
    PGPU_TRY(ctx_alloc(s->ctx, ncap * es, &np));
    if (*p && used) {
        if (hipMemcpy(np, *p, used * es, hipMemcpyDeviceToDevice) != hipSuccess) {
            ctx_free(s->ctx, np);
            return fail(PGPU_EDEVICE, "resident arena copy failed");
        }
    }
    if (*p) PGPU_TRY(ctx_free(s->ctx, *p));
    s->resident += (ncap - *cap) * es;
    *p = np;
    *cap = ncap;
    return PGPU_OK;
}

// Partitions a table's rows into the resident arena (the batched join's
// segmented partition pass over one segment).
int partition_table(pgpu_server* s, Table& t) {
    t.bits = batch::bits_for(t.n, kPartTarget);
    const uint64_t nb = (1ull << t.bits) + 1;
    PGPU_TRY(grow(s, &s->R, &s->R_cap, s->R_used, s->R_used + std::max<uint64_t>(t.n, 1), 16));
    PGPU_TRY(grow(s, (void**)&s->RB, &s->RB_cap, s->RB_used, s->RB_used + nb, 4));
    t.rdst = s->R_used;
    t.rbnd = s->RB_used;
    batch::Layout lay;
    lay.add(t.key, t.pay, t.n, t.bits, 0);
    Seg sg = lay.segs[0];
    sg.dst = t.rdst;
    sg.bnd = t.rbnd;
    void *d_seg = nullptr, *d_chunks = nullptr, *d_H = nullptr;
    int rc = ctx_alloc(s->ctx, sizeof(Seg), &d_seg);
    if (rc == PGPU_OK) rc = ctx_alloc(s->ctx, std::max<size_t>(lay.chunks.size(), 1) * sizeof(Chunk), &d_chunks);
    if (rc == PGPU_OK) rc = ctx_alloc(s->ctx, std::max<uint64_t>(lay.hist, 1) * 4, &d_H);
    hipStream_t st = s->ctx->stream;
    if (rc == PGPU_OK &&
        (hipMemcpy(d_seg, &sg, sizeof sg, hipMemcpyHostToDevice) != hipSuccess ||
         (!lay.chunks.empty() &&
          hipMemcpy(d_chunks, lay.chunks.data(), lay.chunks.size() * sizeof(Chunk), hipMemcpyHostToDevice) !=
              hipSuccess)))
        rc = fail(PGPU_EDEVICE, "table partition upload failed");
    const uint32_t nc = (uint32_t)lay.chunks.size();
    if (rc == PGPU_OK) rc = batch::launch_hist(st, 8, (Seg*)d_seg, (Chunk*)d_chunks, 0, nc, (uint32_t*)d_H);
    if (rc == PGPU_OK) rc = batch::launch_scan(st, t.bits, (Seg*)d_seg, 0, 1, (uint32_t*)d_H, s->RB);
    if (rc == PGPU_OK)
        rc = batch::launch_scatter(st, 8, true, t.bits, (Seg*)d_seg, (Chunk*)d_chunks, 0, nc, (uint32_t*)d_H, s->R);
    if (rc == PGPU_OK && hipStreamSynchronize(st) != hipSuccess) rc = fail(PGPU_EDEVICE, "table partition failed");
    if (d_seg) ctx_free(s->ctx, d_seg);
    if (d_chunks) ctx_free(s->ctx, d_chunks);
    if (d_H) ctx_free(s->ctx, d_H);
    if (rc != PGPU_OK) return rc;
    s->R_used += std::max<uint64_t>(t.n, 1);
    s->RB_used += nb;
    return PGPU_OK;
}

// Bytes of a radix plan for builds of maxR rows and probes of maxS rows (its
// radix buffers, as pgpu_join_plan_create sizes them, rounded up).
uint64_t radix_plan_bytes(uint64_t maxR, uint64_t maxS) {
    const uint64_t nb = 1ull << 14, blocks = 512;
    return 2 * 16 * std::max<uint64_t>(maxR, 1) + 2 * 8 * std::max<uint64_t>(maxS, 1) + 2 * blocks * nb * 4 +
           (2 * blocks + 2) * 256 * 8 + 64 * nb + 32 * (nb + ceil_div(maxS, 32768) + 1) + (1 << 20);
}

// The table's rows in the radix layout (and the plan that probes it, grown to
// the largest table), for tables whose segmented layout has more than 1024
// partitions, when the budget has room; other tables keep only the segmented
// layout.  Below that fan-out the one-pass segmented partitioning of a probe
// beats two radix passes (a 2^25-row probe of a 2^20-row table: 0.39 ms
// segmented, 0.46 ms radix; of a 2^23-row table: 0.63 / 0.55 ms).  Runs while
// no query is in flight.
constexpr uint32_t kRadixMinBits = 11;
int radix_table(pgpu_server* s, Table& t) {
    if (!t.n || t.bits < kRadixMinBits) return PGPU_OK;
    int bits = (int)ilog2_ceil(ceil_div(t.n, 1024));
    bits = std::max(2, std::min(14, bits));
    const uint64_t nb = 1ull << bits;
    const uint64_t cache = 16 * t.n + 8 * (nb + 1) + 4 * nb;
    const uint64_t maxR = std::max(s->rplan_maxR, t.n);
    const uint64_t plan = maxR > s->rplan_maxR ? radix_plan_bytes(maxR, s->caps.rows) : 0;
    if (s->resident - s->rplan_bytes + plan + cache > s->cfg.hbm_budget) return PGPU_OK;
    if (plan) {
        pgpu_join_cfg c{8, 0, 0, PGPU_ALGO_RADIX, 0};
        pgpu_join_plan* np = nullptr;
        if (pgpu_join_plan_create(s->ctx, &c, maxR, s->caps.rows, &np) != PGPU_OK) return PGPU_OK;  // no room
        if (s->rplan) pgpu_join_plan_destroy(s->rplan);
        s->rplan = np;
        s->resident = s->resident - s->rplan_bytes + plan;
        s->rplan_bytes = plan;
        s->rplan_maxR = maxR;
    }
    int rc = ctx_alloc(s->ctx, 16 * t.n, &t.rrows);
    if (rc == PGPU_OK) rc = ctx_alloc(s->ctx, 8 * (nb + 1), (void**)&t.rbase);
    if (rc == PGPU_OK) rc = ctx_alloc(s->ctx, 4 * nb, (void**)&t.rtot);
    if (rc == PGPU_OK)
        rc = radix_build_resident(s->rplan, s->ctx->stream, (const int64_t*)t.key, (const int64_t*)t.pay, t.n, bits,
                                  t.rrows, t.rbase, t.rtot);
    if (rc == PGPU_OK && hipStreamSynchronize(s->ctx->stream) != hipSuccess)
        rc = fail(PGPU_EDEVICE, "radix layout of a table failed");
    if (rc != PGPU_OK) {
        for (void* p : {t.rrows, (void*)t.rbase, (void*)t.rtot})
            if (p) (void)ctx_free(s->ctx, p);
        t.rrows = nullptr;
        t.rbase = nullptr;
        t.rtot = nullptr;
        return rc;
    }
    t.rbits = bits;
    t.rbytes = cache;
    s->resident += cache;
    return PGPU_OK;
}

template <typename T>
T* carve(char*& p, uint64_t n) {
    T* r = (T*)p;
    p += (n * sizeof(T) + 255) & ~uint64_t(255);
    return r;
}

Staging carve_all(char* base, const pgpu_server_cfg& c, const Caps& k) {
    Staging s;
    char* p = base;
    s.items = carve<QItem>(p, c.max_items);
    s.segs = carve<Seg>(p, c.max_batch);
    s.chunks = carve<Chunk>(p, k.chunks);
    s.jobs = carve<Job>(p, c.max_batch);
    s.item_job = carve<uint32_t>(p, k.pitems);
    s.res = carve<unsigned long long>(p, 2ull * c.max_batch);
    return s;
}

uint64_t staging_bytes(const pgpu_server_cfg& c, const Caps& k) {
    uint64_t b = 0;
    b += (c.max_items * sizeof(QItem) + 255) & ~uint64_t(255);
    b += (c.max_batch * sizeof(Seg) + 255) & ~uint64_t(255);
    b += (k.chunks * sizeof(Chunk) + 255) & ~uint64_t(255);
    b += (c.max_batch * sizeof(Job) + 255) & ~uint64_t(255);
    b += (k.pitems * 4 + 255) & ~uint64_t(255);
    b += (2ull * c.max_batch * 8 + 255) & ~uint64_t(255);
    return b;
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
    if (s->cfg.slots) {
        int rc = batch::join_grid(8, false, kVariant, &s->join_grid);
        if (rc != PGPU_OK) return cleanup(rc);
    }
    s->slots.resize(s->cfg.slots);
    const uint64_t sbytes = staging_bytes(s->cfg, s->caps);
    for (Slot& sl : s->slots) {
        if (hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess)
            return cleanup(fail(PGPU_EDEVICE, "stream / event creation failed"));
        if (hipHostMalloc(&sl.host, sbytes, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
            return cleanup(fail(PGPU_ENOMEM, "pinned staging allocation failed"));
        void* dview = nullptr;
        if (hipHostGetDevicePointer(&dview, sl.host, 0) != hipSuccess)
            return cleanup(fail(PGPU_EDEVICE, "hipHostGetDevicePointer failed"));
        sl.h = carve_all((char*)sl.host, s->cfg, s->caps);
        sl.d = carve_all((char*)dview, s->cfg, s->caps);
        const uint64_t res_words = 4ull * s->cfg.max_batch + 1;
        int rc = ctx_alloc(ctx, res_words * 8, (void**)&sl.d_res);
        if (rc == PGPU_OK) rc = ctx_alloc(ctx, s->caps.hist * 4, (void**)&sl.H);
        if (rc == PGPU_OK) rc = ctx_alloc(ctx, s->caps.bnd * 4, (void**)&sl.bnd);
        if (rc == PGPU_OK) rc = ctx_alloc(ctx, s->caps.rows * 8, &sl.SP);
        if (rc != PGPU_OK) return cleanup(rc);
        s->resident += res_words * 8 + s->caps.hist * 4 + s->caps.bnd * 4 + s->caps.rows * 8;
        hipLaunchKernelGGL(k_szero, dim3(4), dim3(256), 0, sl.st, sl.d_res, res_words);
        if (hipStreamSynchronize(sl.st) != hipSuccess) return cleanup(fail(PGPU_EDEVICE, "slot init failed"));
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
        if (sl.host) (void)hipHostFree(sl.host);
        for (void* p : {(void*)sl.d_res, (void*)sl.H, (void*)sl.bnd, sl.SP})
            if (p) (void)ctx_free(s->ctx, p);
    }
    for (Table& t : s->tables) {
        if (t.slots) (void)ctx_free(s->ctx, t.slots);
        if (t.cnt) (void)ctx_free(s->ctx, t.cnt);
        if (t.special) (void)ctx_free(s->ctx, t.special);
        for (void* p : {t.rrows, (void*)t.rbase, (void*)t.rtot})
            if (p) (void)ctx_free(s->ctx, p);
    }
    if (s->rplan) pgpu_join_plan_destroy(s->rplan);
    if (s->d_tabs) (void)ctx_free(s->ctx, s->d_tabs);
    if (s->R) (void)ctx_free(s->ctx, s->R);
    if (s->RB) (void)ctx_free(s->ctx, s->RB);
    if (s->plain) pgpu_join_plan_destroy(s->plain);
    for (Query* q : s->pending) delete q;
    for (Query* q : s->finished) delete q;
    delete s;
    return PGPU_OK;
}

int pgpu_server_add_table(pgpu_server* s, const void* d_key, const void* d_pay, uint64_t n, uint32_t* table_id) {
    if (!s || !table_id || (n && (!d_key || !d_pay))) return fail(PGPU_EINVAL, "pgpu_server_add_table: null argument");
    if (s->cfg.key_bytes != 8) return fail(PGPU_EUNSUPPORTED, "served tables have int64 keys");
    if (n >= (1ull << 32)) return fail(PGPU_EINVAL, "a served table holds fewer than 2^32 rows");
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
        const uint64_t part_bytes = 16 * std::max<uint64_t>(n, 1) + 4 * ((1ull << batch::bits_for(n, kPartTarget)) + 1);
        if (s->resident + t.bytes + part_bytes > s->cfg.hbm_budget)
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
                               (const long long*)d_key, (const long long*)d_pay, n, 1u, t.slots, t.cnt, 64 - logcap,
                               t.cap - 1, t.special);
        PGPU_LAUNCH_CHECK();
        PGPU_HIP(hipStreamSynchronize(st));
        s->resident += t.bytes;
        PGPU_TRY(partition_table(s, t));
        PGPU_TRY(radix_table(s, t));
    } else {
        s->plain_maxR = std::max(s->plain_maxR, n);  // plain mode keeps only the columns
    }
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
    pgpu_join_cfg c{8, 0, 0, PGPU_ALGO_AUTO, 0};
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
    if (m >= (1ull << 32)) return fail(PGPU_EINVAL, "a probe batch holds fewer than 2^32 rows");
    auto* q = new Query();
    q->tag = tag;
    q->table = table;
    q->probe = (const long long*)d_probe;
    q->m = m;
    q->reserve = m * 8 + 64;  // the probe batch is held in HBM until its result is out
    q->t_arrival = t_arrival > 0 ? s->t0 + t_arrival : now_s();
    if (s->cfg.mode == PGPU_SERVE_BATCHED && !partitioned(s, q) && ceil_div(m, kItemRows) > s->cfg.max_items) {
        delete q;
        return fail(PGPU_EINVAL, "probe batch of %llu rows exceeds one launch's %u work items",
                    (unsigned long long)m, s->cfg.max_items);
    }
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
    st->partitioned = s->part_queries;
    return PGPU_OK;
}

}  // extern "C"

// ====================================================================== persistence
// A served table saved as one memory segment and attached by a server
// without its columns and without the partition pass: HashTable::serialize
// / deserialize (util/hashtable.cpp:99-201), which copy a table's bucket
// memory into a segment made by creatememsegment and back.  The segment
// holds the table's radix-partitioned (key, payload) rows and their bounds;
// the hash slots, about twice as many bytes as the rows and mostly empty,
// are derived again on attach by one build kernel over the rows.
// Layout: a 128-byte header, then the rows and the bounds, each at a 4 KiB
// boundary.
namespace {
struct SegHeader {
    char magic[8];     // "PGPUTBL1"
    uint32_t version;  // 1
    uint32_t key_bytes;
    uint64_t n;        // rows
    uint32_t bits;     // partition bits (2^bits + 1 bounds)
    uint32_t pad;
    uint64_t rows_bytes, bounds_bytes;
    uint64_t reserved[10];
};
static_assert(sizeof(SegHeader) == 128, "segment header");
constexpr uint64_t kSegAlign = 4096;
constexpr char kSegMagic[8] = {'P', 'G', 'P', 'U', 'T', 'B', 'L', '1'};
constexpr int kSegCopyThreads = 8;
constexpr uint64_t kSegStage = 64ull << 20;  // pinned staging block of an attach

inline uint64_t seg_round(uint64_t b) { return (b + kSegAlign - 1) & ~(kSegAlign - 1); }

// As creatememsegment / attachmemsegment (util/hashtable.cpp:123-176): a
// /dev/shm/<name> path goes through shm_open, any other through open.
int seg_fd(const char* path, bool create) {
    const std::string p(path);
    const int flags = create ? (O_RDWR | O_CREAT | O_TRUNC) : O_RDONLY;
    if (p.compare(0, 9, "/dev/shm/") == 0) return shm_open(p.substr(8).c_str(), flags, 0644);
    return open(path, flags, 0644);
}

// pread of `bytes` at file offset `off` into dst, split over
// kSegCopyThreads threads (one thread's copy runs at a fraction of the
// host's memory bandwidth); false on a short read.
bool parallel_read(int fd, char* dst, uint64_t off, uint64_t bytes) {
    auto part = [fd, dst, off](uint64_t a, uint64_t b) {
        while (a < b) {
            const ssize_t r = pread(fd, dst + a, b - a, (off_t)(off + a));
            if (r <= 0) return false;
            a += (uint64_t)r;
        }
        return true;
    };
    if (bytes < (8u << 20)) return part(0, bytes);
    std::thread th[kSegCopyThreads];
    bool ok[kSegCopyThreads];
    const uint64_t per = (bytes / kSegCopyThreads + 4095) & ~4095ull;
    for (int i = 0; i < kSegCopyThreads; ++i) {
        const uint64_t a = std::min<uint64_t>(bytes, i * per), b = std::min<uint64_t>(bytes, a + per);
        th[i] = std::thread([&part, &ok, i, a, b] { ok[i] = part(a, b); });
    }
    bool all = true;
    for (int i = 0; i < kSegCopyThreads; ++i) {
        th[i].join();
        all = all && ok[i];
    }
    return all;
}
}  // namespace

extern "C" {

int pgpu_server_save_table(pgpu_server* s, uint32_t table, const char* path) {
    if (!s || !path) return fail(PGPU_EINVAL, "pgpu_server_save_table: null argument");
    if (s->cfg.mode != PGPU_SERVE_BATCHED) return fail(PGPU_ESTATE, "only a batched server keeps built tables");
    if (table >= s->tables.size()) return fail(PGPU_EINVAL, "unknown table %u", table);
    PGPU_TRY(ctx_activate(s->ctx));
    const Table& t = s->tables[table];
    SegHeader h;
    memset(&h, 0, sizeof h);
    memcpy(h.magic, kSegMagic, 8);
    h.version = 1;
    h.key_bytes = 8;
    h.n = t.n;
    h.bits = t.bits;
    h.rows_bytes = 16 * std::max<uint64_t>(t.n, 1);
    h.bounds_bytes = 4 * ((1ull << t.bits) + 1);
    const uint64_t size = kSegAlign + seg_round(h.rows_bytes) + seg_round(h.bounds_bytes);
    const int fd = seg_fd(path, true);
    if (fd < 0) return fail(PGPU_EINVAL, "pgpu_server_save_table: cannot create segment %s", path);
    if (ftruncate(fd, (off_t)size) != 0) {
        close(fd);
        return fail(PGPU_ENOMEM, "pgpu_server_save_table: cannot size segment %s to %llu bytes", path,
                    (unsigned long long)size);
    }
    char* m = (char*)mmap(nullptr, size, PROT_READ | PROT_WRITE, MAP_SHARED, fd, 0);
    close(fd);
    if (m == MAP_FAILED) return fail(PGPU_ENOMEM, "pgpu_server_save_table: cannot map segment %s", path);
    memcpy(m, &h, sizeof h);
    int rc = PGPU_OK;
    if (hipMemcpy(m + kSegAlign, (const char*)s->R + 16 * t.rdst, h.rows_bytes, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(m + kSegAlign + seg_round(h.rows_bytes), s->RB + t.rbnd, h.bounds_bytes, hipMemcpyDeviceToHost) !=
            hipSuccess)
        rc = fail(PGPU_EDEVICE, "pgpu_server_save_table: copy of table %u failed", table);
    munmap(m, size);
    return rc;
}

int pgpu_server_load_table(pgpu_server* s, const char* path, uint32_t* table_id) {
    if (!s || !path || !table_id) return fail(PGPU_EINVAL, "pgpu_server_load_table: null argument");
    if (s->cfg.mode != PGPU_SERVE_BATCHED) return fail(PGPU_ESTATE, "only a batched server attaches built tables");
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->submitted != s->completed || !s->pending.empty())
            return fail(PGPU_ESTATE, "tables are registered before queries are submitted");
    }
    const int fd = seg_fd(path, false);
    if (fd < 0) return fail(PGPU_EINVAL, "pgpu_server_load_table: no segment %s", path);
    struct stat sb;
    SegHeader h;
    if (fstat(fd, &sb) != 0 || (uint64_t)sb.st_size < kSegAlign || !parallel_read(fd, (char*)&h, 0, sizeof h)) {
        close(fd);
        return fail(PGPU_EINVAL, "pgpu_server_load_table: %s is not a table segment", path);
    }
    const uint64_t size = (uint64_t)sb.st_size;
    const bool ok = memcmp(h.magic, kSegMagic, 8) == 0 && h.version == 1 && h.key_bytes == 8 &&
                    h.n < (1ull << 32) && h.bits <= 16 && h.rows_bytes == 16 * std::max<uint64_t>(h.n, 1) &&
                    h.bounds_bytes == 4 * ((1ull << h.bits) + 1) &&
                    kSegAlign + seg_round(h.rows_bytes) + seg_round(h.bounds_bytes) <= size;
    if (!ok) {
        close(fd);
        return fail(PGPU_EINVAL, "pgpu_server_load_table: %s is not a valid table segment", path);
    }
    int rc = ctx_activate(s->ctx);
    Table t;
    t.n = h.n;
    t.bits = h.bits;
    t.cap = 1ull << std::max<uint32_t>(4, ilog2_ceil(2 * std::max<uint64_t>(t.n, 1)));
    t.bytes = t.cap * (sizeof(SSlot) + 4) + 32;
    const uint64_t nb = (1ull << t.bits) + 1;
    if (rc == PGPU_OK && s->resident + t.bytes + h.rows_bytes + h.bounds_bytes > s->cfg.hbm_budget)
        rc = fail(PGPU_EBUDGET, "resident table of %llu rows exceeds the server's HBM budget", (unsigned long long)t.n);
    if (rc == PGPU_OK) rc = ctx_alloc(s->ctx, t.cap * sizeof(SSlot), (void**)&t.slots);
    if (rc == PGPU_OK) rc = ctx_alloc(s->ctx, t.cap * 4, (void**)&t.cnt);
    if (rc == PGPU_OK) rc = ctx_alloc(s->ctx, 32, (void**)&t.special);
    if (rc == PGPU_OK) rc = grow(s, &s->R, &s->R_cap, s->R_used, s->R_used + std::max<uint64_t>(t.n, 1), 16);
    if (rc == PGPU_OK) rc = grow(s, (void**)&s->RB, &s->RB_cap, s->RB_used, s->RB_used + nb, 4);
    // rows: the segment into two pinned staging blocks in turn (several
    // threads copy one block while the previous one's DMA runs)
    char* pin[2] = {nullptr, nullptr};
    uint64_t pin_cap = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    for (int i = 0; i < 2 && rc == PGPU_OK; ++i) {
        rc = ctx_host_alloc_reuse(s->ctx, kSegStage, (void**)&pin[i], &pin_cap);
        if (rc == PGPU_OK && hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess)
            rc = fail(PGPU_EDEVICE, "hipEventCreateWithFlags failed");
    }
    hipStream_t st = s->ctx->stream;
    if (rc == PGPU_OK) {
        t.rdst = s->R_used;
        t.rbnd = s->RB_used;
        char* dst = (char*)s->R + 16 * t.rdst;
        uint64_t i = 0;
        for (uint64_t off = 0; off < h.rows_bytes && rc == PGPU_OK; off += kSegStage, ++i) {
            const uint64_t len = std::min(kSegStage, h.rows_bytes - off);
            if (i >= 2 && hipEventSynchronize(ev[i & 1]) != hipSuccess) {
                rc = fail(PGPU_EDEVICE, "pgpu_server_load_table: staging wait failed");
                break;
            }
            if (!parallel_read(fd, pin[i & 1], kSegAlign + off, len)) {
                rc = fail(PGPU_EINVAL, "pgpu_server_load_table: short read of %s", path);
                break;
            }
            if (hipMemcpyAsync(dst + off, pin[i & 1], len, hipMemcpyHostToDevice, st) != hipSuccess ||
                hipEventRecord(ev[i & 1], st) != hipSuccess)
                rc = fail(PGPU_EDEVICE, "pgpu_server_load_table: upload of %s failed", path);
        }
        // the bounds (at most 256 KiB) through the first block once its DMA is done
        if (rc == PGPU_OK && (hipStreamSynchronize(st) != hipSuccess ||
                              !parallel_read(fd, pin[0], kSegAlign + seg_round(h.rows_bytes), h.bounds_bytes) ||
                              hipMemcpyAsync(s->RB + t.rbnd, pin[0], h.bounds_bytes, hipMemcpyHostToDevice, st) !=
                                  hipSuccess))
            rc = fail(PGPU_EDEVICE, "pgpu_server_load_table: upload of %s failed", path);
    }
    if (rc == PGPU_OK) {  // the hash slots, derived from the partitioned rows
        const int logcap = (int)ilog2_ceil(t.cap);
        const long long* rows = (const long long*)((const char*)s->R + 16 * t.rdst);
        hipLaunchKernelGGL(k_tclear, dim3((unsigned)std::min<uint64_t>(ceil_div(t.cap, 256), 4096)), dim3(256), 0, st,
                           t.slots, t.cnt, t.cap, t.special);
        if (t.n)
            hipLaunchKernelGGL(k_tbuild, dim3((unsigned)std::min<uint64_t>(ceil_div(t.n, 256), 4096)), dim3(256), 0,
                               st, rows, rows + 1, t.n, 2u, t.slots, t.cnt, 64 - logcap, t.cap - 1, t.special);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
            rc = fail(PGPU_EDEVICE, "pgpu_server_load_table: table build failed");
    }
    if (hipStreamSynchronize(st) != hipSuccess && rc == PGPU_OK)
        rc = fail(PGPU_EDEVICE, "pgpu_server_load_table: upload of %s failed", path);
    for (int i = 0; i < 2; ++i) {
        if (pin[i]) ctx_host_free_reuse(s->ctx, pin[i]);
        if (ev[i]) (void)hipEventDestroy(ev[i]);
    }
    close(fd);
    if (rc != PGPU_OK) {
        if (t.slots) ctx_free(s->ctx, t.slots);
        if (t.cnt) ctx_free(s->ctx, t.cnt);
        if (t.special) ctx_free(s->ctx, t.special);
        return rc;
    }
    s->resident += t.bytes;
    s->R_used += std::max<uint64_t>(t.n, 1);
    s->RB_used += nb;
    s->tables.push_back(t);
    *table_id = (uint32_t)(s->tables.size() - 1);
    return upload_tabs(s);
}

}  // extern "C"
