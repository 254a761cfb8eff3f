// Batched join plans (C ABI): many independent joins bound at creation,
// replayed as one hipGraph, plus the autotuner that picks the join
// workgroup shape and partition size for a batch.
#include <algorithm>
#include <cstring>
#include <numeric>

#include "batch_internal.h"

using namespace pgpu;
using namespace pgpu::batch;

struct pgpu_batch_plan {
    pgpu_ctx* ctx = nullptr;
    pgpu_batch_cfg cfg{};
    uint32_t n = 0;
    std::vector<uint32_t> bits;
    Layout lr, ls;  // build / probe sides
    uint32_t n_items = 0;
    int grid = 0;
    Seg* segs = nullptr;      // build segs then probe segs
    Chunk* chunks = nullptr;  // build chunks then probe chunks
    uint32_t* H = nullptr;
    uint32_t* bnd = nullptr;  // build bounds then probe bounds
    void* RP = nullptr;
    void* SP = nullptr;
    Job* jobs = nullptr;
    uint32_t* item_job = nullptr;
    unsigned long long* res = nullptr;  // [n][4] then two words of queue counters (three used)
    uint32_t* redo = nullptr;           // variant 3: partitions handed to the general join kernel
    unsigned long long* h_res = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool pending = false;
    KernelTimer* timer = nullptr;
    std::vector<void*> owned;
};

namespace {

struct Timed {
    KernelTimer* t;
    int idx;
    Timed(KernelTimer* t_, const char* name) : t(t_), idx(t_->start(name)) {}
    ~Timed() { t->stop(idx); }
};

int balloc(pgpu_batch_plan* p, uint64_t bytes, void** out) {
    PGPU_TRY(ctx_alloc(p->ctx, std::max<uint64_t>(bytes, 16), out));
    p->owned.push_back(*out);
    return PGPU_OK;
}

int enqueue(pgpu_batch_plan* p, bool timed) {
    hipStream_t st = p->ctx->stream;
    KernelTimer* tm = timed ? p->timer : nullptr;
    const int kb = (int)p->cfg.key_bytes;
    const bool hvs = p->cfg.probe_payload != 0;
    const uint32_t nbc = (uint32_t)p->lr.chunks.size(), nsc = (uint32_t)p->ls.chunks.size();
    const uint64_t words = 4ull * p->n + 2;
    auto timed_run = [&](const char* name, auto&& fn) -> int {
        if (!tm) return fn();
        Timed t(tm, name);
        return fn();
    };
    PGPU_TRY(timed_run("batch_zero", [&] { return launch_zero(st, p->res, words); }));
    PGPU_TRY(timed_run("batch_hist", [&] { return launch_hist(st, kb, p->segs, p->chunks, 0, nbc + nsc, p->H); }));
    const uint32_t mb = p->bits.empty() ? 0u : *std::max_element(p->bits.begin(), p->bits.end());
    PGPU_TRY(timed_run("batch_scan", [&] { return launch_scan(st, mb, p->segs, 0, 2 * p->n, p->H, p->bnd); }));
    PGPU_TRY(timed_run("batch_scatter_build",
                       [&] { return launch_scatter(st, kb, true, mb, p->segs, p->chunks, 0, nbc, p->H, p->RP); }));
    PGPU_TRY(timed_run("batch_scatter_probe",
                       [&] { return launch_scatter(st, kb, hvs, mb, p->segs, p->chunks, nbc, nsc, p->H, p->SP); }));
    PGPU_TRY(timed_run("batch_join", [&] {
        return launch_join(st, kb, hvs, (int)p->cfg.variant, p->grid, p->jobs, p->item_job, p->n_items, p->RP, p->SP,
                           p->bnd, p->bnd + p->lr.bnd, (uint32_t*)(p->res + 4ull * p->n), p->res, p->redo);
    }));
    PGPU_HIP(hipMemcpyAsync(p->h_res, p->res, 4ull * p->n * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    return PGPU_OK;
}

}  // namespace

extern "C" {

int pgpu_batch_plan_create(pgpu_ctx* ctx, const pgpu_batch_cfg* cfg, const pgpu_batch_join* joins, uint32_t n,
                           pgpu_batch_plan** out) {
    if (!ctx || !cfg || !out || (n && !joins)) return fail(PGPU_EINVAL, "pgpu_batch_plan_create: null argument");
    *out = nullptr;
    if (cfg->key_bytes != 4 && cfg->key_bytes != 8)
        return fail(PGPU_EUNSUPPORTED, "batched join keys must be 4 or 8 bytes (got %u)", cfg->key_bytes);
    if (n == 0) return fail(PGPU_EINVAL, "pgpu_batch_plan_create: empty batch");
    pgpu_batch_cfg c = *cfg;
    if (c.variant == 0) c.variant = c.key_bytes == 4 ? 3 : 2;
    if (!variant_ok((int)c.variant, (int)c.key_bytes))
        return fail(PGPU_EINVAL, "join variant %u with %u-byte keys: unknown (variants 1 and 2 take any keys, 3 "
                    "takes 4-byte keys)", c.variant, c.key_bytes);
    if (c.part_rows == 0) c.part_rows = variant_target((int)c.variant);
    if (c.part_rows > (uint32_t)variant_piece((int)c.variant))
        return fail(PGPU_EINVAL, "part_rows %u exceeds the LDS piece of variant %u (%d rows)", c.part_rows,
                    c.variant, variant_piece((int)c.variant));
    for (uint32_t i = 0; i < n; ++i) {
        const pgpu_batch_join& j = joins[i];
        if (j.n_build >= (1ull << 32) || j.n_probe >= (1ull << 32))
            return fail(PGPU_EINVAL, "join %u: each side must hold fewer than 2^32 rows", i);
        if ((j.n_build && (!j.d_rkey || !j.d_rpay)) || (j.n_probe && !j.d_skey) ||
            (j.n_probe && c.probe_payload && !j.d_spay))
            return fail(PGPU_EINVAL, "join %u: null column", i);
    }
    PGPU_TRY(ctx_activate(ctx));
    auto* p = new pgpu_batch_plan();
    p->ctx = ctx;
    p->cfg = c;
    p->n = n;
    p->timer = new KernelTimer(ctx->stream);
    auto cleanup = [&](int rc) {
        pgpu_batch_plan_destroy(p);
        return rc;
    };
    // Work items of the largest joins first, so the tail of the persistent
    // queue holds the smallest items.
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return joins[a].n_build + joins[a].n_probe > joins[b].n_build + joins[b].n_probe;
    });
    p->bits.resize(n);
    for (uint32_t i = 0; i < n; ++i) p->bits[i] = bits_for(joins[i].n_build, c.part_rows);
    // chunks of one scatter tile: every tile writes its runs at offsets of its own
    const uint32_t mb = *std::max_element(p->bits.begin(), p->bits.end());
    const uint32_t crR = scatter_tile_rows((int)c.key_bytes, true, mb);
    const uint32_t crS = scatter_tile_rows((int)c.key_bytes, c.probe_payload != 0, mb);
    for (uint32_t i = 0; i < n; ++i)
        p->lr.add(joins[i].d_rkey, joins[i].d_rpay, joins[i].n_build, p->bits[i], 0, crR);
    const uint32_t nbc = (uint32_t)p->lr.chunks.size();
    for (uint32_t i = 0; i < n; ++i)
        p->ls.add(joins[i].d_skey, c.probe_payload ? joins[i].d_spay : nullptr, joins[i].n_probe, p->bits[i], nbc,
                  crS);
    // probe segments index the combined segment table after the build ones,
    // and their histograms / bounds follow the build side's
    for (auto& ch : p->ls.chunks) ch.seg += n;
    for (auto& s : p->ls.segs) {
        s.hist += p->lr.hist;
        s.bnd += p->lr.bnd;
    }
    std::vector<Job> jobs(n);
    std::vector<uint32_t> item_job;
    for (uint32_t i : order) {
        Job& jb = jobs[i];
        jb.rdst = p->lr.segs[i].dst;
        jb.sdst = p->ls.segs[i].dst;
        jb.rbnd = p->lr.segs[i].bnd;
        jb.sbnd = p->ls.segs[i].bnd - p->lr.bnd;  // the join kernel gets the probe bounds array
        jb.bits = p->bits[i];
        jb.item0 = (uint32_t)item_job.size();
        item_job.insert(item_job.end(), 1ull << p->bits[i], i);
    }
    p->n_items = (uint32_t)item_job.size();
    int rc;
    if ((rc = join_grid((int)c.key_bytes, c.probe_payload != 0, (int)c.variant, &p->grid)) != PGPU_OK)
        return cleanup(rc);
    p->grid = (int)std::min<uint64_t>((uint64_t)p->grid, std::max<uint32_t>(p->n_items, 1));
    std::vector<Seg> segs = p->lr.segs;
    segs.insert(segs.end(), p->ls.segs.begin(), p->ls.segs.end());
    std::vector<Chunk> chunks = p->lr.chunks;
    chunks.insert(chunks.end(), p->ls.chunks.begin(), p->ls.chunks.end());
    const uint64_t kb = c.key_bytes;
    const uint64_t esR = 2 * kb, esS = c.probe_payload ? 2 * kb : kb;
    struct {
        void** ptr;
        uint64_t bytes;
    } a[] = {
        {(void**)&p->segs, segs.size() * sizeof(Seg)},
        {(void**)&p->chunks, chunks.size() * sizeof(Chunk)},
        {(void**)&p->H, (p->lr.hist + p->ls.hist) * 4},
        {(void**)&p->bnd, (p->lr.bnd + p->ls.bnd) * 4},
        {&p->RP, p->lr.rows * esR},
        {&p->SP, p->ls.rows * esS},
        {(void**)&p->jobs, n * sizeof(Job)},
        {(void**)&p->item_job, item_job.size() * 4},
        {(void**)&p->res, (4ull * n + 2) * 8},
        {(void**)&p->redo, c.variant == 3 ? (uint64_t)p->n_items * 4 : 0},
    };
    for (auto& x : a)
        if ((rc = balloc(p, x.bytes, x.ptr)) != PGPU_OK) return cleanup(rc);
    if (hipHostMalloc((void**)&p->h_res, 4ull * n * 8, hipHostMallocDefault) != hipSuccess)
        return cleanup(fail(PGPU_ENOMEM, "hipHostMalloc for the batch results failed"));
    hipStream_t st = ctx->stream;
    if (hipMemcpyAsync(p->segs, segs.data(), segs.size() * sizeof(Seg), hipMemcpyHostToDevice, st) != hipSuccess ||
        (!chunks.empty() && hipMemcpyAsync(p->chunks, chunks.data(), chunks.size() * sizeof(Chunk),
                                           hipMemcpyHostToDevice, st) != hipSuccess) ||
        hipMemcpyAsync(p->jobs, jobs.data(), n * sizeof(Job), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(p->item_job, item_job.data(), item_job.size() * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return cleanup(fail(PGPU_EDEVICE, "uploading the batch tables failed"));
    if (c.graph) {
        if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return cleanup(fail(PGPU_EDEVICE, "hipStreamBeginCapture failed"));
        rc = enqueue(p, false);
        const hipError_t e = hipStreamEndCapture(st, &p->graph);
        if (rc != PGPU_OK) return cleanup(rc);
        if (e != hipSuccess) return cleanup(fail(PGPU_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e)));
        if (hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0) != hipSuccess)
            return cleanup(fail(PGPU_EDEVICE, "hipGraphInstantiate failed"));
    }
    *out = p;
    return PGPU_OK;
}

int pgpu_batch_plan_destroy(pgpu_batch_plan* p) {
    if (!p) return PGPU_OK;
    (void)hipSetDevice(p->ctx->device);
    (void)hipStreamSynchronize(p->ctx->stream);
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    delete p->timer;
    for (void* x : p->owned) (void)ctx_free(p->ctx, x);
    if (p->h_res) (void)hipHostFree(p->h_res);
    delete p;
    return PGPU_OK;
}

int pgpu_batch_run_async(pgpu_batch_plan* p) {
    if (!p) return fail(PGPU_EINVAL, "pgpu_batch_run_async: null plan");
    PGPU_TRY(ctx_activate(p->ctx));
    if (p->exec && !p->timer->enabled()) {
        PGPU_HIP(hipGraphLaunch(p->exec, p->ctx->stream));
    } else {
        PGPU_TRY(enqueue(p, p->timer->enabled()));
    }
    p->pending = true;
    return PGPU_OK;
}

int pgpu_batch_wait(pgpu_batch_plan* p, pgpu_join_result* out) {
    if (!p) return fail(PGPU_EINVAL, "pgpu_batch_wait: null plan");
    if (!p->pending) return fail(PGPU_ESTATE, "pgpu_batch_wait: nothing enqueued");
    PGPU_TRY(ctx_activate(p->ctx));
    PGPU_HIP(hipStreamSynchronize(p->ctx->stream));
    p->pending = false;
    if (out) {
        for (uint32_t i = 0; i < p->n; ++i) {
            memset(&out[i], 0, sizeof out[i]);
            out[i].matches = p->h_res[4 * i];
            out[i].sum_build = p->h_res[4 * i + 1];
            out[i].sum_probe = p->h_res[4 * i + 2];
            out[i].radix_bits = p->bits[i];
            out[i].lds_passes = std::max<uint32_t>(1, (uint32_t)p->h_res[4 * i + 3]);
        }
    }
    return PGPU_OK;
}

int pgpu_batch_info(pgpu_batch_plan* p, pgpu_batch_info_t* info) {
    if (!p || !info) return fail(PGPU_EINVAL, "pgpu_batch_info: null argument");
    memset(info, 0, sizeof *info);
    info->cfg = p->cfg;
    info->items = p->n_items;
    info->grid = (uint32_t)p->grid;
    info->build_rows = p->lr.rows;
    info->probe_rows = p->ls.rows;
    info->chunks = p->lr.chunks.size() + p->ls.chunks.size();
    info->graph = p->exec != nullptr;
    return PGPU_OK;
}

int pgpu_batch_timing(pgpu_batch_plan* p, int enable) {
    if (!p) return fail(PGPU_EINVAL, "pgpu_batch_timing: null plan");
    PGPU_TRY(ctx_activate(p->ctx));
    p->timer->enable(enable != 0);
    return PGPU_OK;
}

int pgpu_batch_times(pgpu_batch_plan* p, pgpu_kernel_time* out, int cap, int* n) {
    if (!p || !n || (cap > 0 && !out)) return fail(PGPU_EINVAL, "pgpu_batch_times: null argument");
    PGPU_TRY(ctx_activate(p->ctx));
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

int pgpu_batch_autotune(pgpu_ctx* ctx, const pgpu_batch_cfg* base, const pgpu_batch_join* joins, uint32_t n,
                        int reps, pgpu_batch_cfg* best, pgpu_tune_entry* entries, int cap, int* n_entries) {
    if (!ctx || !base || !best || (cap > 0 && !entries)) return fail(PGPU_EINVAL, "pgpu_batch_autotune: null argument");
    if (reps < 1) reps = 1;
    struct Cand {
        uint32_t variant, part_rows;
    };
    const Cand all[] = {{1, 1024}, {1, 1536}, {2, 2048}, {2, 3072}, {2, 4096}, {3, 4096}};
    std::vector<Cand> cands;
    for (const Cand& cd : all)
        if (variant_ok((int)cd.variant, (int)base->key_bytes)) cands.push_back(cd);
    PGPU_TRY(ctx_activate(ctx));
    hipEvent_t e0, e1;
    PGPU_HIP(hipEventCreate(&e0));
    PGPU_HIP(hipEventCreate(&e1));
    double best_ms = -1;
    int count = 0, rc = PGPU_OK;
    for (const Cand& cd : cands) {
        pgpu_batch_cfg c = *base;
        c.variant = cd.variant;
        c.part_rows = cd.part_rows;
        c.graph = 1;
        pgpu_batch_plan* p = nullptr;
        if ((rc = pgpu_batch_plan_create(ctx, &c, joins, n, &p)) != PGPU_OK) break;
        double ms = 0;
        rc = pgpu_batch_run_async(p);  // warm-up
        if (rc == PGPU_OK) rc = pgpu_batch_wait(p, nullptr);
        if (rc == PGPU_OK && hipEventRecord(e0, ctx->stream) != hipSuccess) rc = fail(PGPU_EDEVICE, "event record");
        for (int r = 0; rc == PGPU_OK && r < reps; ++r) rc = pgpu_batch_run_async(p);
        if (rc == PGPU_OK && hipEventRecord(e1, ctx->stream) != hipSuccess) rc = fail(PGPU_EDEVICE, "event record");
        if (rc == PGPU_OK) rc = pgpu_batch_wait(p, nullptr);
        float f = 0;
        if (rc == PGPU_OK && hipEventElapsedTime(&f, e0, e1) != hipSuccess) rc = fail(PGPU_EDEVICE, "event time");
        ms = (double)f / reps;
        pgpu_batch_plan_destroy(p);
        if (rc != PGPU_OK) break;
        if (count < cap) entries[count] = pgpu_tune_entry{c.variant, c.part_rows, ms};
        ++count;
        if (best_ms < 0 || ms < best_ms) {
            best_ms = ms;
            *best = c;
            best->graph = base->graph;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (n_entries) *n_entries = count;
    return rc;
}

}  // extern "C"
