#include "registry.h"

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>

#include "operators.h"

namespace pythia {

namespace {
std::mutex g_mu;
struct Table {
    unsigned row_bytes = 0;
    std::vector<char> rows;
    uint64_t generation = 0;
};
std::map<std::string, Table> g_tables;
uint64_t g_generation = 0;
struct Resident {
    pgpu_ctx* ctx = nullptr;
    void* d = nullptr;
};
std::mutex g_res_mu;  // held while a source is copied; never taken under g_mu
std::map<std::pair<int, std::string>, Resident> g_resident;
// A device context outlives the queries that use it (an idle one is kept
// with refs == 0): its pooled streams, pinned words and device blocks serve
// the next query, so a repeated query allocates nothing.
struct Dev {
    pgpu_ctx* ctx = nullptr;
    int refs = 0;
    uint64_t budget = 0;  // as requested at creation
};
std::map<int, Dev> g_devs;
}  // namespace

void register_table(const std::string& name, const void* rows, uint64_t nrows, unsigned row_bytes) {
    std::lock_guard<std::mutex> g(g_mu);
    auto& t = g_tables[name];
    t.row_bytes = row_bytes;
    t.rows.assign((const char*)rows, (const char*)rows + nrows * row_bytes);
    t.generation = ++g_generation;
}

void clear_tables() {
    std::lock_guard<std::mutex> g(g_mu);
    g_tables.clear();
}

bool lookup_table(const std::string& name, unsigned row_bytes, std::vector<char>& out, uint64_t* generation) {
    std::lock_guard<std::mutex> g(g_mu);
    auto it = g_tables.find(name);
    if (it == g_tables.end() || it->second.row_bytes != row_bytes) return false;
    out = it->second.rows;
    if (generation) *generation = it->second.generation;
    return true;
}

void check(int rc, const char* what) {
    if (rc != PGPU_OK) throw QueryExecutionError(std::string(what) + ": " + pgpu_last_error());
}

pgpu_ctx* acquire_device(int device, uint64_t hbm_budget) {
    std::lock_guard<std::mutex> g(g_mu);
    Dev& d = g_devs[device];
    if (d.ctx && d.budget != hbm_budget) {
        if (d.refs == 0) {  // idle, other budget: start afresh
            pgpu_ctx_destroy(d.ctx);
            d.ctx = nullptr;
        } else {
            // The live context (held by a running query or by device copies
            // kept across queries) enforces the budget it was created with;
            // silently running under it would ignore the one asked for.
            throw QueryExecutionError("gpu.hbmbudget " + std::to_string(hbm_budget) + " on device " +
                                      std::to_string(device) + ": the device is in use under an HBM budget of " +
                                      std::to_string(d.budget) +
                                      " bytes (0 = all free memory); finish those queries or release the "
                                      "resident device copies (pythia_clear_resident) first");
        }
    }
    if (!d.ctx) {
        pgpu_ctx* c = nullptr;
        const int rc = pgpu_ctx_create(device, hbm_budget, &c);
        if (rc != PGPU_OK) g_devs.erase(device);
        check(rc, "pgpu_ctx_create");
        d.ctx = c;
        d.budget = hbm_budget;
    }
    ++d.refs;
    return d.ctx;
}

void release_device(pgpu_ctx* ctx) {
    if (!ctx) return;
    std::lock_guard<std::mutex> g(g_mu);
    for (auto& e : g_devs)
        if (e.second.ctx == ctx && e.second.refs > 0) {
            --e.second.refs;  // at zero the context stays, idle, for the next query
            return;
        }
}

const void* resident_rows(int device, uint64_t hbm_budget, const std::string& key, const char* rows, uint64_t bytes) {
    std::lock_guard<std::mutex> g(g_res_mu);
    auto it = g_resident.find({device, key});
    if (it != g_resident.end()) return it->second.d;
    Resident r;
    r.ctx = acquire_device(device, hbm_budget);
    int rc = pgpu_alloc(r.ctx, std::max<uint64_t>(bytes, 256), &r.d);
    if (rc == PGPU_OK && bytes) rc = pgpu_h2d(r.ctx, r.d, rows, bytes);
    if (rc != PGPU_OK) {
        const std::string msg = std::string("resident copy of a scan source: ") + pgpu_last_error();
        if (r.d) pgpu_free(r.ctx, r.d);
        release_device(r.ctx);
        throw QueryExecutionError(msg);
    }
    g_resident[{device, key}] = r;
    return r.d;
}

namespace {
struct ColKey {
    int device;
    std::string key;
    unsigned offset, bytes;
    bool operator<(const ColKey& o) const {
        return std::tie(device, key, offset, bytes) < std::tie(o.device, o.key, o.offset, o.bytes);
    }
};
std::map<ColKey, Resident> g_columns;  // under g_res_mu

template <typename P>
struct PlanEntry {
    std::mutex mu;  // held while a query uses the plan
    pgpu_ctx* ctx = nullptr;
    P* plan = nullptr;
    uint64_t a = 0, b = 0, c = 0;  // capacities the plan was created with
};
std::mutex g_plan_mu;  // guards the maps (not the entries)
std::map<std::tuple<int, uint32_t, uint32_t>, std::unique_ptr<PlanEntry<pgpu_join_plan>>> g_join_plans;
std::map<int, std::unique_ptr<PlanEntry<pgpu_groupjoin>>> g_groupjoins;

struct Pinned {
    int device;
    pgpu_ctx* ctx;
    uint64_t bytes;
    bool busy;
};
std::mutex g_pin_mu;
std::map<void*, Pinned> g_pinned;
}  // namespace

const void* resident_column(int device, uint64_t hbm_budget, const std::string& key, const void* d_rows, uint64_t n,
                            unsigned row_bytes, unsigned offset, unsigned bytes) {
    std::lock_guard<std::mutex> g(g_res_mu);
    const ColKey ck{device, key, offset, bytes};
    auto it = g_columns.find(ck);
    if (it != g_columns.end()) return it->second.d;
    Resident r;
    r.ctx = acquire_device(device, hbm_budget);
    int rc = pgpu_alloc(r.ctx, std::max<uint64_t>(n * bytes, 256), &r.d);
    if (rc == PGPU_OK) rc = pgpu_rows_column(r.ctx, d_rows, n, row_bytes, offset, bytes, r.d);
    if (rc != PGPU_OK) {
        const std::string msg = std::string("resident column of a scan source: ") + pgpu_last_error();
        if (r.d) pgpu_free(r.ctx, r.d);
        release_device(r.ctx);
        throw QueryExecutionError(msg);
    }
    g_columns[ck] = r;
    return r.d;
}

void with_join_plan(int device, uint64_t hbm_budget, const pgpu_join_cfg& cfg, uint64_t max_build,
                    uint64_t max_probe, const std::function<void(pgpu_join_plan*)>& fn) {
    PlanEntry<pgpu_join_plan>* e;
    {
        std::lock_guard<std::mutex> g(g_plan_mu);
        auto& slot = g_join_plans[std::make_tuple(device, cfg.key_bytes, cfg.probe_payload)];
        if (!slot) slot.reset(new PlanEntry<pgpu_join_plan>());
        e = slot.get();
    }
    std::lock_guard<std::mutex> lk(e->mu);
    if (!e->plan || e->a < max_build || e->b < max_probe) {
        if (e->plan) pgpu_join_plan_destroy(e->plan);
        e->plan = nullptr;
        if (!e->ctx) e->ctx = acquire_device(device, hbm_budget);
        // grow geometrically so a series of slightly larger inputs does not
        // recreate the plan every time
        const uint64_t nb = std::max(max_build, e->a + e->a / 2), np = std::max(max_probe, e->b + e->b / 2);
        check(pgpu_join_plan_create(e->ctx, &cfg, nb, np, &e->plan), "pgpu_join_plan_create");
        e->a = nb;
        e->b = np;
    }
    fn(e->plan);
}

void with_groupjoin(int device, uint64_t hbm_budget, uint64_t max_dim, uint64_t max_fact, uint64_t max_groups,
                    const std::function<void(pgpu_groupjoin*)>& fn) {
    PlanEntry<pgpu_groupjoin>* e;
    {
        std::lock_guard<std::mutex> g(g_plan_mu);
        auto& slot = g_groupjoins[device];
        if (!slot) slot.reset(new PlanEntry<pgpu_groupjoin>());
        e = slot.get();
    }
    std::lock_guard<std::mutex> lk(e->mu);
    if (!e->plan || e->a < max_dim || e->b < max_fact || e->c < max_groups) {
        if (e->plan) pgpu_groupjoin_destroy(e->plan);
        e->plan = nullptr;
        if (!e->ctx) e->ctx = acquire_device(device, hbm_budget);
        const uint64_t nd = std::max(max_dim, e->a), nf = std::max(max_fact, e->b), ng = std::max(max_groups, e->c);
        check(pgpu_groupjoin_create(e->ctx, nd, nf, ng, &e->plan), "pgpu_groupjoin_create");
        e->a = nd;
        e->b = nf;
        e->c = ng;
    }
    fn(e->plan);
}

void* acquire_pinned(int device, uint64_t hbm_budget, uint64_t bytes) {
    {
        std::lock_guard<std::mutex> g(g_pin_mu);
        void* best = nullptr;
        for (auto& kv : g_pinned)
            if (!kv.second.busy && kv.second.device == device && kv.second.bytes >= bytes &&
                (!best || kv.second.bytes < g_pinned[best].bytes))
                best = kv.first;
        if (best) {
            g_pinned[best].busy = true;
            return best;
        }
    }
    pgpu_ctx* ctx = acquire_device(device, hbm_budget);
    void* p = nullptr;
    const int rc = pgpu_host_alloc(ctx, bytes, &p);
    if (rc != PGPU_OK) {
        release_device(ctx);
        check(rc, "pgpu_host_alloc");
    }
    std::lock_guard<std::mutex> g(g_pin_mu);
    g_pinned[p] = Pinned{device, ctx, bytes, true};
    return p;
}

void release_pinned(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> g(g_pin_mu);
    auto it = g_pinned.find(p);
    if (it != g_pinned.end()) it->second.busy = false;
}

void clear_resident() {
    {
        std::lock_guard<std::mutex> g(g_res_mu);
        for (auto& e : g_resident) {
            pgpu_free(e.second.ctx, e.second.d);
            release_device(e.second.ctx);
        }
        g_resident.clear();
        for (auto& e : g_columns) {
            pgpu_free(e.second.ctx, e.second.d);
            release_device(e.second.ctx);
        }
        g_columns.clear();
    }
    {
        std::lock_guard<std::mutex> g(g_plan_mu);
        for (auto& kv : g_join_plans) {
            std::lock_guard<std::mutex> lk(kv.second->mu);
            if (kv.second->plan) pgpu_join_plan_destroy(kv.second->plan);
            if (kv.second->ctx) release_device(kv.second->ctx);
        }
        g_join_plans.clear();
        for (auto& kv : g_groupjoins) {
            std::lock_guard<std::mutex> lk(kv.second->mu);
            if (kv.second->plan) pgpu_groupjoin_destroy(kv.second->plan);
            if (kv.second->ctx) release_device(kv.second->ctx);
        }
        g_groupjoins.clear();
    }
    std::lock_guard<std::mutex> g(g_pin_mu);
    for (auto it = g_pinned.begin(); it != g_pinned.end();) {
        if (it->second.busy) {  // still held by a running query: freed by a later call
            ++it;
            continue;
        }
        pgpu_host_free(it->second.ctx, it->first);
        release_device(it->second.ctx);
        it = g_pinned.erase(it);
    }
}

}  // namespace pythia
