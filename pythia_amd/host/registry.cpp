#include "registry.h"

#include <algorithm>
#include <map>
#include <mutex>

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
    if (d.ctx && d.refs == 0 && d.budget != hbm_budget) {  // idle, other budget: start afresh
        pgpu_ctx_destroy(d.ctx);
        d.ctx = nullptr;
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

void clear_resident() {
    std::lock_guard<std::mutex> g(g_res_mu);
    for (auto& e : g_resident) {
        pgpu_free(e.second.ctx, e.second.d);
        release_device(e.second.ctx);
    }
    g_resident.clear();
}

}  // namespace pythia
