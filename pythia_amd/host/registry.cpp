#include "registry.h"

#include <map>
#include <mutex>

#include "operators.h"

namespace pythia {

namespace {
std::mutex g_mu;
std::map<std::string, std::pair<unsigned, std::vector<char>>> g_tables;
struct Dev {
    pgpu_ctx* ctx = nullptr;
    int refs = 0;
};
std::map<int, Dev> g_devs;
}  // namespace

void register_table(const std::string& name, const void* rows, uint64_t nrows, unsigned row_bytes) {
    std::lock_guard<std::mutex> g(g_mu);
    auto& t = g_tables[name];
    t.first = row_bytes;
    t.second.assign((const char*)rows, (const char*)rows + nrows * row_bytes);
}

void clear_tables() {
    std::lock_guard<std::mutex> g(g_mu);
    g_tables.clear();
}

bool lookup_table(const std::string& name, unsigned row_bytes, std::vector<char>& out) {
    std::lock_guard<std::mutex> g(g_mu);
    auto it = g_tables.find(name);
    if (it == g_tables.end() || it->second.first != row_bytes) return false;
    out = it->second.second;
    return true;
}

void check(int rc, const char* what) {
    if (rc != PGPU_OK) throw QueryExecutionError(std::string(what) + ": " + pgpu_last_error());
}

pgpu_ctx* acquire_device(int device, uint64_t hbm_budget) {
    std::lock_guard<std::mutex> g(g_mu);
    Dev& d = g_devs[device];
    if (!d.ctx) check(pgpu_ctx_create(device, hbm_budget, &d.ctx), "pgpu_ctx_create");
    ++d.refs;
    return d.ctx;
}

void release_device(pgpu_ctx* ctx) {
    if (!ctx) return;
    std::lock_guard<std::mutex> g(g_mu);
    for (auto it = g_devs.begin(); it != g_devs.end(); ++it) {
        if (it->second.ctx == ctx && --it->second.refs == 0) {
            pgpu_ctx_destroy(ctx);
            g_devs.erase(it);
            return;
        }
    }
}

}  // namespace pythia
