#include "query.h"

#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "gpu_operators.h"
#include "pythia_host.h"
#include "registry.h"

namespace {

// sanitycheck (query.cpp:37-60): the branch exists, names a node, and that
// node has a type.
void sanitycheck(libconfig::Config& cfg, libconfig::Setting& node, const std::string& branch) {
    if (!node.exists(branch)) throw MissingParameterException("Cannot find `" + branch + "' attribute in query subtree.");
    if (!node[branch].exists("name")) throw MissingParameterException("Cannot find `name' attribute in query subtree.");
    std::string name = node[branch]["name"];
    if (!cfg.getRoot().exists(name)) throw MissingParameterException("Cannot find description for node `" + name + "'.");
    if (!cfg.getRoot()[name].exists("type"))
        throw MissingParameterException("Cannot find mandatory `type' parameter in description for node `" + name +
                                        "'.");
}

// constructsubtree (query.cpp:62-233): pick the class by type string, wire the
// children, then init this node.
void constructsubtree(libconfig::Config& cfg, libconfig::Setting& node, Operator** root) {
    std::string name = node["name"];
    std::string type = cfg.getRoot()[name]["type"];
    if (type == "scan" || type == "parallelscan" || type == "partitionedscan") {
        *root = new ParallelScanOp();
    } else if (type == "memscan") {
        *root = new MemoryScanOp();
    } else if (type == "gpuhashjoin") {
        auto* op = new GpuHashJoinOp();
        *root = op;
        sanitycheck(cfg, node, "build");
        constructsubtree(cfg, node["build"], &op->buildOp);
        sanitycheck(cfg, node, "probe");
        constructsubtree(cfg, node["probe"], &op->probeOp);
    } else if (type == "hashjoin" || type == "aggregate_sum" || type == "aggregate_count" || type == "partition") {
        throw MissingParameterException("`" + type + "' is the reference engine's CPU operator; this engine runs `gpu" +
                                        type + "'.");
    } else {
        SingleInputOp* op = nullptr;
        if (type == "gpuaggregate_sum")
            op = new GpuAggregateSum();
        else if (type == "gpuaggregate_count")
            op = new GpuAggregateCount();
        else if (type == "merge")
            op = new MergeOp();
        else if (type == "checker_callstate")
            op = new CallStateChecker();
        else if (type == "gpupartition")
            op = new GpuPartitionOp();
        else if (type == "threadidprepend")
            op = new ThreadIdPrependOp();
        else
            throw MissingParameterException("`" + type + "' is neither a built-in nor a user-defined type.");
        *root = op;
        sanitycheck(cfg, node, "input");
        constructsubtree(cfg, node["input"], &op->nextOp);
    }
    (*root)->init(cfg, cfg.lookup(name));
}

// Fused forms taken and join output bytes pulled to the host, over the tree.
void gpu_report(Operator* op, uint32_t* fused, uint64_t* d2h) {
    if (!op) return;
    if (auto* a = dynamic_cast<GpuAggregate*>(op)) *fused |= (uint32_t)a->fused();
    if (auto* j = dynamic_cast<GpuHashJoinOp*>(op)) *d2h += j->pulledBytes();
    if (auto* d = dynamic_cast<DualInputOp*>(op)) {
        gpu_report(d->buildOp, fused, d2h);
        gpu_report(d->probeOp, fused, d2h);
    } else if (auto* s1 = dynamic_cast<SingleInputOp*>(op)) {
        gpu_report(s1->nextOp, fused, d2h);
    }
}

// "<what>", followed by the failing merge worker's message when there is one.
std::string failure(Operator* root, const char* what) {
    std::string m = what;
    if (auto* merge = dynamic_cast<MergeOp*>(root)) {
        const std::string e = merge->firstError();
        if (!e.empty()) m += ": " + e;
    }
    return m;
}

thread_local std::string g_err;

int set_err(const std::string& m) {
    g_err = m;
    return -1;
}

}  // namespace

void Query::create(libconfig::Config& cfg) {
    if (!cfg.getRoot().exists("treeroot")) throw MissingParameterException("Cannot find `treeroot' in the plan.");
    libconfig::Setting& tr = cfg.getRoot()["treeroot"];
    if (!tr.exists("name")) throw MissingParameterException("Cannot find `name' attribute in query subtree.");
    constructsubtree(cfg, tr, &tree);
}

static std::string describe(const Schema& s) {
    std::string d;
    for (unsigned i = 0; i < s.columns(); ++i) {
        if (i) d += ",";
        switch (s.getColumnType(i)) {
            case CT_INTEGER: d += "int"; break;
            case CT_LONG: d += "long"; break;
            case CT_DECIMAL: d += "dec"; break;
            case CT_DATE: d += "date"; break;
            case CT_CHAR: d += "char(" + std::to_string(s.getColumnWidth(i) - 1) + ")"; break;
            default: d += "?";
        }
    }
    return d;
}

namespace {

// The driver copies every output page into its result buffer before it asks
// for the next page (a page stays valid only until then).  Pages of at least
// kMinSplit bytes are copied by the driver thread and helper threads
// together: on a 16-thread EPYC share one thread moves ~26 GB/s out of the
// pinned pages the GPU operators emit, eight ~45 GB/s (1.5 M group rows,
// 18 MB, in 0.4 instead of 0.7 ms).  Helpers are started on the first large
// page, spin briefly for the next one, then sleep.
class PageCopier {
  public:
    ~PageCopier() {
        {
            std::lock_guard<std::mutex> g(mu_);
            quit_ = true;
            gen_.fetch_add(1, std::memory_order_release);
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }
    void copy(char* dst, const char* src, size_t n) {
        if (n < kMinSplit || helpers_ == 0) {
            memcpy(dst, src, n);
            return;
        }
        std::lock_guard<std::mutex> one(call_mu_);
        if (th_.empty()) {
            owner_ = getpid();
            for (int h = 0; h < helpers_; ++h) th_.emplace_back(&PageCopier::run, this, h);
        } else if (owner_ != getpid()) {  // a forked child has no helper threads
            memcpy(dst, src, n);
            return;
        }
        dst_ = dst;
        src_ = src;
        n_ = n;
        left_.store(helpers_, std::memory_order_relaxed);
        {
            std::lock_guard<std::mutex> g(mu_);
            gen_.fetch_add(1, std::memory_order_release);
        }
        cv_.notify_all();
        size_t b, e;
        piece(0, &b, &e);
        memcpy(dst + b, src + b, e - b);
        while (left_.load(std::memory_order_acquire) != 0) std::this_thread::yield();
    }

  private:
    static constexpr size_t kMinSplit = 256u << 10;
    static constexpr int kSpin = 4096;  // polls of the page counter before a helper sleeps
    // Up to 7 helpers, leaving half the host threads to the operators.
    const int helpers_ = std::max(0, std::min(7, (int)std::thread::hardware_concurrency() / 2 - 1));
    void piece(int i, size_t* b, size_t* e) const {
        const size_t parts = (size_t)helpers_ + 1;
        const size_t step = ((n_ + parts - 1) / parts + 63) & ~(size_t)63;
        *b = std::min(n_, step * (size_t)i);
        *e = std::min(n_, step * (size_t)(i + 1));
    }
    void run(int h) {
        uint64_t seen = 0;
        while (true) {
            uint64_t g = gen_.load(std::memory_order_acquire);
            for (int spin = 0; g == seen && spin < kSpin; ++spin) {
                __builtin_ia32_pause();
                g = gen_.load(std::memory_order_acquire);
            }
            if (g == seen) {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return gen_.load(std::memory_order_acquire) != seen; });
                g = gen_.load(std::memory_order_acquire);
            }
            seen = g;
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (quit_) return;
            }
            size_t b, e;
            piece(h + 1, &b, &e);
            memcpy(dst_ + b, src_ + b, e - b);
            left_.fetch_sub(1, std::memory_order_release);
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_, call_mu_;
    std::condition_variable cv_;
    std::atomic<uint64_t> gen_{0};
    std::atomic<int> left_{0};
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    size_t n_ = 0;
    bool quit_ = false;
    pid_t owner_ = 0;
};

PageCopier g_copier;

// The result rows of a query: grown by doubling and kept between queries on
// a thread, so a repeated query writes into memory already mapped.
struct RowBuf {
    char* p = nullptr;
    size_t len = 0, cap = 0;
    ~RowBuf() { free(p); }
    char* grow(size_t n) {
        if (len + n > cap) {
            const size_t c = std::max(cap * 2, len + n);
            char* q = (char*)realloc(p, c);
            if (!q) throw std::bad_alloc();
            p = q;
            cap = c;
        }
        char* d = p + len;
        len += n;
        return d;
    }
};

}  // namespace

extern "C" {

const char* pythia_last_error(void) { return g_err.c_str(); }

int pythia_register_table(const char* name, const void* rows, uint64_t nrows, uint32_t row_bytes) {
    if (!name || (nrows && !rows) || !row_bytes) return set_err("pythia_register_table: bad argument");
    pythia::register_table(name, rows, nrows, row_bytes);
    return 0;
}

int pythia_clear_tables(void) {
    pythia::clear_tables();
    return 0;
}

int pythia_clear_resident(void) {
    pythia::clear_resident();
    return 0;
}

// The executequery driver loop (drivers/executequery.cpp:32-101): create,
// threadInit, scanStart, getNext until Finished, scanStop, threadClose,
// destroy; every output row is collected.
int pythia_run_query(const char* plan_text, pythia_result* out) {
    if (!plan_text || !out) return set_err("pythia_run_query: null argument");
    memset(out, 0, sizeof *out);
    Query q;
    int rc = 0;
    static thread_local RowBuf rows;
    rows.len = 0;
    try {
        libconfig::Config cfg;
        cfg.readString(plan_text);
        q.create(cfg);
        snprintf(out->schema, sizeof out->schema, "%s", describe(q.getOutSchema()).c_str());
        out->row_bytes = q.getOutSchema().getTupleSize();
        q.threadInit();
        auto t0 = std::chrono::steady_clock::now();
        if (q.scanStart() == Operator::Error) throw QueryExecutionError(failure(q.tree, "scanStart returned Error"));
        Operator::GetNextResultT r{Operator::Ready, nullptr};
        uint64_t calls = 0;
        while (r.first == Operator::Ready) {
            r = q.getNext();
            ++calls;
            if (r.first == Operator::Error) throw QueryExecutionError(failure(q.tree, "getNext returned Error"));
            const size_t nb = r.second->getUsedSpace();
            if (nb) g_copier.copy(rows.grow(nb), (const char*)r.second->data(), nb);
        }
        if (q.scanStop() == Operator::Error) throw QueryExecutionError(failure(q.tree, "scanStop returned Error"));
        auto t1 = std::chrono::steady_clock::now();
        out->seconds = std::chrono::duration<double>(t1 - t0).count();
        out->getnext_calls = calls;
        gpu_report(q.tree, &out->gpu_fused, &out->join_d2h_bytes);
        q.threadClose();
    } catch (const std::exception& e) {
        rc = set_err(e.what());
    }
    if (q.tree) {
        try {
            q.destroy();
        } catch (const std::exception& e) {
            if (!rc) rc = set_err(e.what());
        }
    }
    if (rc) return rc;
    out->nrows = out->row_bytes ? rows.len / out->row_bytes : 0;
    out->rows = malloc(rows.len ? rows.len : 1);
    if (!out->rows) return set_err("pythia_run_query: out of host memory for the result rows");
    if (rows.len) memcpy(out->rows, rows.p, rows.len);
    return 0;
}

void pythia_free_result(pythia_result* r) {
    if (r && r->rows) free(r->rows);
    if (r) r->rows = nullptr;
}

}  // extern "C"
