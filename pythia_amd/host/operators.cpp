// Operator base, lifecycle visitors, scans, merge and the call-state checker
// of the host engine the GPU operators plug into.
#include "operators.h"

#include <sys/stat.h>

#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <thread>

#include "registry.h"

TupleBuffer EmptyPage(nullptr, 0, nullptr, 1);

void Operator::init(libconfig::Config& root, libconfig::Setting& node) {
    // operators/base.cpp:37-40: every operator reads the global page size.
    buffsize = (unsigned int)(int)root.getRoot()["buffsize"];
}

void SingleInputOp::accept(Visitor* v) { v->visit(this); }
void DualInputOp::accept(Visitor* v) { v->visit(this); }
void ZeroInputOp::accept(Visitor* v) { v->visit(this); }
void MergeOp::accept(Visitor* v) { v->visit(this); }

// ------------------------------------------------------------ visitors
void ThreadInitVisitor::visit(SingleInputOp* op) {
    op->threadInit(threadid);
    op->nextOp->accept(this);
}
void ThreadInitVisitor::visit(DualInputOp* op) {
    op->threadInit(threadid);
    op->buildOp->accept(this);
    op->probeOp->accept(this);
}
void ThreadInitVisitor::visit(ZeroInputOp* op) { op->threadInit(threadid); }
// MergeOp initialises its subtree once per worker itself.
void ThreadInitVisitor::visit(MergeOp* op) { op->threadInit(threadid); }

void ThreadCloseVisitor::visit(SingleInputOp* op) {
    op->threadClose(threadid);
    op->nextOp->accept(this);
}
void ThreadCloseVisitor::visit(DualInputOp* op) {
    op->threadClose(threadid);
    op->buildOp->accept(this);
    op->probeOp->accept(this);
}
void ThreadCloseVisitor::visit(ZeroInputOp* op) { op->threadClose(threadid); }
void ThreadCloseVisitor::visit(MergeOp* op) { op->threadClose(threadid); }

// Destroy / free tolerate a partially built tree (a plan that failed in
// Query::create leaves null children).
void RecursiveDestroyVisitor::visit(SingleInputOp* op) {
    op->destroy();
    if (op->nextOp) op->nextOp->accept(this);
}
void RecursiveDestroyVisitor::visit(DualInputOp* op) {
    op->destroy();
    if (op->buildOp) op->buildOp->accept(this);
    if (op->probeOp) op->probeOp->accept(this);
}
void RecursiveDestroyVisitor::visit(ZeroInputOp* op) { op->destroy(); }
void RecursiveDestroyVisitor::visit(MergeOp* op) {
    op->destroy();
    if (op->nextOp) op->nextOp->accept(this);
}

void RecursiveFreeVisitor::visit(SingleInputOp* op) {
    if (op->nextOp) op->nextOp->accept(this);
    delete op;
}
void RecursiveFreeVisitor::visit(DualInputOp* op) {
    if (op->buildOp) op->buildOp->accept(this);
    if (op->probeOp) op->probeOp->accept(this);
    delete op;
}
void RecursiveFreeVisitor::visit(ZeroInputOp* op) { delete op; }
void RecursiveFreeVisitor::visit(MergeOp* op) {
    if (op->nextOp) op->nextOp->accept(this);
    delete op;
}

// ------------------------------------------------------------ scans
static void parse_text(const std::string& text, const Schema& s, const std::string& seps, std::vector<char>& out) {
    std::istringstream in(text);
    std::string line;
    const unsigned ts = s.getTupleSize();
    std::vector<char> row(ts);
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        std::vector<std::string> tok;
        size_t st = 0;
        while (true) {
            size_t e = line.find_first_of(seps, st);
            tok.push_back(line.substr(st, e == std::string::npos ? std::string::npos : e - st));
            if (e == std::string::npos) break;
            st = e + 1;
        }
        if (tok.size() < s.columns()) throw InvalidParameter("text row with too few fields: " + line);
        std::fill(row.begin(), row.end(), 0);
        for (unsigned c = 0; c < s.columns(); ++c) {
            char* d = row.data() + s.getColumnOffset(c);
            const char* t = tok[c].c_str();
            switch (s.getColumnType(c)) {
                case CT_INTEGER: {
                    CtInt v = atoi(t);
                    memcpy(d, &v, 4);
                    break;
                }
                case CT_LONG:
                case CT_DATE: {
                    CtLong v = atoll(t);
                    memcpy(d, &v, 8);
                    break;
                }
                case CT_DECIMAL: {
                    CtDecimal v = atof(t);
                    memcpy(d, &v, 8);
                    break;
                }
                case CT_CHAR: strncpy(d, t, s.getColumnWidth(c) - 1); break;
                default: throw IllegalSchemaDeclarationException();
            }
        }
        out.insert(out.end(), row.begin(), row.end());
    }
}

void ParallelScanOp::init(libconfig::Config& root, libconfig::Setting& node) {
    Operator::init(root, node);
    schema = Schema::create(node["schema"]);
    std::string path;
    root.getRoot().lookupValue("path", path);
    std::string type = "binary";
    node.lookupValue("filetype", type);
    std::string seps = "|";
    node.lookupValue("separators", seps);
    std::vector<std::string> files;
    if (node.exists("files")) {
        for (int i = 0; i < node["files"].getLength(); ++i) files.push_back(node["files"][i]);
    } else if (node.exists("file")) {
        files.push_back(node["file"]);
    } else {
        throw MissingParameterException("scan needs `file' or `files'");
    }
    // mapping: one list of thread ids per file (operators/parallelscan.cpp:
    // 42-88); without it every thread shares every file.
    std::vector<std::vector<int>> mapping;
    if (node.exists("mapping")) {
        libconfig::Setting& m = node["mapping"];
        if (m.getLength() == 0)
            throw MissingParameterException("ParallelScanOp `mapping' parameter cannot have a length of zero.");
        if ((size_t)m.getLength() != files.size()) throw InvalidParameter("mapping needs one thread list per file");
        for (int i = 0; i < m.getLength(); ++i) {
            mapping.emplace_back();
            for (int k = 0; k < m[i].getLength(); ++k) mapping.back().push_back((int)m[i][k]);
        }
    }
    std::vector<std::vector<char>> per_file;
    source_key_ = "scan:" + type + ":" + seps + ":";
    for (unsigned i = 0; i < schema.columns(); ++i)
        source_key_ += std::to_string((int)schema.getColumnType(i)) + "/" + std::to_string(schema.getColumnWidth(i)) + ",";
    for (auto& f : files) {
        struct stat sb {};
        if (stat((path + f).c_str(), &sb) == 0)
            source_key_ += "|" + path + f + ":" + std::to_string((long long)sb.st_size) + ":" +
                           std::to_string((long long)sb.st_mtim.tv_sec) + "." + std::to_string((long long)sb.st_mtim.tv_nsec);
        std::ifstream in(path + f, std::ios::binary);
        if (!in) throw MissingParameterException("cannot open " + path + f);
        std::stringstream ss;
        ss << in.rdbuf();
        std::string data = ss.str();
        per_file.emplace_back();
        if (type == "text") {
            parse_text(data, schema, seps, per_file.back());
        } else {
            if (data.size() % schema.getTupleSize())
                throw InvalidParameter("binary file size is not a multiple of the tuple size");
            per_file.back().assign(data.begin(), data.end());
        }
    }
    std::vector<char> all;
    if (mapping.empty())
        for (auto& r : per_file) all.insert(all.end(), r.begin(), r.end());
    setRows(std::move(all));
    if (mapping.empty()) return;
    // a mapped scan: group 0 stays empty for unmapped threads
    for (size_t g = 0; g < per_file.size(); ++g) {
        for (int tid : mapping[g]) {
            if (tid < 0 || tid >= MAX_THREADS || group_of_[tid] != 0) throw InvalidParameter("bad scan mapping");
            group_of_[tid] = (int)g + 1;
        }
        groups_.emplace_back(new Group());
        groups_.back()->rows = std::move(per_file[g]);
        groups_.back()->n = groups_.back()->rows.size() / schema.getTupleSize();
    }
}

bool ParallelScanOp::sharedRows(const char** rows, uint64_t* nrows, std::string* key) const {
    if (groups_.size() != 1) return false;
    *rows = groups_[0]->rows.data();
    *nrows = groups_[0]->n;
    *key = source_key_;
    return true;
}

void ParallelScanOp::setRows(std::vector<char>&& rows) {
    groups_.clear();
    groups_.emplace_back(new Group());
    groups_[0]->rows = std::move(rows);
    groups_[0]->n = groups_[0]->rows.size() / schema.getTupleSize();
    group_of_.assign(MAX_THREADS, 0);
    rows_per_page_ = std::max<uint64_t>(1, buffsize / schema.getTupleSize());
    pages_.assign(MAX_THREADS, nullptr);
}

void ParallelScanOp::threadInit(unsigned short tid) {
    // MergeOp runs every worker's threadInit before any scanStart, so the
    // first worker rewinds the shared cursors for this query run.
    if (tid == 0)
        for (auto& g : groups_) g->cursor = 0;
}

Operator::ResultCode ParallelScanOp::scanStart(unsigned short, Page*, Schema&) { return Ready; }

Operator::GetNextResultT ParallelScanOp::getNext(unsigned short tid) {
    delete pages_[tid];
    pages_[tid] = nullptr;
    Group& g = *groups_.at(group_of_.at(tid));
    const uint64_t start = g.cursor.fetch_add(rows_per_page_);
    if (start >= g.n) return {Finished, &EmptyPage};
    const uint64_t n = std::min(rows_per_page_, g.n - start);
    const unsigned ts = schema.getTupleSize();
    pages_[tid] = new Page(g.rows.data() + start * ts, n * ts, nullptr, ts);
    return {start + n >= g.n ? Finished : Ready, pages_[tid]};
}

void ParallelScanOp::threadClose(unsigned short tid) {
    delete pages_[tid];
    pages_[tid] = nullptr;
}

void ParallelScanOp::destroy() {
    for (auto& p : pages_) {
        delete p;
        p = nullptr;
    }
}

void MemoryScanOp::init(libconfig::Config& root, libconfig::Setting& node) {
    Operator::init(root, node);
    schema = Schema::create(node["schema"]);
    std::string name = node["table"];
    std::vector<char> rows;
    uint64_t generation = 0;
    if (!pythia::lookup_table(name, schema.getTupleSize(), rows, &generation))
        throw MissingParameterException("memscan: no registered table `" + name + "' with this row width");
    setRows(std::move(rows));
    source_key_ = "memscan:" + name + ":" + std::to_string(generation) + ":" + std::to_string(schema.getTupleSize());
}

// ------------------------------------------------------------ merge
MergeOp::~MergeOp() {
    for (size_t w = 0; w < workers_.size(); ++w) {
        if (workers_[w].th) {
            command((int)w, Exit);
            workers_[w].th->join();
            delete workers_[w].th;
            workers_[w].th = nullptr;
        }
    }
}

void MergeOp::init(libconfig::Config& root, libconfig::Setting& node) {
    Operator::init(root, node);
    schema = nextOp->getOutSchema();
    nthreads_ = (int)node["threads"];
    if (nthreads_ < 1 || nthreads_ > MAX_THREADS) throw InvalidParameter("merge: threads must be in 1..128");
}

void MergeOp::command(int w, Cmd c) {
    std::unique_lock<std::mutex> g(mu_);
    cv_.wait(g, [&] { return !workers_[w].busy; });
    workers_[w].cmd = c;
    workers_[w].busy = true;
    wcv_[w]->notify_one();  // only worker w waits on its own variable
}

void MergeOp::waitIdle(int w) {
    std::unique_lock<std::mutex> g(mu_);
    cv_.wait(g, [&] { return !workers_[w].busy; });
}

void MergeOp::workerMain(int w) {
    {
        ThreadInitVisitor v((unsigned short)w);
        try {
            nextOp->accept(&v);
        } catch (const std::exception& e) {
            std::lock_guard<std::mutex> g(mu_);
            workers_[w].status = Error;
            workers_[w].error = e.what();
        }
        std::lock_guard<std::mutex> g(mu_);
        workers_[w].busy = false;
        cv_.notify_all();
    }
    while (true) {
        Cmd c;
        {
            std::unique_lock<std::mutex> g(mu_);
            wcv_[w]->wait(g, [&] { return workers_[w].busy; });
            c = workers_[w].cmd;
        }
        if (c == Exit) {
            std::lock_guard<std::mutex> g(mu_);
            workers_[w].busy = false;
            cv_.notify_all();
            return;
        }
        GetNextResultT r{Error, &EmptyPage};
        ResultCode st = Ready;
        std::string err;
        try {
            if (c == DoScanStart) {
                Schema empty;
                st = nextOp->scanStart((unsigned short)w, nullptr, empty);
                if (st != Error) r = nextOp->getNext((unsigned short)w);
            } else if (c == DoGetNext) {
                r = nextOp->getNext((unsigned short)w);
            } else if (c == DoScanStop) {
                st = nextOp->scanStop((unsigned short)w);
            } else if (c == DoThreadClose) {
                ThreadCloseVisitor v((unsigned short)w);
                nextOp->accept(&v);
            }
        } catch (const std::exception& e) {
            st = Error;
            r = {Error, &EmptyPage};
            err = e.what();
        }
        std::lock_guard<std::mutex> g(mu_);
        Worker& wk = workers_[w];
        if (c == DoScanStart || c == DoGetNext) {
            if (st == Error) r.first = Error;
            wk.result = r;
            wk.has_result = true;
        }
        wk.status = st == Error ? Error : wk.status;
        if (!err.empty()) wk.error = err;
        wk.busy = false;
        cv_.notify_all();
    }
}

void MergeOp::threadInit(unsigned short) {
    workers_.assign(nthreads_, Worker());
    wcv_.clear();
    for (int w = 0; w < nthreads_; ++w) wcv_.emplace_back(new std::condition_variable());
    for (int w = 0; w < nthreads_; ++w) {
        workers_[w].busy = true;
        workers_[w].th = new std::thread(&MergeOp::workerMain, this, w);
    }
    for (int w = 0; w < nthreads_; ++w) waitIdle(w);
    for (int w = 0; w < nthreads_; ++w)
        if (workers_[w].status == Error) throw QueryExecutionError("merge worker threadInit: " + workers_[w].error);
}

Operator::ResultCode MergeOp::scanStart(unsigned short, Page*, Schema&) {
    remaining_ = nthreads_;
    prev_ = -1;
    finished_ = false;
    for (int w = 0; w < nthreads_; ++w) {
        workers_[w].finished = false;
        workers_[w].has_result = false;
        workers_[w].status = Ready;
    }
    for (int w = 0; w < nthreads_; ++w) command(w, DoScanStart);
    return Ready;
}

Operator::GetNextResultT MergeOp::getNext(unsigned short) {
    if (finished_) return {Finished, &EmptyPage};
    // The page handed out last time stays valid until now: let its producer
    // continue (operators/merge.cpp:447-461).
    if (prev_ >= 0 && !workers_[prev_].finished) command(prev_, DoGetNext);
    std::unique_lock<std::mutex> g(mu_);
    while (true) {
        for (int i = 0; i < nthreads_; ++i) {
            const int w = (prev_ + 1 + i + nthreads_) % nthreads_;
            Worker& wk = workers_[w];
            if (wk.busy || wk.finished || !wk.has_result) continue;
            prev_ = w;
            wk.has_result = false;
            GetNextResultT r = wk.result;
            if (r.first == Ready) return r;
            wk.finished = true;
            --remaining_;
            if (r.first == Error) {
                finished_ = remaining_ == 0;
                return r;
            }
            if (remaining_ > 0) return {Ready, r.second};
            finished_ = true;
            return r;
        }
        cv_.wait(g);
    }
}

Operator::ResultCode MergeOp::scanStop(unsigned short) {
    for (int w = 0; w < nthreads_; ++w) waitIdle(w);
    for (int w = 0; w < nthreads_; ++w) command(w, DoScanStop);
    ResultCode rc = Ready;
    for (int w = 0; w < nthreads_; ++w) {
        waitIdle(w);
        if (workers_[w].status == Error) rc = Error;
    }
    return rc;
}

std::string MergeOp::firstError() {
    std::lock_guard<std::mutex> g(mu_);
    for (auto& wk : workers_)
        if (!wk.error.empty()) return wk.error;
    return std::string();
}

void MergeOp::threadClose(unsigned short) {
    for (int w = 0; w < nthreads_; ++w) command(w, DoThreadClose);
    for (int w = 0; w < nthreads_; ++w) {
        command(w, Exit);
        waitIdle(w);
        workers_[w].th->join();
        delete workers_[w].th;
        workers_[w].th = nullptr;
    }
}

// ------------------------------------------------------------ call-state checker
void CallStateChecker::init(libconfig::Config& root, libconfig::Setting& node) {
    Operator::init(root, node);
    schema = nextOp->getOutSchema();
    state_.assign(MAX_THREADS, Init);
}

void CallStateChecker::bad(const char* what, unsigned short tid) {
    throw QueryExecutionError(std::string("call-state violation: ") + what + " on thread " + std::to_string(tid));
}

void CallStateChecker::threadInit(unsigned short tid) {
    if (state_[tid] != Init && state_[tid] != Closed) bad("threadInit twice", tid);
    state_[tid] = ThreadInit;
}

Operator::ResultCode CallStateChecker::scanStart(unsigned short tid, Page* p, Schema& s) {
    if (state_[tid] != ThreadInit && state_[tid] != ScanStopped) bad("scanStart before threadInit or twice", tid);
    state_[tid] = ScanStarted;
    return nextOp->scanStart(tid, p, s);
}

Operator::GetNextResultT CallStateChecker::getNext(unsigned short tid) {
    if (state_[tid] != ScanStarted && state_[tid] != GotFinished) bad("getNext outside a scan", tid);
    GetNextResultT r = nextOp->getNext(tid);
    if (r.first == Error) return r;
    if (r.second == nullptr) bad("null page", tid);
    if (state_[tid] == GotFinished) {
        if (r.first != Finished) bad("Ready after Finished", tid);
        if (r.second->getNumTuples() != 0) bad("non-empty page after Finished", tid);
    }
    if (r.first == Finished) state_[tid] = GotFinished;
    return r;
}

Operator::ResultCode CallStateChecker::scanStop(unsigned short tid) {
    if (state_[tid] != ScanStarted && state_[tid] != GotFinished) bad("scanStop without scanStart", tid);
    state_[tid] = ScanStopped;
    return nextOp->scanStop(tid);
}

void CallStateChecker::threadClose(unsigned short tid) {
    if (state_[tid] == Init || state_[tid] == Closed) bad("threadClose without threadInit", tid);
    state_[tid] = Closed;
}

// ------------------------------------------------------------------ threadidprepend
void ThreadIdPrependOp::init(libconfig::Config& root, libconfig::Setting& node) {
    Operator::init(root, node);
    schema.add(CT_INTEGER);
    Schema& in = nextOp->getOutSchema();
    for (unsigned i = 0; i < in.columns(); ++i) schema.add(in.get(i));
    st_.assign(MAX_THREADS, State());
}

void ThreadIdPrependOp::threadInit(unsigned short tid) {
    if (buffsize < schema.getTupleSize()) throw InvalidParameter("buffsize smaller than one output tuple");
    st_[tid] = State();
    st_[tid].out = new Page(buffsize, schema.getTupleSize());
}

Operator::ResultCode ThreadIdPrependOp::scanStart(unsigned short tid, Page* p, Schema& s) {
    State& st = st_[tid];
    st.in = nullptr;
    st.rc = Ready;
    st.off = 0;
    return nextOp->scanStart(tid, p, s);
}

Operator::GetNextResultT ThreadIdPrependOp::getNext(unsigned short tid) {
    State& st = st_[tid];
    Page* out = st.out;
    out->clear();
    const unsigned insz = nextOp->getOutSchema().getTupleSize();
    const CtInt id = tid;
    for (;;) {
        if (st.in) {
            while (const void* t = st.in->getTupleOffset(st.off)) {
                void* dst = out->allocateTuple();
                if (!dst) return {Ready, out};  // page full; resume at st.off
                schema.writeData(dst, 0, &id);
                memcpy(schema.calcOffset(dst, 1), t, insz);
                ++st.off;
            }
        }
        if (st.rc == Finished) {
            st.in = nullptr;
            return {Finished, out};
        }
        GetNextResultT r = nextOp->getNext(tid);
        if (r.first == Error) return {Error, out};
        st.in = r.second;
        st.rc = r.first;
        st.off = 0;
    }
}

void ThreadIdPrependOp::threadClose(unsigned short tid) {
    delete st_[tid].out;
    st_[tid].out = nullptr;
}
