// The Operator lifecycle of Pythia's plan engine (operators/operators.h:
// 79-409), restated for the GPU operators' host side: the same class names,
// virtual signatures, ResultCode protocol and Page type, so an operator
// written against this header also compiles against the reference's.
//
//   init(root, node)   once, children already wired
//   threadInit(tid)    once per worker thread (ThreadInitVisitor)
//   scanStart(tid,...) propagate down, then produce
//   getNext(tid)       -> (Ready|Finished|Error, Page*); after Finished every
//                      call returns Finished with an empty page
//   scanStop(tid)      propagate down
//   threadClose(tid)   once per worker thread (ThreadCloseVisitor)
//   destroy()          once
#pragma once

#include <pthread.h>

#include <atomic>
#include <memory>
#include <thread>
#include <condition_variable>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "page.h"
#include "plan_config.h"
#include "schema.h"

static const unsigned short MAX_THREADS = 128;

class MissingParameterException : public std::runtime_error {
  public:
    explicit MissingParameterException(const std::string& m = "missing parameter") : std::runtime_error(m) {}
};
class InvalidParameter : public std::runtime_error {
  public:
    explicit InvalidParameter(const std::string& m = "invalid parameter") : std::runtime_error(m) {}
};
class NotYetImplemented : public std::runtime_error {
  public:
    explicit NotYetImplemented(const std::string& m = "not yet implemented") : std::runtime_error(m) {}
};
class QueryExecutionError : public std::runtime_error {
  public:
    explicit QueryExecutionError(const std::string& m = "query execution error") : std::runtime_error(m) {}
};
class UnknownHashException : public std::runtime_error {
  public:
    UnknownHashException() : std::runtime_error("unknown hash function") {}
};
class UnknownComparisonException : public std::runtime_error {
  public:
    UnknownComparisonException() : std::runtime_error("no comparison between these column types") {}
};

class Visitor;

class Operator {
  public:
    enum ResultCode {
        Ready = 0,  // more data follows; the page is valid
        Finished,   // no more data; the page is valid (may be empty)
        Error       // the page must not be used
    };
    typedef TupleBuffer Page;
    typedef std::pair<Operator::ResultCode, Page*> GetNextResultT;

    Operator() : buffsize(0) {}
    virtual ~Operator() {}

    // Reads the global `buffsize` (page bytes) as operators/base.cpp:37-40.
    virtual void init(libconfig::Config& root, libconfig::Setting& node);
    virtual void threadInit(unsigned short threadid) {}
    virtual ResultCode scanStart(unsigned short threadid, Page* indexdatapage, Schema& indexdataschema) = 0;
    virtual GetNextResultT getNext(unsigned short threadid) = 0;
    virtual ResultCode scanStop(unsigned short threadid) = 0;
    virtual void threadClose(unsigned short threadid) {}
    virtual void destroy() {}
    virtual void accept(Visitor* v) = 0;
    virtual Schema& getOutSchema() { return schema; }

  protected:
    Schema schema;
    unsigned int buffsize;
};

class SingleInputOp : public virtual Operator {
  public:
    SingleInputOp() : nextOp(nullptr) {}
    void accept(Visitor* v) override;
    ResultCode scanStart(unsigned short tid, Page* p, Schema& s) override { return nextOp->scanStart(tid, p, s); }
    ResultCode scanStop(unsigned short tid) override { return nextOp->scanStop(tid); }
    Operator* nextOp;
};

class DualInputOp : public virtual Operator {
  public:
    DualInputOp() : buildOp(nullptr), probeOp(nullptr) {}
    void accept(Visitor* v) override;
    ResultCode scanStart(unsigned short tid, Page* p, Schema& s) override {
        ResultCode a = buildOp->scanStart(tid, p, s);
        ResultCode b = probeOp->scanStart(tid, p, s);
        return a == b ? a : Error;
    }
    ResultCode scanStop(unsigned short tid) override {
        ResultCode a = buildOp->scanStop(tid);
        ResultCode b = probeOp->scanStop(tid);
        return a == b ? a : Error;
    }
    Operator* buildOp;
    Operator* probeOp;
};

class ZeroInputOp : public virtual Operator {
  public:
    void accept(Visitor* v) override;
    ResultCode scanStart(unsigned short, Page*, Schema&) override { return Ready; }
    ResultCode scanStop(unsigned short) override { return Ready; }
};

class MergeOp;

// Visitor over the operator categories.  Lifecycle visitors recurse into
// children; MergeOp runs its own subtree per worker thread.
class Visitor {
  public:
    virtual ~Visitor() {}
    virtual void visit(SingleInputOp* op) = 0;
    virtual void visit(DualInputOp* op) = 0;
    virtual void visit(ZeroInputOp* op) = 0;
    virtual void visit(MergeOp* op) = 0;
};

class ThreadInitVisitor : public Visitor {
  public:
    explicit ThreadInitVisitor(unsigned short t) : threadid(t) {}
    void visit(SingleInputOp* op) override;
    void visit(DualInputOp* op) override;
    void visit(ZeroInputOp* op) override;
    void visit(MergeOp* op) override;
    unsigned short threadid;
};

class ThreadCloseVisitor : public Visitor {
  public:
    explicit ThreadCloseVisitor(unsigned short t) : threadid(t) {}
    void visit(SingleInputOp* op) override;
    void visit(DualInputOp* op) override;
    void visit(ZeroInputOp* op) override;
    void visit(MergeOp* op) override;
    unsigned short threadid;
};

class RecursiveDestroyVisitor : public Visitor {
  public:
    void visit(SingleInputOp* op) override;
    void visit(DualInputOp* op) override;
    void visit(ZeroInputOp* op) override;
    void visit(MergeOp* op) override;
};

class RecursiveFreeVisitor : public Visitor {
  public:
    void visit(SingleInputOp* op) override;
    void visit(DualInputOp* op) override;
    void visit(ZeroInputOp* op) override;
    void visit(MergeOp* op) override;
};

// Generation-counted barrier (the reference's PThreadLockCVBarrier has no
// generation counter, Barrier.cpp:71-88).  break_() releases every waiter
// with `false`, so an error in one thread never strands the others.
class Barrier {
  public:
    void init(int n) {
        std::lock_guard<std::mutex> g(mu_);
        n_ = n;
        count_ = 0;
        broken_ = false;
    }
    bool Arrive() {
        std::unique_lock<std::mutex> g(mu_);
        if (broken_) return false;
        const uint64_t gen = gen_;
        if (++count_ == n_) {
            count_ = 0;
            ++gen_;
            cv_.notify_all();
            return true;
        }
        cv_.wait(g, [&] { return gen_ != gen || broken_; });
        return !broken_ || gen_ != gen;
    }
    void break_() {
        std::lock_guard<std::mutex> g(mu_);
        broken_ = true;
        cv_.notify_all();
    }
    bool broken() {
        std::lock_guard<std::mutex> g(mu_);
        return broken_;
    }

  private:
    std::mutex mu_;
    std::condition_variable cv_;
    int n_ = 1, count_ = 0;
    uint64_t gen_ = 0;
    bool broken_ = false;
};

extern TupleBuffer EmptyPage;

// ------------------------------------------------------------------ scans
// Shared page cursor over a relation held in memory, handed out page by page
// to any worker thread (ParallelScanOp + Table::atomicReadNext semantics,
// operators/parallelscan.cpp:176-190).
class ParallelScanOp : public ZeroInputOp {
  public:
    // parallelscan / scan: schema = (...); files = (...); filetype = "text" |
    // "binary"; separators = "|".  Text rows are parsed with atoi/atoll/atof.
    void init(libconfig::Config& root, libconfig::Setting& node) override;
    void threadInit(unsigned short tid) override;
    ResultCode scanStart(unsigned short tid, Page*, Schema&) override;
    GetNextResultT getNext(unsigned short tid) override;
    void threadClose(unsigned short tid) override;
    void destroy() override;

    // An unmapped scan's whole relation and a key naming its contents (file
    // paths, sizes and modification times, or a registered table and its
    // generation), for device-resident copies; false for a mapped scan.
    bool sharedRows(const char** rows, uint64_t* nrows, std::string* key) const;

  protected:
    void setRows(std::vector<char>&& rows);
    std::string source_key_;
    struct Group {
        std::vector<char> rows;
        uint64_t n = 0;
        std::atomic<uint64_t> cursor{0};
    };
    // group_of_[tid]: index into groups_ of the rows that thread reads
    // (0 = the shared relation of an unmapped scan).
    std::vector<int> group_of_ = std::vector<int>(MAX_THREADS, 0);
    std::vector<std::unique_ptr<Group>> groups_;
    uint64_t rows_per_page_ = 1;
    std::vector<Page*> pages_;
};

// In-memory relation registered by name (pythia_register_table); plan:
// type = "memscan"; table = "name"; schema = (...).
class MemoryScanOp : public ParallelScanOp {
  public:
    void init(libconfig::Config& root, libconfig::Setting& node) override;
};

// ------------------------------------------------------------------ merge
// Runs its subtree on `threads` worker threads and serves their pages to a
// single consumer (operators/merge.cpp:123-710): a worker's Finished is
// turned into Ready until the last worker finishes; a worker Error is
// propagated.
class MergeOp : public SingleInputOp {
  public:
    ~MergeOp() override;
    void accept(Visitor* v) override;
    void init(libconfig::Config& root, libconfig::Setting& node) override;
    void threadInit(unsigned short tid) override;
    ResultCode scanStart(unsigned short tid, Page* p, Schema& s) override;
    GetNextResultT getNext(unsigned short tid) override;
    ResultCode scanStop(unsigned short tid) override;
    void threadClose(unsigned short tid) override;
    int threads() const { return nthreads_; }
    // The message of the first worker that failed (empty if none did).
    std::string firstError();

  private:
    enum Cmd { None, DoScanStart, DoGetNext, DoScanStop, DoThreadClose, Exit };
    struct Worker {
        std::thread* th = nullptr;
        Cmd cmd = None;
        bool busy = false;       // executing a command
        bool has_result = false; // produced page not yet taken
        bool finished = false;
        GetNextResultT result{Error, nullptr};
        ResultCode status = Ready;
        std::string error;
    };
    void workerMain(int w);
    void command(int w, Cmd c);
    void waitIdle(int w);

    int nthreads_ = 0;
    int remaining_ = 0;
    int prev_ = -1;
    bool finished_ = false;
    std::vector<Worker> workers_;
    std::mutex mu_;
    std::condition_variable cv_;  // the consumer waits here for workers' results / idleness
    std::vector<std::unique_ptr<std::condition_variable>> wcv_;  // worker w waits on wcv_[w] for a command
};

// threadidprepend (ThreadIdPrependOp, operators/threadidprepend.cpp:42-110):
// each output tuple is (int id of the worker thread that produced it, the
// input tuple's columns).
class ThreadIdPrependOp : public SingleInputOp {
  public:
    void init(libconfig::Config& root, libconfig::Setting& node) override;
    void threadInit(unsigned short tid) override;
    ResultCode scanStart(unsigned short tid, Page* p, Schema& s) override;
    GetNextResultT getNext(unsigned short tid) override;
    void threadClose(unsigned short tid) override;

  private:
    struct State {
        Page* out = nullptr;
        Page* in = nullptr;  // child page being mapped
        ResultCode rc = Ready;
        uint64_t off = 0;
    };
    std::vector<State> st_;
};

// CallStateChecker contract (operators/checker_callstate.cpp:53-159): fails
// the query when a parent violates the lifecycle protocol of its child.
class CallStateChecker : public SingleInputOp {
  public:
    void init(libconfig::Config& root, libconfig::Setting& node) override;
    void threadInit(unsigned short tid) override;
    ResultCode scanStart(unsigned short tid, Page* p, Schema& s) override;
    GetNextResultT getNext(unsigned short tid) override;
    ResultCode scanStop(unsigned short tid) override;
    void threadClose(unsigned short tid) override;

  private:
    enum State { Init, ThreadInit, ScanStarted, GotFinished, ScanStopped, Closed };
    std::vector<int> state_;
    void bad(const char* what, unsigned short tid);
};
