// `gpuhashjoin`, `gpuaggregate_sum`, `gpuaggregate_count` and `gpupartition`:
// drop-ins for Pythia's `hashjoin`, `aggregate_sum`, `aggregate_count` and
// `partition` plan types.
// They read exactly the reference operators' configuration keys and produce
// the same output schema and Page protocol; the join / aggregation itself
// runs on the GPU through the C ABI of pgpu.h (nothing else crosses to the
// device).  Optional extra keys, ignored by the CPU types, live in a `gpu`
// group: device, hbmbudget (bytes), probebatch (probe rows per device call),
// resident (gpuhashjoin over unmapped scans: the scanned relations are kept in
// HBM across queries and joined there; a gpuaggregate directly above such a
// join folds its output in HBM).
#pragma once

#include <atomic>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "operators.h"
#include "pgpu.h"

struct GpuKnobs {
    int device = 0;
    uint64_t hbmbudget = 0;
    uint64_t probebatch = 1 << 20;
    bool resident = false;
    void read(libconfig::Setting& node);
};

class GpuHashJoinOp : public DualInputOp {
  public:
    enum JoinSrcT { BuildSide, ProbeSide };
    typedef std::pair<JoinSrcT, unsigned int> JoinPrjT;

    void accept(Visitor* v) override { v->visit(static_cast<DualInputOp*>(this)); }
    // Keys as HashJoinOp::init / JoinOp::init (operators/join.cpp:184-366):
    // hash{fn, buckets, ...}, tuplesperbucket, buildjattr, probejattr,
    // projection, threadgroups, allocpolicy, stripeon.
    void init(libconfig::Config& root, libconfig::Setting& node) override;
    void threadInit(unsigned short threadid) override;
    ResultCode scanStart(unsigned short threadid, Page* indexdatapage, Schema& indexdataschema) override;
    GetNextResultT getNext(unsigned short threadid) override;
    ResultCode scanStop(unsigned short threadid) override;
    void threadClose(unsigned short threadid) override;
    void destroy() override;

    const Schema& buildSchema() const { return sbuild; }
    pgpu_hj_stats stats(unsigned short group);

    // Resident mode (gpu { resident = true; }): both inputs are device copies
    // of the child scans' relations, one thread group, probe rows taken from
    // a shared cursor in probebatch chunks.
    bool residentMode() const { return resident_; }
    int device() const { return knobs.device; }
    // Resident mode: joins thread tid's next probe chunk and hands out the
    // output rows in HBM (valid until the thread's next call); false when the
    // probe relation is exhausted.  Used between scanStart and scanStop in
    // place of getNext.
    bool nextDeviceBatch(unsigned short tid, const void** d_rows, uint64_t* n);

    // Fused consumers.  An aggregate directly above a resident join picks
    // one of these during its init (after the join's); the join then builds
    // no row-format table, and the aggregate calls the matching method below
    // between the join's scanStart and scanStop instead of getNext.
    //   FuseRows:     join output rows handed over in HBM (nextDeviceBatch)
    //   FuseCountSum: no GROUP BY, count or sum of one projected integer
    //                 column of the key's width: the 2-pass radix join with
    //                 its fused count + sum over dense device columns
    //   FuseGroup:    sum of a projected dec probe column grouped by one
    //                 projected int build column, int keys: the partitioned
    //                 group join
    enum FuseKind { FuseNone = 0, FuseRows = 1, FuseCountSum = 2, FuseGroup = 4 };
    bool canFuseCountSum(int outcol) const;              // outcol < 0: count only
    bool canFuseGroup(int groupcol, int sumcol) const;
    void setFused(FuseKind k) { fused_ = k; }
    FuseKind fused() const { return fused_; }
    struct CountSum {
        uint64_t matches = 0;
        uint64_t sum = 0;  // modulo 2^64 (the low 32 bits for an int column)
    };
    // In the fused forms the threads do not meet at a barrier: after its
    // scanStart each thread asks claimRun(), and the one thread that gets
    // true for this run of the plan runs the join (countSum / groupRows) and
    // emits the whole result; the others have nothing to emit.
    bool claimRun(unsigned short tid);
    CountSum countSum(int outcol);
    // The group join's result rows of this run (int group key, FP64 sum: 12
    // bytes each) in pinned host memory owned by the join until the next run
    // or destroy(): the thread that claims the run computes them, the others
    // wait for the count (every thread then emits its share).  The rows are
    // written out in order after the count is published; waitGroupRows(upto)
    // returns once rows [0, upto) are in place.
    const char* groupRows(unsigned short tid, int groupcol, int sumcol, uint64_t* n);
    void waitGroupRows(uint64_t upto);
    // Threads of the (single, in resident mode) thread group and tid's place in it.
    unsigned short groupThreads() const { return groupsize.empty() ? 1 : groupsize[0]; }
    unsigned short groupPosition(unsigned short tid) const { return threadposingrp[tid]; }
    // Join output bytes copied to the host by pgpu_hj_pull during this query.
    uint64_t pulledBytes() const { return pulled_bytes_; }

  private:
    struct ThreadState {
        Page* out = nullptr;
        std::vector<char> probe;   // staged probe rows (child pages are transient)
        uint64_t nprobe = 0;
        uint64_t pending = 0;      // device output rows not yet pulled
        bool depleted = false;     // probe child returned Finished
        bool done = false;
    };
    bool stageProbe(unsigned short tid);  // false when nothing more to stage
    void fail_group(unsigned short group);

    std::vector<JoinPrjT> projection;
    unsigned int joinattr1 = 0, joinattr2 = 0;
    std::vector<unsigned short> threadgroups, threadposingrp, groupleader, groupsize;
    std::vector<Barrier*> barriers;
    Schema sbuild;
    std::vector<pgpu_hj*> tables;
    std::vector<ThreadState> state;
    GpuKnobs knobs;
    pgpu_ctx* ctx = nullptr;
    std::mutex mu;
    // resident mode
    bool resident_ = false;
    const char* d_probe_ = nullptr;
    uint64_t nprobe_rows_ = 0;
    std::atomic<uint64_t> probe_cursor_{0};
    ResultCode scanStartResident(unsigned short tid, Page* idx, Schema& idxschema);
    bool probeNextResident(unsigned short tid, uint64_t* nout);
    const void* residentColumn(bool build, unsigned col);
    const char* d_build_ = nullptr;
    uint64_t nbuild_rows_ = 0;
    std::string build_key_, probe_key_;
    // fused consumers
    FuseKind fused_ = FuseNone;
    std::mutex fuse_mu_;
    std::vector<uint64_t> runs_;             // per thread: scanStart calls so far
    std::atomic<uint64_t> claimed_run_{0};   // the last run a thread claimed
    void residentInputs();                   // device copies of both relations (fuse_mu_ held)
    char* fuse_rows_ = nullptr;              // pinned, from the registry's pool
    uint64_t fuse_rows_cap_ = 0;
    std::condition_variable fuse_cv_;        // signals a published run (under fuse_mu_)
    uint64_t published_run_ = 0;             // the last run whose group rows are out
    uint64_t published_rows_ = 0;
    uint64_t copied_rows_ = 0;               // rows of the published run already in fuse_rows_
    std::string run_error_;                  // the claimer's failure, rethrown by the waiters
    std::atomic<uint64_t> pulled_bytes_{0};
};

class GpuAggregate : public SingleInputOp {
  public:
    void accept(Visitor* v) override { v->visit(static_cast<SingleInputOp*>(this)); }
    // Keys as GenericAggregate::init (operators/genericaggregate.cpp:51-173):
    // field | fields, hash{...}, global + threads, presorted (not implemented).
    void init(libconfig::Config& root, libconfig::Setting& node) override;
    void threadInit(unsigned short threadid) override;
    ResultCode scanStart(unsigned short threadid, Page* indexdatapage, Schema& indexdataschema) override;
    GetNextResultT getNext(unsigned short threadid) override;
    ResultCode scanStop(unsigned short) override { return Ready; }
    void threadClose(unsigned short threadid) override;
    void destroy() override;

  protected:
    // The aggregate's own columns (AggregateSum / AggregateCount foldinit).
    virtual Schema foldinit(libconfig::Setting& node) = 0;
    virtual uint32_t kind() const = 0;
    virtual uint32_t valOffset() const { return 0; }
    virtual int sumField() const { return -1; }  // the summed input column (-1: a count)

    std::vector<unsigned short> aggfields;

  public:
    // The fused form chosen at init over a resident gpuhashjoin (FuseNone otherwise).
    GpuHashJoinOp::FuseKind fused() const { return fuse_; }

  private:
    struct ThreadState {
        Page* out = nullptr;
        pgpu_agg* agg = nullptr;  // ThreadLocal table
        uint64_t first = 0, end = 0, pos = 0;
        bool done = false;
        std::vector<char> row;      // FuseCountSum: the one output row, if this thread emits it
        const char* rows = nullptr; // FuseGroup: the join's result rows
        std::shared_ptr<Page> view; // FuseGroup: overlay page over `rows`
    };
    GpuHashJoinOp::FuseKind fuse_ = GpuHashJoinOp::FuseNone;
    GpuHashJoinOp* join_ = nullptr;  // the join below, when fused
    ResultCode scanStartFused(unsigned short tid, Page* idx, Schema& idxschema);
    pgpu_agg* newTable();
    bool global_ = false;
    int threads_ = 0;
    pgpu_agg* shared_ = nullptr;
    uint64_t shared_groups_ = 0;
    Barrier barrier_;
    std::vector<ThreadState> state;
    GpuKnobs knobs;
    pgpu_ctx* ctx = nullptr;
    std::mutex mu, init_mu;
};

class GpuAggregateSum : public GpuAggregate {
  protected:
    // sumfield must be int, long or dec; the accumulator has its type
    // (operators/aggregatesum.cpp:37-51).
    Schema foldinit(libconfig::Setting& node) override;
    uint32_t kind() const override { return kind_; }
    uint32_t valOffset() const override { return valoff_; }
    int sumField() const override { return sumfield_; }

  private:
    uint32_t kind_ = 0, valoff_ = 0;
    int sumfield_ = -1;
};

class GpuAggregateCount : public GpuAggregate {
  protected:
    Schema foldinit(libconfig::Setting& node) override;  // one CT_LONG count
    uint32_t kind() const override { return PGPU_AGG_COUNT; }
};

class GpuPartitionOp : public SingleInputOp {
  public:
    void accept(Visitor* v) override { v->visit(static_cast<SingleInputOp*>(this)); }
    // Keys as PartitionOp::init (operators/partition.cpp:44-101): attr,
    // range = [min, max] and buckets of its exactrange hash (buckets = the
    // threads running the operator, < 128), maxtuples | maxtuplesinM (required
    // there to size the staging area; read and ignored here), sort = "yes" |
    // "no", sortattr (defaults to attr).
    void init(libconfig::Config& root, libconfig::Setting& node) override;
    void threadInit(unsigned short threadid) override;
    // Drains the child (scanStart .. scanStop), waits for every thread, the
    // first thread partitions all rows on the device, and each thread pulls
    // its own partition (operators/partition.cpp:271-359).
    ResultCode scanStart(unsigned short threadid, Page* indexdatapage, Schema& indexdataschema) override;
    // Pages of at most buffsize / tuplesize rows over the thread's partition;
    // the last one (possibly empty) comes with Finished (partition.cpp:376-404).
    GetNextResultT getNext(unsigned short threadid) override;
    ResultCode scanStop(unsigned short threadid) override;
    void threadClose(unsigned short threadid) override;
    void destroy() override;

  private:
    struct ThreadState {
        std::vector<char> rows;  // this thread's partition
        uint64_t n = 0, pos = 0;
        std::unique_ptr<Page> view;
    };
    unsigned attribute_ = 0, sortattr_ = 0, buckets_ = 0;
    int64_t rmin_ = 0, rmax_ = 0;
    bool sort_ = false;
    Barrier barrier_;
    pgpu_part* part_ = nullptr;
    std::vector<uint64_t> counts_;
    std::vector<ThreadState> state;
    GpuKnobs knobs;
    pgpu_ctx* ctx = nullptr;
    std::mutex mu;
};
