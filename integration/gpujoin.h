/*
 * gpujoin.h - the GPU plan operators as classes of the reference engine
 * (sblanas/pythia), for a maintainer to add to its tree next to
 * operators/join.cpp.  They are written against the reference's own headers
 * (operators/operators.h:79-186, schema.h, hash.h, util/buffer.h) in its
 * dialect (gnu++98) and reach the GPU only through the C ABI of pgpu.h
 * (libpgpu.so).
 *
 *   gpuhashjoin         GpuHashJoinOp      in place of HashJoinOp
 *   gpuaggregate_sum    GpuAggregateSum    in place of AggregateSum
 *   gpuaggregate_count  GpuAggregateCount  in place of AggregateCount
 *
 * They read the same configuration keys as the operators they replace and
 * produce the same output schema and Page protocol; optional keys live in a
 * `gpu` group (device, hbmbudget, probebatch).  Registration: one branch per
 * type string in constructsubtree (query.cpp:130-203) creating the object.
 */
#ifndef PYTHIA_INTEGRATION_GPUJOIN_H
#define PYTHIA_INTEGRATION_GPUJOIN_H

#include <pthread.h>

#include <vector>

#include "operators/operators.h"
#include "pgpu.h"

/* A barrier that counts generations (a spurious wakeup cannot release it
 * early) and can be broken: a thread whose GPU call failed breaks it so its
 * group returns Error instead of waiting forever. */
class GpuGroupBarrier {
  public:
    GpuGroupBarrier();
    ~GpuGroupBarrier();
    void init(int n);
    bool arrive();  /* false once broken */
    void breakAll();

  private:
    pthread_mutex_t mu_;
    pthread_cond_t cv_;
    int n_, waiting_;
    unsigned long generation_;
    bool broken_;
};

/* Optional `gpu { device; hbmbudget; probebatch; }` keys. */
struct GpuPlanKnobs {
    GpuPlanKnobs() : device(0), hbmbudget(0), probebatch(1 << 20) {}
    void read(libconfig::Setting& node);
    int device;
    unsigned long long hbmbudget;
    unsigned long long probebatch;
};

/* Keys as HashJoinOp::init (operators/join.cpp:273-366): hash{fn, buckets,
 * ...}, tuplesperbucket, buildjattr, probejattr, projection, threadgroups
 * (both parsed by JoinOp::init), allocpolicy.  Join keys compare as the
 * reference's Comparator does: int / long / dec by value, date with date,
 * char with char (strncmp). */
class GpuHashJoinOp : public JoinOp {
  public:
    GpuHashJoinOp();
    virtual ~GpuHashJoinOp();
    virtual void init(libconfig::Config& root, libconfig::Setting& node);
    virtual void threadInit(unsigned short threadid);
    virtual ResultCode scanStart(unsigned short threadid, Page* indexdatapage, Schema& indexdataschema);
    virtual GetNextResultT getNext(unsigned short threadid);
    virtual ResultCode scanStop(unsigned short threadid);
    virtual void threadClose(unsigned short threadid);
    virtual void destroy();

  private:
    struct ThreadState {
        ThreadState() : out(0), nprobe(0), pending(0), depleted(false), done(false) {}
        Page* out;
        std::vector<char> probe;   /* staged probe rows: child pages are transient */
        unsigned long long nprobe, pending;
        bool depleted, done;
    };
    bool stageProbe(unsigned short tid);
    void failGroup(unsigned short g);

    Schema sbuild;                       /* join key + projected build columns (join.cpp:278-286) */
    std::vector<unsigned> sbuild_src;    /* their offsets in the build child's tuples */
    std::vector<GpuGroupBarrier*> groupbarriers;
    std::vector<pgpu_hj*> tables;        /* one per thread group */
    std::vector<ThreadState> state;
    GpuPlanKnobs knobs;
    pgpu_ctx* ctx;
};

/* Keys as GenericAggregate::init (operators/genericaggregate.cpp:51-173):
 * field | fields, hash{...}, global + threads (Global mode: one table, each
 * group emitted by one thread), presorted (NotYetImplemented). */
class GpuAggregate : public virtual SingleInputOp {
  public:
    GpuAggregate();
    virtual ~GpuAggregate();
    virtual void init(libconfig::Config& root, libconfig::Setting& node);
    virtual void threadInit(unsigned short threadid);
    virtual ResultCode scanStart(unsigned short threadid, Page* indexdatapage, Schema& indexdataschema);
    virtual GetNextResultT getNext(unsigned short threadid);
    virtual ResultCode scanStop(unsigned short threadid) { return Ready; }
    virtual void threadClose(unsigned short threadid);
    virtual void destroy();

  protected:
    /* The aggregate's own columns (AggregateSum / AggregateCount foldinit);
     * sets kind_ and valoffset_. */
    virtual void foldschema(libconfig::Setting& node, Schema& out) = 0;
    unsigned kind_;        /* enum pgpu_agg_kind */
    unsigned valoffset_;   /* summed column's offset in the input tuple */

  private:
    struct ThreadState {
        ThreadState() : out(0), agg(0), first(0), end(0), pos(0) {}
        Page* out;
        pgpu_agg* agg;     /* ThreadLocal mode */
        unsigned long long first, end, pos;
    };
    pgpu_agg* newTable();
    std::vector<unsigned short> aggfields;
    bool global_;
    int threads_;
    pgpu_agg* shared_;
    pthread_mutex_t mu_;
    GpuGroupBarrier barrier_;
    std::vector<ThreadState> state;
    GpuPlanKnobs knobs;
    pgpu_ctx* ctx;
};

/* sumfield: an int, long or dec column; the accumulator has its type
 * (operators/aggregatesum.cpp:37-51), an int sum wrapping mod 2^32. */
class GpuAggregateSum : public GpuAggregate {
  protected:
    virtual void foldschema(libconfig::Setting& node, Schema& out);
};

/* One long count (operators/aggregatecount.cpp:37-42). */
class GpuAggregateCount : public GpuAggregate {
  protected:
    virtual void foldschema(libconfig::Setting& node, Schema& out);
};

#endif
