// TEST INFRASTRUCTURE ONLY: shared by the drivers that oracle/ref.mk links
// against the reference engine compiled from its sources.
#ifndef ORACLE_REF_MEMSCAN_H
#define ORACLE_REF_MEMSCAN_H

#include <stdint.h>
#include <time.h>

#include <vector>

#include "operators/operators.h"

using namespace libconfig;

// Hands out buffsize-sized overlay pages of a packed relation to any thread
// through a shared atomic cursor (the Table::atomicReadNext pattern).
class MemScan : public ZeroInputOp {
  public:
    MemScan(char* rows, uint64_t n, Schema s) : rows_(rows), n_(n), s_(s), cursor_(0) {}
    virtual void init(Config& root, Setting& node) {
        Operator::init(root, node);
        schema = s_;
        rpp_ = buffsize / schema.getTupleSize();
        if (rpp_ == 0) rpp_ = 1;
        for (int i = 0; i < MAX_THREADS; ++i) pages_.push_back(NULL);
        empty_ = new Page(rows_, 0, rows_, schema.getTupleSize());
    }
    virtual GetNextResultT getNext(unsigned short tid) {
        delete pages_[tid];
        pages_[tid] = NULL;
        uint64_t start = __sync_fetch_and_add(&cursor_, rpp_);
        if (start >= n_) return std::make_pair(Finished, empty_);
        uint64_t k = n_ - start < rpp_ ? n_ - start : rpp_;
        const unsigned ts = schema.getTupleSize();
        pages_[tid] = new Page(rows_ + start * ts, k * ts, NULL, ts);
        return std::make_pair(start + k >= n_ ? Finished : Ready, pages_[tid]);
    }

  private:
    char* rows_;
    uint64_t n_;
    Schema s_;
    volatile uint64_t cursor_;
    uint64_t rpp_;
    std::vector<Page*> pages_;
    Page* empty_;
};

static double now() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

#endif
