/*
 * gpujoin.cpp - see gpujoin.h.  gnu++98, against the reference engine's
 * headers; every device action is a pgpu_* call.
 */
#include "gpujoin.h"

#include <stdio.h>
#include <string.h>

#include <string>

#include "hash.h"

namespace {

// A failed pgpu call: its message to stderr (the reference's exceptions
// carry none), then the reference's own execution error.
void gpu_check(int rc, const char* what) {
    if (rc != PGPU_OK) {
        fprintf(stderr, "%s: %s\n", what, pgpu_last_error());
        throw QueryExecutionError();
    }
}

// enum pgpu_col_type of a join key column (int, long and date: integers by value)
unsigned key_type(ColumnType t) {
    switch (t) {
        case CT_DECIMAL: return PGPU_COL_F64;
        case CT_CHAR: return PGPU_COL_CHAR;
        case CT_INTEGER: return PGPU_COL_INT32;
        default: return PGPU_COL_INT64;
    }
}

// Byte offset of column `pos` in a tuple of schema s.
unsigned column_offset(Schema& s, unsigned pos) {
    return (unsigned)(reinterpret_cast<unsigned long long>(s.calcOffset(0, pos)));
}

}  // namespace

// ------------------------------------------------------------------ barrier
GpuGroupBarrier::GpuGroupBarrier() : n_(0), waiting_(0), generation_(0), broken_(false) {
    pthread_mutex_init(&mu_, 0);
    pthread_cond_init(&cv_, 0);
}

GpuGroupBarrier::~GpuGroupBarrier() {
    pthread_cond_destroy(&cv_);
    pthread_mutex_destroy(&mu_);
}

void GpuGroupBarrier::init(int n) {
    pthread_mutex_lock(&mu_);
    n_ = n;
    waiting_ = 0;
    broken_ = false;
    pthread_mutex_unlock(&mu_);
}

bool GpuGroupBarrier::arrive() {
    pthread_mutex_lock(&mu_);
    if (broken_) {
        pthread_mutex_unlock(&mu_);
        return false;
    }
    const unsigned long gen = generation_;
    if (++waiting_ == n_) {
        waiting_ = 0;
        ++generation_;
        pthread_cond_broadcast(&cv_);
    } else {
        while (gen == generation_ && !broken_) pthread_cond_wait(&cv_, &mu_);
    }
    const bool ok = !broken_ || gen != generation_;
    pthread_mutex_unlock(&mu_);
    return ok;
}

void GpuGroupBarrier::breakAll() {
    pthread_mutex_lock(&mu_);
    broken_ = true;
    pthread_cond_broadcast(&cv_);
    pthread_mutex_unlock(&mu_);
}

void GpuPlanKnobs::read(libconfig::Setting& node) {
    if (!node.exists("gpu")) return;
    libconfig::Setting& g = node["gpu"];
    g.lookupValue("device", device);
    long v = 0;  // libconfig 1.2 reads integers into long (64 bits on x86-64)
    if (g.lookupValue("hbmbudget", v) && v >= 0) hbmbudget = (unsigned long long)v;
    if (g.lookupValue("probebatch", v) && v > 0) probebatch = (unsigned long long)v;
}

// ===================================================================== join
GpuHashJoinOp::GpuHashJoinOp() : ctx(0) {}

GpuHashJoinOp::~GpuHashJoinOp() {}

void GpuHashJoinOp::init(libconfig::Config& root, libconfig::Setting& node) {
    // projection, buildjattr / probejattr, threadgroups and the output schema
    // exactly as the CPU join (operators/join.cpp:184-260)
    JoinOp::init(root, node);
    Schema& bs = buildOp->getOutSchema();
    Schema& ps = probeOp->getOutSchema();
    sbuild.add(bs.get(joinattr1));
    sbuild_src.push_back(column_offset(bs, joinattr1));
    for (unsigned i = 0; i < projection.size(); ++i) {
        if (projection[i].first != BuildSide) continue;
        sbuild.add(bs.get(projection[i].second));
        sbuild_src.push_back(column_offset(bs, projection[i].second));
    }
    // the hash node is checked by building both hashers as HashJoinOp::init
    // does (join.cpp:290-300); the device places tuples with its own hash,
    // the join result does not depend on it
    node["hash"].add("field", libconfig::Setting::TypeInt) = (int)joinattr1;
    TupleHasher hb = TupleHasher::create(bs, node["hash"]);
    node["hash"].remove("field");
    node["hash"].add("field", libconfig::Setting::TypeInt) = (int)joinattr2;
    TupleHasher hp = TupleHasher::create(ps, node["hash"]);
    node["hash"].remove("field");
    hb.destroy();
    hp.destroy();
    (void)(int)node["tuplesperbucket"];
    std::string policy = node["allocpolicy"];
    (void)policy;
    // the key pairs Comparator::init accepts (comparator.cpp:43-413); the
    // reference itself throws the same exception from createComparator
    {
        const ColumnType bt = bs.getColumnType(joinattr1), pt = ps.getColumnType(joinattr2);
        const bool num = (bt == CT_INTEGER || bt == CT_LONG || bt == CT_DECIMAL) &&
                         (pt == CT_INTEGER || pt == CT_LONG || pt == CT_DECIMAL);
        if (!num && !(bt == CT_DATE && pt == CT_DATE) && !(bt == CT_CHAR && pt == CT_CHAR))
            throw UnknownComparisonException();
    }
    knobs.read(node);

    gpu_check(pgpu_ctx_create(knobs.device, knobs.hbmbudget, &ctx), "pgpu_ctx_create");
    pgpu_hj_desc d;
    memset(&d, 0, sizeof d);
    d.build_row_bytes = sbuild.getTupleSize();
    d.build_key_offset = 0;
    d.build_key_bytes = sbuild.getColumnWidth(0);
    d.probe_row_bytes = ps.getTupleSize();
    d.probe_key_offset = column_offset(ps, joinattr2);
    d.probe_key_bytes = ps.getColumnWidth(joinattr2);
    d.build_key_type = key_type(bs.getColumnType(joinattr1));
    d.probe_key_type = key_type(ps.getColumnType(joinattr2));
    d.nproj = projection.size();
    if (d.nproj > PGPU_MAX_PROJ) throw InvalidParameter();
    unsigned buildcol = 1;
    for (unsigned j = 0; j < d.nproj; ++j) {
        if (projection[j].first == BuildSide) {
            d.proj[j].side = 0;
            d.proj[j].src_offset = column_offset(sbuild, buildcol);
            d.proj[j].bytes = sbuild.getColumnWidth(buildcol);
            ++buildcol;
        } else {
            d.proj[j].side = 1;
            d.proj[j].src_offset = column_offset(ps, projection[j].second);
            d.proj[j].bytes = ps.getColumnWidth(projection[j].second);
        }
    }
    d.slots = MAX_THREADS;  // one probe stream per worker thread
    for (unsigned g = 0; g < groupleader.size(); ++g) {
        pgpu_hj* t = 0;
        gpu_check(pgpu_hj_create(ctx, &d, &t), "pgpu_hj_create");
        tables.push_back(t);
        groupbarriers.push_back(new GpuGroupBarrier());
        groupbarriers.back()->init(groupsize[g]);
    }
    state.resize(MAX_THREADS);
}

void GpuHashJoinOp::failGroup(unsigned short g) { groupbarriers.at(g)->breakAll(); }

void GpuHashJoinOp::threadInit(unsigned short tid) {
    if (threadgroups.at(tid) == (unsigned short)-1) throw InvalidParameter();
    state[tid] = ThreadState();
    state[tid].out = new Page(buffsize, schema.getTupleSize(), this);
}

// Build: every build tuple projected into sbuild and staged (buildFromPage,
// join.cpp:641-667); the group's barrier; the leader builds the device
// table; the barrier; the first probe batch (an empty probe side is
// Finished, join.cpp:444-459).
Operator::ResultCode GpuHashJoinOp::scanStart(unsigned short tid, Page* idx, Schema& idxschema) {
    const unsigned short g = threadgroups[tid];
    if (buildOp->scanStart(tid, idx, idxschema) == Error) {
        failGroup(g);
        return Error;
    }
    const unsigned sb = sbuild.getTupleSize();
    std::vector<char> rows;
    GetNextResultT r(Ready, (Page*)0);
    try {
        while (r.first == Ready) {
            r = buildOp->getNext(tid);
            if (r.first == Error) break;
            Page::Iterator it = r.second->createIterator();
            void* tup;
            while ((tup = it.next())) {
                const size_t base = rows.size();
                rows.resize(base + sb);
                unsigned o = 0;
                for (unsigned c = 0; c < sbuild_src.size(); ++c) {
                    const unsigned w = sbuild.getColumnWidth(c);
                    memcpy(&rows[base + o], (const char*)tup + sbuild_src[c], w);
                    o += w;
                }
            }
            if (rows.size() >= (32u << 20)) {
                gpu_check(pgpu_hj_build_append(tables[g], &rows[0], rows.size() / sb), "pgpu_hj_build_append");
                rows.clear();
            }
        }
        if (r.first != Error && !rows.empty())
            gpu_check(pgpu_hj_build_append(tables[g], &rows[0], rows.size() / sb), "pgpu_hj_build_append");
    } catch (...) {
        failGroup(g);
        throw;
    }
    if (r.first == Error || buildOp->scanStop(tid) == Error) {
        failGroup(g);
        return Error;
    }
    if (!groupbarriers[g]->arrive()) return Error;
    if (groupleader[g] == tid && pgpu_hj_build_finish(tables[g]) != PGPU_OK) {
        fprintf(stderr, "pgpu_hj_build_finish: %s\n", pgpu_last_error());
        failGroup(g);
        return Error;
    }
    if (!groupbarriers[g]->arrive()) return Error;
    ResultCode rc = probeOp->scanStart(tid, idx, idxschema);
    if (rc == Error) return Error;
    ThreadState& st = state[tid];
    Page* keep = st.out;
    st = ThreadState();
    st.out = keep;
    if (!stageProbe(tid)) {
        st.done = true;
        return Finished;
    }
    return rc;
}

// Copies probe tuples (child pages are only valid until the next pull) until
// probebatch rows are staged or the probe side is depleted.
bool GpuHashJoinOp::stageProbe(unsigned short tid) {
    ThreadState& st = state[tid];
    st.probe.clear();
    st.nprobe = 0;
    const unsigned ts = probeOp->getOutSchema().getTupleSize();
    while (!st.depleted && st.nprobe < knobs.probebatch) {
        GetNextResultT r = probeOp->getNext(tid);
        if (r.first == Error) throw QueryExecutionError();  // readNextTupleFromProbe (join.cpp:584-586)
        const unsigned long long n = r.second->getUsedSpace() / ts;
        if (n) {
            const char* src = (const char*)r.second->getTupleOffset(0);
            st.probe.insert(st.probe.end(), src, src + n * ts);
        }
        st.nprobe += n;
        if (r.first == Finished) st.depleted = true;
    }
    return st.nprobe > 0;
}

// Fills the thread's page with output rows: those of the last probe batch
// still on the device first, then the next batch is probed.  Ready with a
// full page; Finished (with the last rows) when the probe side is done, and
// Finished with an empty page on every later call.
Operator::GetNextResultT GpuHashJoinOp::getNext(unsigned short tid) {
    ThreadState& st = state[tid];
    Page* out = st.out;
    out->clear();
    if (st.done) return GetNextResultT(Finished, out);
    const unsigned ts = schema.getTupleSize();
    pgpu_hj* t = tables[threadgroups[tid]];
    while (true) {
        if (st.pending) {
            const unsigned long long room = (out->capacity() - out->getUsedSpace()) / ts;
            const unsigned long long k = room < st.pending ? room : st.pending;
            unsigned long long got = 0;
            if (k) {
                void* dst = out->allocate(k * ts);
                gpu_check(pgpu_hj_pull(t, tid, dst, k, (uint64_t*)&got), "pgpu_hj_pull");
            }
            st.pending -= got;
            if (st.pending) return GetNextResultT(Ready, out);  // the page is full
            continue;
        }
        if (st.nprobe) {
            unsigned long long nout = 0;
            gpu_check(pgpu_hj_probe(t, tid, &st.probe[0], st.nprobe, (uint64_t*)&nout), "pgpu_hj_probe");
            st.pending = nout;
            st.nprobe = 0;
            continue;
        }
        if (!stageProbe(tid)) {
            st.done = true;
            return GetNextResultT(Finished, out);
        }
    }
}

Operator::ResultCode GpuHashJoinOp::scanStop(unsigned short tid) { return probeOp->scanStop(tid); }

void GpuHashJoinOp::threadClose(unsigned short tid) {
    delete state[tid].out;
    state[tid].out = 0;
}

void GpuHashJoinOp::destroy() {
    for (unsigned i = 0; i < tables.size(); ++i) pgpu_hj_destroy(tables[i]);
    tables.clear();
    for (unsigned i = 0; i < groupbarriers.size(); ++i) delete groupbarriers[i];
    groupbarriers.clear();
    if (ctx) pgpu_ctx_destroy(ctx);
    ctx = 0;
}

// ================================================================ aggregate
GpuAggregate::GpuAggregate() : kind_(0), valoffset_(0), global_(false), threads_(0), shared_(0), ctx(0) {
    pthread_mutex_init(&mu_, 0);
}

GpuAggregate::~GpuAggregate() { pthread_mutex_destroy(&mu_); }

void GpuAggregate::init(libconfig::Config& root, libconfig::Setting& node) {
    Operator::init(root, node);
    if (node.exists("fields")) {
        libconfig::Setting& f = node["fields"];
        for (int i = 0; i < f.getLength(); ++i) aggfields.push_back((unsigned short)(int)f[i]);
    } else {
        aggfields.push_back((unsigned short)(int)node["field"]);
    }
    // no GROUP BY: one group, hashed by alwayszero (genericaggregate.cpp:80-87)
    if (aggfields.empty() && !node.exists("hash")) {
        libconfig::Setting& h = node.add("hash", libconfig::Setting::TypeGroup);
        h.add("fn", libconfig::Setting::TypeString) = "alwayszero";
    }
    Schema& in = nextOp->getOutSchema();
    TupleHasher th = TupleHasher::create(in, node["hash"]);  // validates the node (hash.cpp:99-234)
    th.destroy();
    for (unsigned i = 0; i < aggfields.size(); ++i) schema.add(in.get(aggfields[i]));
    Schema uds;
    foldschema(node, uds);
    for (unsigned i = 0; i < uds.columns(); ++i) schema.add(uds.get(i));
    if (node.exists("presorted")) throw NotYetImplemented();
    if (node.exists("global")) {
        global_ = true;
        threads_ = node["threads"];
        barrier_.init(threads_);
    }
    knobs.read(node);
    gpu_check(pgpu_ctx_create(knobs.device, knobs.hbmbudget, &ctx), "pgpu_ctx_create");
    state.resize(MAX_THREADS);
}

pgpu_agg* GpuAggregate::newTable() {
    Schema& in = nextOp->getOutSchema();
    pgpu_agg_desc d;
    memset(&d, 0, sizeof d);
    d.row_bytes = in.getTupleSize();
    d.nkeys = aggfields.size();
    if (d.nkeys > 8) throw InvalidParameter();
    for (unsigned i = 0; i < d.nkeys; ++i) {
        d.key_offset[i] = column_offset(in, aggfields[i]);
        d.key_bytes[i] = in.getColumnWidth(aggfields[i]);
    }
    d.kind = kind_;
    d.val_offset = valoffset_;
    pgpu_agg* a = 0;
    gpu_check(pgpu_agg_create(ctx, &d, &a), "pgpu_agg_create");
    return a;
}

void GpuAggregate::threadInit(unsigned short tid) {
    state[tid] = ThreadState();
    state[tid].out = new Page(buffsize, schema.getTupleSize(), this);
    if (!global_) {
        state[tid].agg = newTable();
        return;
    }
    pthread_mutex_lock(&mu_);
    try {
        if (!shared_) shared_ = newTable();
    } catch (...) {
        pthread_mutex_unlock(&mu_);
        barrier_.breakAll();
        throw;
    }
    pthread_mutex_unlock(&mu_);
}

// Drains the input into the device table (remember, genericaggregate.cpp:
// 240-282), then compacts the groups; in Global mode each thread emits a
// contiguous share of them (genericaggregate.cpp:324-412).
Operator::ResultCode GpuAggregate::scanStart(unsigned short tid, Page* idx, Schema& idxschema) {
    ThreadState& st = state[tid];
    ResultCode rc = nextOp->scanStart(tid, idx, idxschema);
    if (rc == Error) {
        if (global_) barrier_.breakAll();
        return Error;
    }
    pgpu_agg* t = global_ ? shared_ : st.agg;
    const unsigned ts = nextOp->getOutSchema().getTupleSize();
    std::vector<char> rows;
    GetNextResultT r(rc, (Page*)0);
    try {
        if (rc == Ready) {
            do {
                r = nextOp->getNext(tid);
                if (r.first == Error) break;
                const unsigned long long n = r.second->getUsedSpace() / ts;
                if (n) {
                    const char* src = (const char*)r.second->getTupleOffset(0);
                    rows.insert(rows.end(), src, src + n * ts);
                }
                if (rows.size() >= (64u << 20)) {
                    gpu_check(pgpu_agg_append(t, &rows[0], rows.size() / ts), "pgpu_agg_append");
                    rows.clear();
                }
            } while (r.first == Ready);
        }
        if (r.first != Error && !rows.empty())
            gpu_check(pgpu_agg_append(t, &rows[0], rows.size() / ts), "pgpu_agg_append");
    } catch (...) {
        if (global_) barrier_.breakAll();
        throw;
    }
    if (nextOp->scanStop(tid) == Error || r.first == Error) {
        if (global_) barrier_.breakAll();
        return Error;
    }
    unsigned long long ng = 0;
    if (global_ && !barrier_.arrive()) return Error;  // every thread has folded its input
    gpu_check(pgpu_agg_finish(t, (uint64_t*)&ng, 0), "pgpu_agg_finish");
    if (global_) {
        st.first = (unsigned long long)tid * ng / threads_;
        st.end = (unsigned long long)(tid + 1) * ng / threads_;
    } else {
        st.first = 0;
        st.end = ng;
    }
    st.pos = st.first;
    return Ready;
}

Operator::GetNextResultT GpuAggregate::getNext(unsigned short tid) {
    ThreadState& st = state[tid];
    Page* out = st.out;
    out->clear();
    if (st.pos >= st.end) return GetNextResultT(Finished, out);
    const unsigned ts = schema.getTupleSize();
    unsigned long long k = out->capacity() / ts;
    if (k > st.end - st.pos) k = st.end - st.pos;
    void* dst = out->allocate(k * ts);
    gpu_check(pgpu_agg_pull(global_ ? shared_ : st.agg, st.pos, k, dst), "pgpu_agg_pull");
    st.pos += k;
    return GetNextResultT(st.pos >= st.end ? Finished : Ready, out);
}

void GpuAggregate::threadClose(unsigned short tid) {
    delete state[tid].out;
    state[tid].out = 0;
    if (state[tid].agg) pgpu_agg_destroy(state[tid].agg);
    state[tid].agg = 0;
}

void GpuAggregate::destroy() {
    if (shared_) pgpu_agg_destroy(shared_);
    shared_ = 0;
    if (ctx) pgpu_ctx_destroy(ctx);
    ctx = 0;
}

void GpuAggregateSum::foldschema(libconfig::Setting& node, Schema& out) {
    const unsigned sumfield = (int)node["sumfield"];
    Schema& in = nextOp->getOutSchema();
    if (sumfield >= in.columns()) throw InvalidParameter();
    switch (in.getColumnType(sumfield)) {
        case CT_INTEGER: kind_ = PGPU_AGG_SUM_I32; break;
        case CT_LONG: kind_ = PGPU_AGG_SUM_I64; break;
        case CT_DECIMAL: kind_ = PGPU_AGG_SUM_F64; break;
        default: throw InvalidParameter();  // AggregateSum sums int, long or dec (aggregatesum.cpp:46-49)
    }
    valoffset_ = column_offset(in, sumfield);
    out.add(in.get(sumfield));
}

void GpuAggregateCount::foldschema(libconfig::Setting&, Schema& out) {
    kind_ = PGPU_AGG_COUNT;
    out.add(CT_LONG);
}
