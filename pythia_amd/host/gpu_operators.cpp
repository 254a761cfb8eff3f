// Host side of the GPU plan operators.  Everything device-related goes
// through pgpu.h; this file only adapts the Operator/Page protocol.
#include "gpu_operators.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <set>
#include <sstream>

#include "registry.h"

using pythia::check;

void GpuKnobs::read(libconfig::Setting& node) {
    if (!node.exists("gpu")) return;
    libconfig::Setting& g = node["gpu"];
    g.lookupValue("device", device);
    long long b = 0;
    if (g.lookupValue("hbmbudget", b)) hbmbudget = (uint64_t)b;
    long long pb = 0;
    if (g.lookupValue("probebatch", pb) && pb > 0) probebatch = (uint64_t)pb;
    g.lookupValue("resident", resident);
}

// The checks TupleHasher::create applies to a hash node
// (hash.cpp:99-234), in the same order, so plans the reference rejects are
// rejected here too.  The GPU operators place tuples with their own hash;
// the join and aggregation results do not depend on it.
static void validate_hash(const libconfig::Setting& node, const Schema& s, int fieldmin, int fieldmax) {
    std::string fn = node["fn"];
    if (fn == "alwayszero") return;
    int buckets = node["buckets"];
    if (buckets == 0) throw MissingParameterException("Number of hash buckets cannot be zero.");
    if (fn == "bytes" || fn == "tpchq1magic") return;
    if (fieldmin != fieldmax) throw IllegalSchemaDeclarationException();
    if (fieldmin < 0 || fieldmin >= (int)s.columns()) throw InvalidParameter("hash field out of range");
    switch (s.getColumnType(fieldmin)) {
        case CT_INTEGER:
        case CT_LONG:
        case CT_DATE: break;
        default: throw IllegalSchemaDeclarationException();
    }
    static const std::set<std::string> known = {"modulo", "range",        "exactrange", "parammodulo",
                                                "knuth",  "tpchorderkey", "willis"};
    if (!known.count(fn)) throw UnknownHashException();
    if (fn == "range" || fn == "exactrange") {
        (void)(int)node["range"][0];
        (void)(int)node["range"][1];
    }
}

// Rows of a scan that shares one relation among its threads (resident mode).
static bool shared_size(Operator* op, uint64_t* n) {
    auto* scan = dynamic_cast<ParallelScanOp*>(op);
    const char* rows;
    std::string key;
    return scan && scan->sharedRows(&rows, n, &key);
}

static bool is_int_key(ColumnType t) { return t == CT_INTEGER || t == CT_LONG || t == CT_DATE; }

// enum pgpu_col_type of a join key column (int, long and date: integers by value)
static uint32_t key_type(ColumnType t) {
    switch (t) {
        case CT_DECIMAL: return PGPU_COL_F64;
        case CT_CHAR: return PGPU_COL_CHAR;
        case CT_INTEGER: return PGPU_COL_INT32;
        default: return PGPU_COL_INT64;
    }
}

// createProjectionVector (operators/join.cpp:141-182): "B$i" / "P$j".
static std::vector<GpuHashJoinOp::JoinPrjT> parse_projection(const libconfig::Setting& line) {
    std::vector<GpuHashJoinOp::JoinPrjT> ret;
    for (int i = 0; i < line.getLength(); ++i) {
        std::string s = line[i];
        size_t l = s.find('$');
        if (l == std::string::npos || l == 0) throw InvalidParameter("projection entry `" + s + "'");
        std::istringstream ss(s.substr(l + 1));
        unsigned int attr;
        ss >> attr;
        if (!ss) throw InvalidParameter("projection entry `" + s + "'");
        switch (s.at(l - 1)) {
            case 'B': ret.push_back({GpuHashJoinOp::BuildSide, attr}); break;
            case 'P': ret.push_back({GpuHashJoinOp::ProbeSide, attr}); break;
            default: throw InvalidParameter("projection entry `" + s + "'");
        }
    }
    return ret;
}

// ====================================================================== join
void GpuHashJoinOp::init(libconfig::Config& root, libconfig::Setting& node) {
    Operator::init(root, node);
    projection = parse_projection(node["projection"]);
    joinattr1 = (int)node["buildjattr"];
    joinattr2 = (int)node["probejattr"];
    Schema& bs = buildOp->getOutSchema();
    Schema& ps = probeOp->getOutSchema();
    if (joinattr1 >= bs.columns() || joinattr2 >= ps.columns()) throw InvalidParameter("join attribute out of range");

    // thread -> (group, position), group -> leader, size, barrier
    threadgroups.assign(MAX_THREADS, (unsigned short)-1);
    threadposingrp.assign(MAX_THREADS, (unsigned short)-1);
    libconfig::Setting& part = node["threadgroups"];
    for (int i = 0; i < part.getLength(); ++i) {
        for (int j = 0; j < part[i].getLength(); ++j) {
            int tid = part[i][j];
            if (tid < 0 || tid >= MAX_THREADS || threadgroups[tid] != (unsigned short)-1)
                throw InvalidParameter("thread " + std::to_string(tid) + " in two thread groups or out of range");
            threadgroups[tid] = (unsigned short)i;
            threadposingrp[tid] = (unsigned short)j;
        }
        groupleader.push_back((unsigned short)(int)part[i][0]);
        groupsize.push_back((unsigned short)part[i].getLength());
        barriers.push_back(new Barrier());
        barriers.back()->init(part[i].getLength());
    }

    for (auto& p : projection) {
        if (p.first == BuildSide) {
            if (p.second >= bs.columns()) throw InvalidParameter("projected build column out of range");
            schema.add(bs.get(p.second));
        } else {
            if (p.second >= ps.columns()) throw InvalidParameter("projected probe column out of range");
            schema.add(ps.get(p.second));
        }
    }
    // sbuild = join key + projected build columns (operators/join.cpp:278-286)
    sbuild.add(bs.get(joinattr1));
    for (auto& p : projection)
        if (p.first == BuildSide) sbuild.add(bs.get(p.second));

    if (node["hash"].exists("field")) throw InvalidParameter("hash node must not name `field' for a join");
    validate_hash(node["hash"], bs, (int)joinattr1, (int)joinattr1);
    validate_hash(node["hash"], ps, (int)joinattr2, (int)joinattr2);
    (void)(int)node["tuplesperbucket"];
    std::string policy = node["allocpolicy"];
    if (policy == "striped")
        fprintf(stderr, " ** NUMA POLICY WARNING: gpuhashjoin keeps its table in HBM; allocpolicy is ignored.\n");
    // the key pairs Comparator::init accepts for the join's NotEqual
    // (comparator.cpp:43-413): int / long / dec with each other (by value,
    // as doubles when a dec is involved), date with date, char with char
    {
        const ColumnType bt = bs.getColumnType(joinattr1), pt = ps.getColumnType(joinattr2);
        auto num = [](ColumnType t) { return t == CT_INTEGER || t == CT_LONG || t == CT_DECIMAL; };
        if (!((num(bt) && num(pt)) || (bt == CT_DATE && pt == CT_DATE) || (bt == CT_CHAR && pt == CT_CHAR)))
            throw UnknownComparisonException();
    }
    knobs.read(node);
    tables.assign(groupleader.size(), nullptr);
    state.assign(MAX_THREADS, ThreadState());
    runs_.assign(MAX_THREADS, 0);
    if (knobs.resident) {
        const char* rows;
        uint64_t n;
        std::string key;
        auto* b = dynamic_cast<ParallelScanOp*>(buildOp);
        auto* p = dynamic_cast<ParallelScanOp*>(probeOp);
        if (!b || !p || !b->sharedRows(&rows, &n, &key) || !p->sharedRows(&rows, &n, &key))
            throw InvalidParameter("gpu.resident: both join inputs must be scans without a thread mapping");
        if (groupleader.size() != 1) throw InvalidParameter("gpu.resident: the join must have one thread group");
        resident_ = true;
    }
}

void GpuHashJoinOp::fail_group(unsigned short g) { barriers.at(g)->break_(); }

void GpuHashJoinOp::threadInit(unsigned short tid) {
    if (tid >= MAX_THREADS || threadgroups[tid] == (unsigned short)-1)
        throw InvalidParameter("thread " + std::to_string(tid) + " is in no thread group");
    const unsigned short g = threadgroups[tid];
    if (groupleader[g] == tid) {
        try {
            {
                std::lock_guard<std::mutex> lk(mu);
                if (!ctx) ctx = pythia::acquire_device(knobs.device, knobs.hbmbudget);
            }
            pgpu_hj_desc d{};
            Schema& ps = probeOp->getOutSchema();
            Schema& bs = buildOp->getOutSchema();
            // staged build rows are sbuild (key + projected build columns);
            // resident build rows are the scan's own tuples
            d.build_row_bytes = resident_ ? bs.getTupleSize() : sbuild.getTupleSize();
            d.build_key_offset = resident_ ? bs.getColumnOffset(joinattr1) : 0;
            d.build_key_bytes = sbuild.getColumnWidth(0);
            d.probe_row_bytes = ps.getTupleSize();
            d.probe_key_offset = ps.getColumnOffset(joinattr2);
            d.probe_key_bytes = ps.getColumnWidth(joinattr2);
            d.build_key_type = key_type(bs.getColumnType(joinattr1));
            d.probe_key_type = key_type(ps.getColumnType(joinattr2));
            d.nproj = (uint32_t)projection.size();
            if (d.nproj > PGPU_MAX_PROJ) throw InvalidParameter("too many projected columns");
            unsigned buildcol = 1;
            for (uint32_t j = 0; j < d.nproj; ++j) {
                if (projection[j].first == BuildSide) {
                    if (resident_)
                        d.proj[j] = {0, bs.getColumnOffset(projection[j].second), bs.getColumnWidth(projection[j].second)};
                    else
                        d.proj[j] = {0, sbuild.getColumnOffset(buildcol), sbuild.getColumnWidth(buildcol)};
                    ++buildcol;
                } else {
                    d.proj[j] = {1, ps.getColumnOffset(projection[j].second), ps.getColumnWidth(projection[j].second)};
                }
            }
            d.slots = MAX_THREADS;
            check(pgpu_hj_create(ctx, &d, &tables[g]), "pgpu_hj_create");
        } catch (...) {
            fail_group(g);
            throw;
        }
    }
    if (!barriers[g]->Arrive()) throw QueryExecutionError("gpuhashjoin: thread group failed during threadInit");
    if (buffsize < schema.getTupleSize()) throw InvalidParameter("buffsize smaller than one output tuple");
    state[tid] = ThreadState();
    state[tid].out = new Page(buffsize, schema.getTupleSize());
}

Operator::ResultCode GpuHashJoinOp::scanStart(unsigned short tid, Page* idx, Schema& idxschema) {
    if (resident_) return scanStartResident(tid, idx, idxschema);
    const unsigned short g = threadgroups[tid];
    Schema& bs = buildOp->getOutSchema();
    if (buildOp->scanStart(tid, idx, idxschema) == Error) {
        fail_group(g);
        return Error;
    }
    // Build: project each build tuple into sbuild and stage it (buildFromPage,
    // operators/join.cpp:641-667); the library copies before returning.
    const unsigned sb = sbuild.getTupleSize();
    std::vector<unsigned> src_off, width;
    src_off.push_back(bs.getColumnOffset(joinattr1));
    width.push_back(bs.getColumnWidth(joinattr1));
    for (auto& p : projection)
        if (p.first == BuildSide) {
            src_off.push_back(bs.getColumnOffset(p.second));
            width.push_back(bs.getColumnWidth(p.second));
        }
    std::vector<char> rows;
    GetNextResultT r{Ready, nullptr};
    try {
        while (r.first == Ready) {
            r = buildOp->getNext(tid);
            if (r.first == Error) break;
            const uint64_t n = r.second->getNumTuples();
            const size_t base = rows.size();
            rows.resize(base + n * sb);
            char* dst = rows.data() + base;
            const char* src = r.second->data();
            const unsigned ts = bs.getTupleSize();
            for (uint64_t i = 0; i < n; ++i, src += ts, dst += sb)
                for (size_t c = 0, o = 0; c < src_off.size(); o += width[c], ++c) memcpy(dst + o, src + src_off[c], width[c]);
            if (rows.size() >= (32u << 20)) {
                check(pgpu_hj_build_append(tables[g], rows.data(), rows.size() / sb), "pgpu_hj_build_append");
                rows.clear();
            }
        }
        if (r.first != Error && !rows.empty())
            check(pgpu_hj_build_append(tables[g], rows.data(), rows.size() / sb), "pgpu_hj_build_append");
    } catch (...) {
        fail_group(g);
        throw;
    }
    if (r.first == Error || buildOp->scanStop(tid) == Error) {
        fail_group(g);
        return Error;
    }
    // Every thread of the group has staged its build tuples (join.cpp:434).
    if (!barriers[g]->Arrive()) return Error;
    if (groupleader[g] == tid) {
        int rc = pgpu_hj_build_finish(tables[g]);
        if (rc != PGPU_OK) {
            fail_group(g);
            check(rc, "pgpu_hj_build_finish");
        }
    }
    if (!barriers[g]->Arrive()) return Error;

    ResultCode rc = probeOp->scanStart(tid, idx, idxschema);
    if (rc == Error) return Error;
    ThreadState& st = state[tid];
    Page* keep = st.out;
    st = ThreadState();
    st.out = keep;
    // The first probe tuple is read here; an empty probe side is Finished
    // (operators/join.cpp:444-459).
    if (!stageProbe(tid)) {
        st.done = true;
        return Finished;
    }
    return rc;
}

bool GpuHashJoinOp::stageProbe(unsigned short tid) {
    ThreadState& st = state[tid];
    st.probe.clear();
    st.nprobe = 0;
    const unsigned ts = probeOp->getOutSchema().getTupleSize();
    while (!st.depleted && st.nprobe < knobs.probebatch) {
        GetNextResultT r = probeOp->getNext(tid);
        // readNextTupleFromProbe throws on a child error (operators/join.cpp:584-586).
        if (r.first == Error) throw QueryExecutionError("gpuhashjoin: probe input returned Error");
        const uint64_t n = r.second->getNumTuples();
        st.probe.insert(st.probe.end(), r.second->data(), r.second->data() + n * ts);
        st.nprobe += n;
        if (r.first == Finished) st.depleted = true;
    }
    return st.nprobe > 0;
}

Operator::GetNextResultT GpuHashJoinOp::getNext(unsigned short tid) {
    ThreadState& st = state[tid];
    Page* out = st.out;
    out->clear();
    if (st.done) return {Finished, out};
    const unsigned ts = schema.getTupleSize();
    const uint64_t cap = out->capacity() / ts;
    pgpu_hj* t = tables[threadgroups[tid]];
    while (resident_) {
        if (st.pending) {
            const uint64_t used = out->getNumTuples();
            uint64_t got = 0;
            check(pgpu_hj_pull(t, tid, out->data() + used * ts, std::min(cap - used, st.pending), &got),
                  "pgpu_hj_pull");
            pulled_bytes_ += got * ts;
            out->setUsed((used + got) * ts);
            st.pending -= got;
            if (out->getNumTuples() == cap) return {Ready, out};
            continue;
        }
        uint64_t nout = 0;
        if (!probeNextResident(tid, &nout)) {
            st.done = true;
            return {Finished, out};
        }
        st.pending = nout;
    }
    while (true) {
        if (st.pending) {
            const uint64_t used = out->getNumTuples();
            uint64_t got = 0;
            check(pgpu_hj_pull(t, tid, out->data() + used * ts, std::min(cap - used, st.pending), &got),
                  "pgpu_hj_pull");
            pulled_bytes_ += got * ts;
            out->setUsed((used + got) * ts);
            st.pending -= got;
            if (out->getNumTuples() == cap) return {Ready, out};
            continue;
        }
        if (st.nprobe) {
            uint64_t nout = 0;
            check(pgpu_hj_probe(t, tid, st.probe.data(), st.nprobe, &nout), "pgpu_hj_probe");
            st.pending = nout;
            st.nprobe = 0;
            continue;
        }
        if (!stageProbe(tid)) {
            st.done = true;
            return {Finished, out};
        }
    }
}

Operator::ResultCode GpuHashJoinOp::scanStop(unsigned short tid) { return probeOp->scanStop(tid); }

// Resident mode: the leader builds the table straight from the device copy
// of the build relation; the probe relation's device copy is then consumed
// in probebatch chunks from a shared cursor.  The child scans only see
// scanStart / scanStop (their pages are not read).
Operator::ResultCode GpuHashJoinOp::scanStartResident(unsigned short tid, Page* idx, Schema& idxschema) {
    const unsigned short g = threadgroups[tid];
    if (fused_ == FuseCountSum || fused_ == FuseGroup) {
        // the consumer runs the whole join through claimRun(): no table to
        // build, no barrier
        ++runs_[tid];
        if (buildOp->scanStart(tid, idx, idxschema) == Error || buildOp->scanStop(tid) == Error) return Error;
        ResultCode rc = probeOp->scanStart(tid, idx, idxschema);
        if (rc == Error) return Error;
        uint64_t np = 0;
        shared_size(probeOp, &np);
        ThreadState& st = state[tid];
        Page* keep = st.out;
        st = ThreadState();
        st.out = keep;
        if (np == 0) {  // an empty probe side is Finished (operators/join.cpp:444-459)
            st.done = true;
            return Finished;
        }
        return rc;
    }
    if (buildOp->scanStart(tid, idx, idxschema) == Error || buildOp->scanStop(tid) == Error) {
        fail_group(g);
        return Error;
    }
    if (groupleader[g] == tid) {
        try {
            auto* b = dynamic_cast<ParallelScanOp*>(buildOp);
            auto* p = dynamic_cast<ParallelScanOp*>(probeOp);
            const char* rows;
            uint64_t n;
            std::string key;
            b->sharedRows(&rows, &n, &key);
            const unsigned bts = buildOp->getOutSchema().getTupleSize();
            d_build_ = (const char*)pythia::resident_rows(knobs.device, knobs.hbmbudget, key, rows, n * bts);
            nbuild_rows_ = n;
            build_key_ = key;
            check(pgpu_hj_build_device(tables[g], d_build_, n), "pgpu_hj_build_device");
            p->sharedRows(&rows, &n, &key);
            const unsigned pts = probeOp->getOutSchema().getTupleSize();
            d_probe_ = (const char*)pythia::resident_rows(knobs.device, knobs.hbmbudget, key, rows, n * pts);
            nprobe_rows_ = n;
            probe_key_ = key;
            probe_cursor_ = 0;
        } catch (...) {
            fail_group(g);
            throw;
        }
    }
    if (!barriers[g]->Arrive()) return Error;
    ResultCode rc = probeOp->scanStart(tid, idx, idxschema);
    if (rc == Error) return Error;
    ThreadState& st = state[tid];
    Page* keep = st.out;
    st = ThreadState();
    st.out = keep;
    if (nprobe_rows_ == 0) {  // an empty probe side is Finished (operators/join.cpp:444-459)
        st.done = true;
        return Finished;
    }
    return rc;
}

// Joins the next probebatch chunk of the resident probe relation on thread
// tid's slot; false when the relation is exhausted.
bool GpuHashJoinOp::probeNextResident(unsigned short tid, uint64_t* nout) {
    const uint64_t c = probe_cursor_.fetch_add(knobs.probebatch);
    if (c >= nprobe_rows_) return false;
    const uint64_t n = std::min<uint64_t>(knobs.probebatch, nprobe_rows_ - c);
    const unsigned pts = probeOp->getOutSchema().getTupleSize();
    check(pgpu_hj_probe_device(tables[threadgroups[tid]], tid, d_probe_ + c * pts, n, nout), "pgpu_hj_probe_device");
    return true;
}

bool GpuHashJoinOp::nextDeviceBatch(unsigned short tid, const void** d_rows, uint64_t* n) {
    ThreadState& st = state[tid];
    if (st.done) return false;
    uint64_t nout = 0;
    if (!probeNextResident(tid, &nout)) {
        st.done = true;
        return false;
    }
    check(pgpu_hj_output(tables[threadgroups[tid]], tid, d_rows, n), "pgpu_hj_output");
    return true;
}

bool GpuHashJoinOp::canFuseCountSum(int outcol) const {
    if (!resident_) return false;
    Schema& bs = buildOp->getOutSchema();
    Schema& ps = probeOp->getOutSchema();
    const unsigned kb = bs.getColumnWidth(joinattr1);
    if (!is_int_key(bs.getColumnType(joinattr1)) || !is_int_key(ps.getColumnType(joinattr2))) return false;
    if ((kb != 4 && kb != 8) || ps.getColumnWidth(joinattr2) != kb) return false;
    uint64_t nb = 0, np = 0;  // the radix plan takes fewer than 2^32 rows a side
    if (!shared_size(buildOp, &nb) || !shared_size(probeOp, &np) || nb >= (1ull << 32) || np >= (1ull << 32))
        return false;
    if (outcol < 0) return true;
    if ((size_t)outcol >= projection.size()) return false;
    const JoinPrjT& c = projection[outcol];
    Schema& s = c.first == BuildSide ? bs : ps;
    const ColumnType t = s.getColumnType(c.second);
    return s.getColumnWidth(c.second) == kb && (t == CT_INTEGER || t == CT_LONG);
}

bool GpuHashJoinOp::canFuseGroup(int groupcol, int sumcol) const {
    if (!resident_ || groupcol < 0 || sumcol < 0 || (size_t)groupcol >= projection.size() ||
        (size_t)sumcol >= projection.size())
        return false;
    Schema& bs = buildOp->getOutSchema();
    Schema& ps = probeOp->getOutSchema();
    if (bs.getColumnType(joinattr1) != CT_INTEGER || ps.getColumnType(joinattr2) != CT_INTEGER) return false;
    const JoinPrjT& g = projection[groupcol];
    const JoinPrjT& v = projection[sumcol];
    if (g.first != BuildSide || bs.getColumnType(g.second) != CT_INTEGER) return false;
    if (v.first != ProbeSide || ps.getColumnType(v.second) != CT_DECIMAL) return false;
    uint64_t nb = 0, np = 0;  // the group join's dimension side stays under 2^31 rows
    return shared_size(buildOp, &nb) && shared_size(probeOp, &np) && nb < (1ull << 31);
}

const void* GpuHashJoinOp::residentColumn(bool build, unsigned col) {
    Schema& s = build ? buildOp->getOutSchema() : probeOp->getOutSchema();
    return pythia::resident_column(knobs.device, knobs.hbmbudget, build ? build_key_ : probe_key_,
                                   build ? d_build_ : d_probe_, build ? nbuild_rows_ : nprobe_rows_,
                                   s.getTupleSize(), s.getColumnOffset(col), s.getColumnWidth(col));
}

bool GpuHashJoinOp::claimRun(unsigned short tid) {
    // a run nobody claimed (an empty probe side returns before claiming)
    // leaves claimed_run_ behind: any older value is claimable
    const uint64_t run = runs_[tid];
    uint64_t cur = claimed_run_.load();
    while (cur < run)
        if (claimed_run_.compare_exchange_weak(cur, run)) return true;
    return false;
}

void GpuHashJoinOp::residentInputs() {
    const char* rows;
    uint64_t n;
    std::string key;
    dynamic_cast<ParallelScanOp*>(buildOp)->sharedRows(&rows, &n, &key);
    d_build_ = (const char*)pythia::resident_rows(knobs.device, knobs.hbmbudget, key, rows,
                                                  n * buildOp->getOutSchema().getTupleSize());
    nbuild_rows_ = n;
    build_key_ = key;
    dynamic_cast<ParallelScanOp*>(probeOp)->sharedRows(&rows, &n, &key);
    d_probe_ = (const char*)pythia::resident_rows(knobs.device, knobs.hbmbudget, key, rows,
                                                  n * probeOp->getOutSchema().getTupleSize());
    nprobe_rows_ = n;
    probe_key_ = key;
}

// The queryhashjoin plan's join -> aggregate_count / aggregate_sum with no
// GROUP BY (operators/join.cpp:503-562 feeding aggregatesum.cpp:81-107 /
// aggregatecount.cpp:44-52) as one 2-pass radix join with a fused count +
// sum: the summed column's values travel through the partitions as the
// payload of their side; an output row's value is counted once per match.
GpuHashJoinOp::CountSum GpuHashJoinOp::countSum(int outcol) {
    std::lock_guard<std::mutex> lk(fuse_mu_);
    residentInputs();
    CountSum cs;
    if (!nbuild_rows_ || !nprobe_rows_) return cs;
    const void* rk = residentColumn(true, joinattr1);
    const void* sk = residentColumn(false, joinattr2);
    const void* rp = rk;  // a count still streams a build payload: the key itself
    const void* sp = nullptr;
    bool probe_side = false;
    if (outcol >= 0) {
        const JoinPrjT& c = projection[outcol];
        if (c.first == BuildSide)
            rp = residentColumn(true, c.second);
        else {
            sp = residentColumn(false, c.second);
            probe_side = true;
        }
    }
    pgpu_join_cfg cfg{};
    cfg.key_bytes = buildOp->getOutSchema().getColumnWidth(joinattr1);
    cfg.probe_payload = probe_side ? 1 : 0;
    cfg.algo = PGPU_ALGO_AUTO;
    pgpu_join_result r{};
    pythia::with_join_plan(knobs.device, knobs.hbmbudget, cfg, nbuild_rows_, nprobe_rows_, [&](pgpu_join_plan* plan) {
        check(pgpu_join_count_sum(plan, rk, rp, nbuild_rows_, sk, sp, nprobe_rows_, &r), "pgpu_join_count_sum");
    });
    cs.matches = r.matches;
    cs.sum = probe_side ? r.sum_probe : r.sum_build;
    return cs;
}

// hashjoin(build D, probe F) -> aggregate_sum(field = a D column, sumfield =
// an F dec column) (operators/join.cpp:503-562 feeding
// genericaggregate.cpp:240-282) as the partitioned group join; its result
// rows are the aggregate's output rows.
const char* GpuHashJoinOp::groupRows(unsigned short tid, int groupcol, int sumcol, uint64_t* n) {
    const uint64_t run = runs_[tid];
    if (!claimRun(tid)) {
        std::unique_lock<std::mutex> lk(fuse_mu_);
        fuse_cv_.wait(lk, [&] { return published_run_ >= run; });
        if (!run_error_.empty()) throw QueryExecutionError(run_error_);
        *n = published_rows_;
        return fuse_rows_;
    }
    // The claimer alone touches the inputs and the pinned rows.  It publishes
    // the group count as soon as the join has run, then writes the rows out
    // in order, ~3 MiB at a time, so the first threads' shares are emitted
    // (and consumed) while the later ones cross the host link.
    uint64_t ng = 0;
    bool published = false;
    auto publish = [&](const std::string& err, uint64_t rows, uint64_t copied) {
        {
            std::lock_guard<std::mutex> lk(fuse_mu_);
            run_error_ = err;
            published_rows_ = rows;
            copied_rows_ = copied;
            published_run_ = run;
        }
        fuse_cv_.notify_all();
        published = true;
    };
    try {
        residentInputs();
        if (nbuild_rows_ && nprobe_rows_) {
            const auto* dk = (const int32_t*)residentColumn(true, joinattr1);
            const auto* dg = (const int32_t*)residentColumn(true, projection[groupcol].second);
            const auto* fk = (const int32_t*)residentColumn(false, joinattr2);
            const auto* fv = (const double*)residentColumn(false, projection[sumcol].second);
            pythia::with_groupjoin(knobs.device, knobs.hbmbudget, nbuild_rows_, nprobe_rows_, nbuild_rows_,
                                   [&](pgpu_groupjoin* gj) {
                                       check(pgpu_groupjoin_run(gj, dk, dg, nbuild_rows_, fk, fv, nprobe_rows_, &ng,
                                                                nullptr),
                                             "pgpu_groupjoin_run");
                                       if (ng * 12 > fuse_rows_cap_) {
                                           pythia::release_pinned(fuse_rows_);
                                           fuse_rows_ = nullptr;
                                           fuse_rows_cap_ = 0;
                                           fuse_rows_ = (char*)pythia::acquire_pinned(knobs.device, knobs.hbmbudget,
                                                                                     ng * 12);
                                           fuse_rows_cap_ = ng * 12;
                                       }
                                       publish(std::string(), ng, 0);
                                       // chunks of ~3 MiB starting on 4-row (48-byte) boundaries, so
                                       // each lands 16-byte aligned in the pinned block
                                       const uint64_t C = std::max<uint64_t>(1, std::min<uint64_t>(32, ng / 262144));
                                       for (uint64_t c = 0; c < C; ++c) {
                                           const uint64_t a = c * ng / C & ~3ull;
                                           const uint64_t b = c + 1 == C ? ng : (c + 1) * ng / C & ~3ull;
                                           if (b > a)
                                               check(pgpu_groupjoin_rows(gj, PGPU_GJ_ROW_SUM, a, b - a,
                                                                         fuse_rows_ + a * 12),
                                                     "pgpu_groupjoin_rows");
                                           {
                                               std::lock_guard<std::mutex> lk(fuse_mu_);
                                               copied_rows_ = b;
                                           }
                                           fuse_cv_.notify_all();
                                       }
                                   });
        }
        if (!published) publish(std::string(), 0, 0);
    } catch (const std::exception& e) {
        if (published) {
            {
                std::lock_guard<std::mutex> lk(fuse_mu_);
                run_error_ = e.what();
            }
            fuse_cv_.notify_all();
        } else {
            publish(e.what(), 0, 0);
        }
        throw QueryExecutionError(e.what());
    }
    *n = ng;
    return fuse_rows_;
}

void GpuHashJoinOp::waitGroupRows(uint64_t upto) {
    std::unique_lock<std::mutex> lk(fuse_mu_);
    fuse_cv_.wait(lk, [&] { return copied_rows_ >= upto || !run_error_.empty(); });
    if (!run_error_.empty()) throw QueryExecutionError(run_error_);
}

void GpuHashJoinOp::threadClose(unsigned short tid) {
    delete state[tid].out;
    state[tid].out = nullptr;
    const unsigned short g = threadgroups[tid];
    barriers[g]->Arrive();  // a broken barrier returns at once
    if (groupleader[g] == tid) {
        pgpu_hj_destroy(tables[g]);
        tables[g] = nullptr;
    }
}

void GpuHashJoinOp::destroy() {
    for (auto*& t : tables) {
        if (t) pgpu_hj_destroy(t);
        t = nullptr;
    }
    for (auto* b : barriers) delete b;
    barriers.clear();
    pythia::release_pinned(fuse_rows_);  // consumers have emitted every row by now
    fuse_rows_ = nullptr;
    fuse_rows_cap_ = 0;
    if (ctx) pythia::release_device(ctx);
    ctx = nullptr;
}

pgpu_hj_stats GpuHashJoinOp::stats(unsigned short group) {
    pgpu_hj_stats s{};
    if (group < tables.size() && tables[group]) pgpu_hj_stats_get(tables[group], &s);
    return s;
}

// ====================================================================== aggregation
void GpuAggregate::init(libconfig::Config& root, libconfig::Setting& cfg) {
    Operator::init(root, cfg);
    if (cfg.exists("fields")) {
        libconfig::Setting& f = cfg["fields"];
        for (int i = 0; i < f.getLength(); ++i) aggfields.push_back((unsigned short)(int)f[i]);
    } else {
        aggfields.push_back((unsigned short)(int)cfg["field"]);
    }
    if (aggfields.empty() && !cfg.exists("hash")) {
        libconfig::Setting& h = cfg.add("hash", libconfig::Setting::TypeGroup);
        h.add("fn", libconfig::Setting::TypeString) = "alwayszero";
    }
    Schema& in = nextOp->getOutSchema();
    Schema uds = foldinit(cfg);
    for (auto f : aggfields) {
        if (f >= in.columns()) throw InvalidParameter("group column out of range");
        schema.add(in.get(f));
    }
    for (unsigned i = 0; i < uds.columns(); ++i) schema.add(uds.get(i));
    if (cfg.exists("presorted")) throw NotYetImplemented();
    // the hash node names its own field / fieldrange (hash.cpp:120-137)
    libconfig::Setting& h = cfg["hash"];
    std::string fn = h["fn"];
    int fmin = 0, fmax = 0;
    if (fn != "alwayszero") {
        if (h.exists("fieldrange")) {
            fmin = h["fieldrange"][0];
            fmax = h["fieldrange"][1];
        } else {
            fmin = fmax = (int)h["field"];
        }
    }
    validate_hash(h, in, fmin, fmax);
    if (cfg.exists("global")) {
        global_ = true;
        threads_ = (int)cfg["threads"];
        barrier_.init(threads_);
    }
    knobs.read(cfg);
    state.assign(MAX_THREADS, ThreadState());
    // Directly over a resident gpuhashjoin on the same device the join and
    // this aggregate run as one device computation (see FuseKind).
    auto* join = dynamic_cast<GpuHashJoinOp*>(nextOp);
    if (join && join->residentMode() && join->device() == knobs.device) {
        join_ = join;
        fuse_ = GpuHashJoinOp::FuseRows;
        if (aggfields.empty() && join->canFuseCountSum(sumField()))
            fuse_ = GpuHashJoinOp::FuseCountSum;
        else if (aggfields.size() == 1 && kind() == PGPU_AGG_SUM_F64 && join->canFuseGroup(aggfields[0], sumField()))
            fuse_ = GpuHashJoinOp::FuseGroup;
        join->setFused(fuse_);
    }
}

pgpu_agg* GpuAggregate::newTable() {
    {
        std::lock_guard<std::mutex> lk(mu);
        if (!ctx) ctx = pythia::acquire_device(knobs.device, knobs.hbmbudget);
    }
    Schema& in = nextOp->getOutSchema();
    pgpu_agg_desc d{};
    d.row_bytes = in.getTupleSize();
    d.nkeys = (uint32_t)aggfields.size();
    if (d.nkeys > 8) throw InvalidParameter("at most 8 group columns on the GPU");
    for (uint32_t i = 0; i < d.nkeys; ++i) {
        d.key_offset[i] = in.getColumnOffset(aggfields[i]);
        d.key_bytes[i] = in.getColumnWidth(aggfields[i]);
    }
    d.kind = kind();
    d.val_offset = valOffset();
    pgpu_agg* a = nullptr;
    check(pgpu_agg_create(ctx, &d, &a), "pgpu_agg_create");
    return a;
}

void GpuAggregate::threadInit(unsigned short tid) {
    if (buffsize < schema.getTupleSize()) throw InvalidParameter("buffsize smaller than one output tuple");
    state[tid] = ThreadState();
    state[tid].out = new Page(buffsize, schema.getTupleSize());
    if (fuse_ == GpuHashJoinOp::FuseCountSum || fuse_ == GpuHashJoinOp::FuseGroup) return;  // no group table
    if (!global_) {
        state[tid].agg = newTable();
        return;
    }
    try {
        std::lock_guard<std::mutex> lk(init_mu);  // newTable() takes `mu`
        if (!shared_) shared_ = newTable();
    } catch (...) {
        barrier_.break_();
        throw;
    }
    if (!barrier_.Arrive()) throw QueryExecutionError("gpuaggregate: a thread failed during threadInit");
}

// The fused forms: the one thread that claims this run of the plan runs the
// whole join below.  A count / sum without GROUP BY is then one row, emitted
// by that thread (nothing when no tuple joined: no group exists); the other
// threads emit nothing.  The group join's groups are split evenly over the
// join's threads, each emitted once (Global mode's rule, each group by one
// thread, genericaggregate.cpp:324-412; in ThreadLocal mode each group is
// then one thread's partial).
Operator::ResultCode GpuAggregate::scanStartFused(unsigned short tid, Page* idx, Schema& idxschema) {
    ThreadState& st = state[tid];
    st.row.clear();
    st.view.reset();
    st.rows = nullptr;
    st.first = st.end = st.pos = 0;
    ResultCode rc = nextOp->scanStart(tid, idx, idxschema);
    if (rc != Ready) return rc;  // an empty probe side joins nothing
    if (fuse_ == GpuHashJoinOp::FuseCountSum && join_->claimRun(tid)) {
        const GpuHashJoinOp::CountSum cs = join_->countSum(sumField());
        if (cs.matches) {
            st.row.resize(schema.getTupleSize());
            if (kind() == PGPU_AGG_SUM_I32) {
                const int32_t v = (int32_t)(uint32_t)cs.sum;  // the CtInt accumulator wraps mod 2^32
                memcpy(st.row.data(), &v, 4);
            } else {
                const int64_t v = kind() == PGPU_AGG_COUNT ? (int64_t)cs.matches : (int64_t)cs.sum;
                memcpy(st.row.data(), &v, 8);
            }
        }
    }
    if (fuse_ == GpuHashJoinOp::FuseGroup) {
        uint64_t n = 0;
        st.rows = join_->groupRows(tid, aggfields[0], sumField(), &n);
        const uint64_t T = join_->groupThreads(), j = join_->groupPosition(tid);
        st.first = st.pos = j * n / T;
        st.end = (j + 1) * n / T;
    }
    return nextOp->scanStop(tid);
}

Operator::ResultCode GpuAggregate::scanStart(unsigned short tid, Page* idx, Schema& idxschema) {
    if (fuse_ == GpuHashJoinOp::FuseCountSum || fuse_ == GpuHashJoinOp::FuseGroup)
        return scanStartFused(tid, idx, idxschema);
    ThreadState& st = state[tid];
    ResultCode rc = nextOp->scanStart(tid, idx, idxschema);
    if (rc != Ready) {
        if (global_) barrier_.break_();
        return rc;
    }
    pgpu_agg* t = global_ ? shared_ : st.agg;
    const unsigned ts = nextOp->getOutSchema().getTupleSize();
    std::vector<char> rows;
    GetNextResultT r{Ready, nullptr};
    // Over a resident gpuhashjoin on the same device the join's output rows
    // are folded where they are, in HBM (no pages cross the host link).
    const bool fused = fuse_ == GpuHashJoinOp::FuseRows;
    try {
        if (fused) {
            const void* d = nullptr;
            uint64_t n = 0;
            while (join_->nextDeviceBatch(tid, &d, &n))
                if (n) check(pgpu_agg_append_device(t, d, n), "pgpu_agg_append_device");
            r.first = Finished;
        }
        while (!fused) {
            r = nextOp->getNext(tid);
            if (r.first == Error) break;
            const uint64_t n = r.second->getNumTuples();
            rows.insert(rows.end(), r.second->data(), r.second->data() + n * ts);
            if (rows.size() >= (64u << 20)) {
                check(pgpu_agg_append(t, rows.data(), rows.size() / ts), "pgpu_agg_append");
                rows.clear();
            }
            if (r.first != Ready) break;
        }
        if (r.first != Error && !rows.empty())
            check(pgpu_agg_append(t, rows.data(), rows.size() / ts), "pgpu_agg_append");
    } catch (...) {
        if (global_) barrier_.break_();
        throw;
    }
    rc = nextOp->scanStop(tid);
    if (r.first == Error) {
        if (global_) barrier_.break_();
        return Error;
    }
    uint64_t ng = 0;
    if (global_) {
        if (!barrier_.Arrive()) return Error;  // every thread has folded its input
        check(pgpu_agg_finish(t, &ng, nullptr), "pgpu_agg_finish");  // idempotent
        // each group is emitted by exactly one thread (operators/genericaggregate.cpp:324-412)
        st.first = (uint64_t)tid * ng / threads_;
        st.end = (uint64_t)(tid + 1) * ng / threads_;
    } else {
        check(pgpu_agg_finish(t, &ng, nullptr), "pgpu_agg_finish");
        st.first = 0;
        st.end = ng;
    }
    st.pos = st.first;
    st.done = false;
    return rc;
}

Operator::GetNextResultT GpuAggregate::getNext(unsigned short tid) {
    ThreadState& st = state[tid];
    Page* out = st.out;
    out->clear();
    if (fuse_ == GpuHashJoinOp::FuseCountSum) {
        if (!st.row.empty()) {
            memcpy(out->data(), st.row.data(), st.row.size());
            out->setUsed(st.row.size());
            st.row.clear();
        }
        return {Finished, out};
    }
    if (fuse_ == GpuHashJoinOp::FuseGroup) {
        // overlay pages over the join's pinned result rows (valid until destroy)
        if (st.pos >= st.end) return {Finished, out};
        const unsigned ts = schema.getTupleSize();
        const uint64_t k = std::min<uint64_t>(out->capacity() / ts, st.end - st.pos);
        join_->waitGroupRows(st.pos + k);
        st.view.reset(new Page((void*)(st.rows + st.pos * ts), k * ts, nullptr, ts));
        st.pos += k;
        return {st.pos >= st.end ? Finished : Ready, st.view.get()};
    }
    if (st.pos >= st.end) return {Finished, out};
    const unsigned ts = schema.getTupleSize();
    const uint64_t k = std::min<uint64_t>(out->capacity() / ts, st.end - st.pos);
    check(pgpu_agg_pull(global_ ? shared_ : st.agg, st.pos, k, out->data()), "pgpu_agg_pull");
    out->setUsed(k * ts);
    st.pos += k;
    return {st.pos >= st.end ? Finished : Ready, out};
}

void GpuAggregate::threadClose(unsigned short tid) {
    delete state[tid].out;
    state[tid].out = nullptr;
    state[tid].view.reset();
    if (fuse_ == GpuHashJoinOp::FuseCountSum || fuse_ == GpuHashJoinOp::FuseGroup) return;
    if (!global_) {
        pgpu_agg_destroy(state[tid].agg);
        state[tid].agg = nullptr;
        return;
    }
    barrier_.Arrive();
    std::lock_guard<std::mutex> lk(mu);
    if (shared_) {
        pgpu_agg_destroy(shared_);
        shared_ = nullptr;
    }
}

void GpuAggregate::destroy() {
    for (auto& s : state)
        if (s.agg) {
            pgpu_agg_destroy(s.agg);
            s.agg = nullptr;
        }
    if (shared_) pgpu_agg_destroy(shared_);
    shared_ = nullptr;
    if (ctx) pythia::release_device(ctx);
    ctx = nullptr;
}

Schema GpuAggregateSum::foldinit(libconfig::Setting& node) {
    unsigned sumfield = (int)node["sumfield"];
    Schema& in = nextOp->getOutSchema();
    if (sumfield >= in.columns()) throw InvalidParameter("sumfield out of range");
    Schema s;
    switch (in.getColumnType(sumfield)) {
        case CT_INTEGER: kind_ = PGPU_AGG_SUM_I32; break;
        case CT_LONG: kind_ = PGPU_AGG_SUM_I64; break;
        case CT_DECIMAL: kind_ = PGPU_AGG_SUM_F64; break;
        default: throw InvalidParameter("aggregate_sum sums int, long or dec columns");
    }
    valoff_ = in.getColumnOffset(sumfield);
    sumfield_ = (int)sumfield;
    s.add(in.getColumnType(sumfield));
    return s;
}

Schema GpuAggregateCount::foldinit(libconfig::Setting&) {
    Schema s;
    s.add(CT_LONG);
    return s;
}

// ------------------------------------------------------------------ gpupartition
void GpuPartitionOp::init(libconfig::Config& root, libconfig::Setting& cfg) {
    Operator::init(root, cfg);
    schema = nextOp->getOutSchema();
    attribute_ = (unsigned)(int)cfg["attr"];
    if (attribute_ >= schema.columns()) throw InvalidParameter("partition: attr out of range");
    // exactrange as TupleHasher::create builds it (hash.cpp:182-187); the
    // hashed field goes through numericalize (hash.h:215-234): 4 or 8 bytes
    rmin_ = (int)cfg["range"][0];
    rmax_ = (int)cfg["range"][1];
    const int b = (int)cfg["buckets"];
    if (b < 1 || b >= MAX_THREADS) throw InvalidParameter("partition: buckets must be in 1..127 (one thread each)");
    if (rmax_ < rmin_) throw InvalidParameter("partition: range max below min");
    buckets_ = (unsigned)b;
    const unsigned w = schema.getColumnWidth(attribute_);
    if (w != 4 && w != 8) throw IllegalConversionException();
    if (cfg.exists("maxtuplesinM"))
        (void)(int)cfg["maxtuplesinM"];
    else
        (void)(int)cfg["maxtuples"];
    std::string sort = cfg["sort"];
    sort_ = sort == "yes";
    sortattr_ = attribute_;
    if (sort_) {
        int sa = (int)attribute_;
        cfg.lookupValue("sortattr", sa);
        if (sa < 0 || (unsigned)sa >= schema.columns()) throw InvalidParameter("partition: sortattr out of range");
        sortattr_ = (unsigned)sa;
        // sortAllInPage sorts int, long / date and dec keys (partition.cpp:228-249)
        const ColumnType t = schema.getColumnType(sortattr_);
        if (t != CT_INTEGER && t != CT_LONG && t != CT_DATE && t != CT_DECIMAL) throw NotYetImplemented();
    }
    barrier_.init((int)buckets_);
    knobs.read(cfg);
    counts_.assign(buckets_, 0);
    state.clear();
    state.resize(MAX_THREADS);
}

void GpuPartitionOp::threadInit(unsigned short tid) {
    if (tid >= buckets_) throw InvalidParameter("partition: run on exactly `buckets` threads (thread id >= buckets)");
    std::lock_guard<std::mutex> lk(mu);
    if (!ctx) ctx = pythia::acquire_device(knobs.device, knobs.hbmbudget);
    if (!part_) {
        pgpu_part_desc d{};
        d.row_bytes = schema.getTupleSize();
        d.key_offset = schema.getColumnOffset(attribute_);
        d.key_bytes = schema.getColumnWidth(attribute_);
        d.buckets = buckets_;
        d.range_min = rmin_;
        d.range_max = rmax_;
        d.sort = sort_ ? 1 : 0;
        if (sort_) {
            d.sort_offset = schema.getColumnOffset(sortattr_);
            const ColumnType t = schema.getColumnType(sortattr_);
            d.sort_type = t == CT_INTEGER ? PGPU_COL_INT32 : t == CT_DECIMAL ? PGPU_COL_F64 : PGPU_COL_INT64;
        }
        check(pgpu_part_create(ctx, &d, &part_), "pgpu_part_create");
    }
    state[tid].rows.clear();
    state[tid].n = state[tid].pos = 0;
}

Operator::ResultCode GpuPartitionOp::scanStart(unsigned short tid, Page* idx, Schema& idxschema) {
    ThreadState& st = state[tid];
    ResultCode rc = nextOp->scanStart(tid, idx, idxschema);
    if (rc != Ready) {
        barrier_.break_();
        return rc;
    }
    const unsigned ts = schema.getTupleSize();
    std::vector<char> rows;
    GetNextResultT r{Ready, nullptr};
    try {
        do {
            r = nextOp->getNext(tid);
            if (r.first == Error) break;
            const uint64_t n = r.second->getNumTuples();
            rows.insert(rows.end(), r.second->data(), r.second->data() + n * ts);
            if (rows.size() >= (64u << 20)) {
                check(pgpu_part_append(part_, tid, rows.data(), rows.size() / ts), "pgpu_part_append");
                rows.clear();
            }
        } while (r.first == Ready);
        if (r.first != Error && !rows.empty())
            check(pgpu_part_append(part_, tid, rows.data(), rows.size() / ts), "pgpu_part_append");
    } catch (...) {
        barrier_.break_();
        throw;
    }
    rc = nextOp->scanStop(tid);
    if (r.first == Error || rc != Ready) {
        barrier_.break_();
        return Error;
    }
    if (!barrier_.Arrive()) return Error;  // every thread has staged its input
    if (tid == 0) {
        try {
            check(pgpu_part_finish(part_, counts_.data()), "pgpu_part_finish");
        } catch (...) {
            barrier_.break_();
            throw;
        }
    }
    if (!barrier_.Arrive()) return Error;  // the partitions exist
    st.n = counts_[tid];
    st.pos = 0;
    st.rows.resize(st.n * ts);
    check(pgpu_part_pull(part_, tid, 0, st.n, st.rows.data()), "pgpu_part_pull");
    return Ready;
}

Operator::GetNextResultT GpuPartitionOp::getNext(unsigned short tid) {
    ThreadState& st = state[tid];
    const unsigned ts = schema.getTupleSize();
    const uint64_t maxout = buffsize / ts;
    const uint64_t rem = st.n - st.pos;
    const uint64_t k = rem <= maxout ? rem : maxout;
    if (k == 0) {
        st.view.reset();
        return {Finished, &EmptyPage};
    }
    st.view.reset(new Page(st.rows.data() + st.pos * ts, k * ts, nullptr, ts));
    st.pos += k;
    return {rem <= maxout ? Finished : Ready, st.view.get()};
}

Operator::ResultCode GpuPartitionOp::scanStop(unsigned short tid) {
    const bool ok = barrier_.Arrive();  // every thread is done with its partition
    ThreadState& st = state[tid];
    st.view.reset();
    std::vector<char>().swap(st.rows);
    st.n = st.pos = 0;
    if (tid == 0 && part_) check(pgpu_part_reset(part_), "pgpu_part_reset");
    return ok ? Ready : Error;
}

void GpuPartitionOp::threadClose(unsigned short tid) {
    state[tid].view.reset();
    std::vector<char>().swap(state[tid].rows);
}

void GpuPartitionOp::destroy() {
    if (part_) pgpu_part_destroy(part_);
    part_ = nullptr;
    if (ctx) pythia::release_device(ctx);
    ctx = nullptr;
}
