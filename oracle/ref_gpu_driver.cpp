// TEST INFRASTRUCTURE ONLY.  Built by oracle/ref.mk from the reference
// engine's objects (compiled from /root/reference), the GPU operators'
// binding integration/gpujoin.cpp and libpgpu.so; writes only
// oracle/_ref/refgpujoin.
//
// Runs one plan shape inside the reference engine either with its own CPU
// operators or with the binding's GPU operators in their place -- the only
// difference is the class created for the join and aggregate nodes, which is
// what swapping `type = "hashjoin"` for `type = "gpuhashjoin"` does in
// constructsubtree -- and prints one JSON line with an order-independent
// digest of the result, so the two runs can be compared.
//
// usage: refgpujoin c1 <log2 R> <log2 S> <threads> <cpu|gpu> <sum|count|probesum>
//          the queryhashjoin plan (unit_tests/queryhashjoin.cpp) on the C1
//          generators: join -> partial aggregate -> merge, consumer folds
//        refgpujoin rows <R rows> <S rows> <threads> <cpu|gpu> [int|dec|mixed|char|charw]
//          join -> merge with a mixed projection, duplicate keys, misses, and
//          the key types of run_rows; digest of the output multiset
//        refgpujoin groups <D rows> <F rows> <threads> <cpu|gpu> <global 0|1>
//          join -> aggregate_sum(field 0) -> merge with integer-valued dec
//          prices (exact sums); digest of the folded groups
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <string>
#include <vector>

#include "../integration/gpujoin.h"
#include "operators/operators.h"
#include "pgpu_gen.h"
#include "query.h"
#include "ref_memscan.h"
#include "visitors/allvisitors.h"

namespace {

uint64_t mix(uint64_t z) {  // SplitMix64 finalizer
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t row_digest(const char* p, unsigned n) {
    uint64_t h = 0x243F6A8885A308D3ull;
    for (unsigned i = 0; i < n; ++i) h = mix(h ^ (uint64_t)(unsigned char)p[i]);
    return h;
}

Setting& join_node(Config& cfg, int T, const char* const* proj, int nproj) {
    Setting& jn = cfg.getRoot().add("join", Setting::TypeGroup);
    Setting& hn = jn.add("hash", Setting::TypeGroup);
    hn.add("fn", Setting::TypeString) = "modulo";
    hn.add("buckets", Setting::TypeInt) = 4096;
    Setting& tg = jn.add("threadgroups", Setting::TypeList);
    Setting& g0 = tg.add(Setting::TypeArray);
    for (int i = 0; i < T; ++i) g0.add(Setting::TypeInt) = i;
    jn.add("tuplesperbucket", Setting::TypeInt) = 4;
    jn.add("allocpolicy", Setting::TypeString) = "local";
    jn.add("buildjattr", Setting::TypeInt) = 0;
    jn.add("probejattr", Setting::TypeInt) = 0;
    Setting& pj = jn.add("projection", Setting::TypeList);
    for (int i = 0; i < nproj; ++i) pj.add(Setting::TypeString) = proj[i];
    return jn;
}

// Many probe pages per worker: a worker whose first probe read finds the
// shared scan drained makes HashJoinOp::scanStart return Finished
// (operators/join.cpp:451-459), which would end the plan early.
void base_config(Config& cfg, uint64_t probe_bytes, int T) {
    uint64_t bs = probe_bytes / (64 * (uint64_t)T);
    if (bs > (1u << 20)) bs = 1u << 20;
    if (bs < 4096) bs = 4096;
    bs &= ~(uint64_t)15;
    cfg.getRoot().add("path", Setting::TypeString) = "./";
    cfg.getRoot().add("buffsize", Setting::TypeInt) = (int)bs;
}

// Runs q (tree already wired and initialised) and calls fold(tuple) on every
// output tuple.
template <typename F>
int drive(Query& q, F& fold, double* secs) {
    q.threadInit();
    const double t0 = now();
    if (q.scanStart() == Operator::Error) {
        fprintf(stderr, "scanStart failed\n");
        return 1;
    }
    Operator::GetNextResultT r;
    r.first = Operator::Ready;
    while (r.first == Operator::Ready) {
        r = q.getNext();
        if (r.first == Operator::Error) {
            fprintf(stderr, "getNext failed\n");
            return 1;
        }
        Operator::Page::Iterator it = r.second->createIterator();
        void* t;
        while ((t = it.next())) fold(t);
    }
    q.scanStop();
    *secs = now() - t0;
    q.threadClose();
    return 0;
}

struct FoldC1 {
    Schema* s;
    bool is32sum;
    long long value;
    int32_t value32;
    void operator()(void* t) {
        if (is32sum)
            value32 = (int32_t)((uint32_t)value32 + (uint32_t)s->asInt(t, 0));
        else
            value = (long long)((unsigned long long)value + (unsigned long long)s->asLong(t, 0));
    }
};

int run_c1(int argc, char** argv) {
    if (argc < 7) return 2;
    const uint64_t nR = 1ULL << atoi(argv[2]), nS = 1ULL << atoi(argv[3]);
    const int T = atoi(argv[4]);
    const bool gpu = std::string(argv[5]) == "gpu";
    const std::string agg = argv[6];
    const uint64_t seed = PGPU_DEFAULT_SEED;
    std::vector<char> R(nR * 8), S(nS * 8);
    for (uint64_t i = 0; i < nR; ++i) {
        int32_t k = pgpu_c1_rkey(seed, i, nR), p = pgpu_c1_rpay(seed, i);
        memcpy(&R[i * 8], &k, 4);
        memcpy(&R[i * 8 + 4], &p, 4);
    }
    for (uint64_t j = 0; j < nS; ++j) {
        int32_t k = pgpu_c1_sfk(seed, j, nR), p = pgpu_c1_spay(seed, j);
        memcpy(&S[j * 8], &k, 4);
        memcpy(&S[j * 8 + 4], &p, 4);
    }
    Schema rs, ss;
    rs.add(CT_INTEGER);
    rs.add(CT_INTEGER);
    ss.add(CT_INTEGER);
    ss.add(CT_INTEGER);
    Config cfg;
    base_config(cfg, nS * 8, T);
    const char* proj[] = {agg == "probesum" ? "P$1" : "B$1"};
    Setting& jn = join_node(cfg, T, proj, 1);
    Setting& an = cfg.getRoot().add("agg", Setting::TypeGroup);
    an.add("fields", Setting::TypeList);
    if (agg != "count") an.add("sumfield", Setting::TypeInt) = 0;
    Setting& mn = cfg.getRoot().add("merge", Setting::TypeGroup);
    mn.add("threads", Setting::TypeInt) = T;
    Setting& sc = cfg.getRoot().add("scan", Setting::TypeGroup);

    MemScan scanR(&R[0], nR, rs), scanS(&S[0], nS, ss);
    JoinOp* join = gpu ? (JoinOp*)new GpuHashJoinOp() : (JoinOp*)new HashJoinOp();
    SingleInputOp* partial;
    if (agg == "count")
        partial = gpu ? (SingleInputOp*)new GpuAggregateCount() : (SingleInputOp*)new AggregateCount();
    else
        partial = gpu ? (SingleInputOp*)new GpuAggregateSum() : (SingleInputOp*)new AggregateSum();
    MergeOp merge;
    Query q;
    q.tree = &merge;
    merge.nextOp = partial;
    partial->nextOp = join;
    join->buildOp = &scanR;
    join->probeOp = &scanS;
    scanR.init(cfg, sc);
    scanS.init(cfg, sc);
    join->init(cfg, jn);
    partial->init(cfg, an);
    merge.init(cfg, mn);
    FoldC1 f;
    f.s = &q.getOutSchema();
    f.is32sum = agg != "count";
    f.value = 0;
    f.value32 = 0;
    double secs = 0;
    if (drive(q, f, &secs)) return 1;
    q.destroynofree();
    printf("{\"mode\": \"c1\", \"ops\": \"%s\", \"value\": %lld, \"seconds\": %.6f}\n", gpu ? "gpu" : "cpu",
           f.is32sum ? (long long)f.value32 : f.value, secs);
    return 0;
}

struct FoldRows {
    unsigned ts;
    uint64_t n, sum, x;
    void operator()(void* t) {
        const uint64_t h = row_digest((const char*)t, ts);
        ++n;
        sum += h;
        x ^= h;
    }
};

// build (key, long payload, dec value), probe (key, int x): every build key
// used twice, a tenth of the probe keys missing, negative keys.  Key types
// (`keys`): int  int build key against a long probe key (modulo hash)
//           dec  dec against dec, some NaN keys (bytes hash: equal values
//                have equal bytes, no -0 is generated)
//           mixed int build key against a dec probe key (alwayszero hash:
//                one bucket, so the comparator alone decides)
//           char char(7) against char(7), NUL-padded (bytes hash)
//           charw char(5) build against char(9) probe (alwayszero hash)
void put_key(char* dst, ColumnType t, unsigned width, int64_t v, bool nan) {
    if (t == CT_INTEGER) {
        const int32_t x = (int32_t)v;
        memcpy(dst, &x, 4);
    } else if (t == CT_LONG) {
        memcpy(dst, &v, 8);
    } else if (t == CT_DECIMAL) {
        double x = nan ? __builtin_nan("") : (double)v * 0.5;  // halves: exact, never -0
        if (x == 0.0) x = 0.0;
        memcpy(dst, &x, 8);
    } else {
        char buf[32];
        memset(dst, 0, width);
        snprintf(buf, sizeof buf, "k%lld", (long long)v);  // the same text within every width used
        strncpy(dst, buf, width - 1);
    }
}

int run_rows(int argc, char** argv) {
    if (argc < 6) return 2;
    const uint64_t nR = strtoull(argv[2], 0, 10), nS = strtoull(argv[3], 0, 10);
    const int T = atoi(argv[4]);
    const bool gpu = std::string(argv[5]) == "gpu";
    const std::string keys = argc > 6 ? argv[6] : "int";
    ColumnType bt = CT_INTEGER, pt = CT_LONG;
    unsigned bw = 0, pw = 0;
    std::string hashfn = "modulo";
    if (keys == "dec") {
        bt = pt = CT_DECIMAL;
        hashfn = "bytes";
    } else if (keys == "mixed") {
        bt = CT_INTEGER;
        pt = CT_DECIMAL;
        hashfn = "alwayszero";
    } else if (keys == "char") {
        bt = pt = CT_CHAR;
        bw = pw = 8;
        hashfn = "bytes";
    } else if (keys == "charw") {
        bt = pt = CT_CHAR;
        bw = 6;
        pw = 10;
        hashfn = "alwayszero";
    } else if (keys != "int") {
        return 2;
    }
    Schema rs, ss;
    rs.add(bt, bw);
    rs.add(CT_LONG);
    rs.add(CT_DECIMAL);
    ss.add(pt, pw);
    ss.add(CT_INTEGER);
    const unsigned rts = rs.getTupleSize(), sts = ss.getTupleSize();
    const unsigned bkw = rs.getColumnWidth(0), pkw = ss.getColumnWidth(0);
    std::vector<char> R(nR * rts), S(nS * sts);
    // charw: the build's char(5) holds at most 4 characters + NUL ("k" and up
    // to 3 digits), so keys stay below 1000 there
    const int64_t keys_n = keys == "charw" ? 900 : (int64_t)(nR / 2 ? nR / 2 : 1);
    for (uint64_t i = 0; i < nR; ++i) {
        char* r = &R[i * rts];
        const int64_t k = (int64_t)(i % keys_n) * 3 - (keys == "char" || keys == "charw" ? 0 : keys_n);
        put_key(r, bt, bkw, k, keys == "dec" && i % 97 == 5);
        const int64_t p = (int64_t)mix(i);
        const double v = (double)(mix(i + 7) % 100000) / 8.0;
        memcpy(r + bkw, &p, 8);
        memcpy(r + bkw + 8, &v, 8);
    }
    for (uint64_t j = 0; j < nS; ++j) {
        const uint64_t z = mix(j + 1000003);
        const int64_t k = (int64_t)(z % (uint64_t)(keys_n + keys_n / 10 + 1)) * 3 -
                          (keys == "char" || keys == "charw" ? 0 : keys_n);
        char* s = &S[j * sts];
        put_key(s, pt, pkw, k, keys == "dec" && j % 89 == 3);
        const int32_t x = (int32_t)(z >> 40);
        memcpy(s + pkw, &x, 4);
    }
    Config cfg;
    base_config(cfg, nS * sts, T);
    const char* proj[] = {"P$1", "B$2", "B$1", "P$0", "B$0"};
    Setting& jn = join_node(cfg, T, proj, 5);
    jn["hash"]["fn"] = hashfn;
    Setting& mn = cfg.getRoot().add("merge", Setting::TypeGroup);
    mn.add("threads", Setting::TypeInt) = T;
    Setting& sc = cfg.getRoot().add("scan", Setting::TypeGroup);
    MemScan scanR(&R[0], nR, rs), scanS(&S[0], nS, ss);
    JoinOp* join = gpu ? (JoinOp*)new GpuHashJoinOp() : (JoinOp*)new HashJoinOp();
    MergeOp merge;
    Query q;
    q.tree = &merge;
    merge.nextOp = join;
    join->buildOp = &scanR;
    join->probeOp = &scanS;
    scanR.init(cfg, sc);
    scanS.init(cfg, sc);
    join->init(cfg, jn);
    merge.init(cfg, mn);
    FoldRows f;
    f.ts = q.getOutSchema().getTupleSize();
    f.n = f.sum = f.x = 0;
    double secs = 0;
    if (drive(q, f, &secs)) return 1;
    q.destroynofree();
    printf("{\"mode\": \"rows\", \"keys\": \"%s\", \"ops\": \"%s\", \"rows\": %llu, \"row_bytes\": %u, "
           "\"sum\": %llu, \"xor\": %llu, \"seconds\": %.6f}\n",
           keys.c_str(), gpu ? "gpu" : "cpu", (unsigned long long)f.n, f.ts, (unsigned long long)f.sum,
           (unsigned long long)f.x, secs);
    return 0;
}

struct FoldGroups {
    Schema* s;
    std::map<int32_t, double> sums;
    void operator()(void* t) { sums[s->asInt(t, 0)] += s->asDecimal(t, 1); }
};

// D (int orderkey unique, int custkey), F (int orderkey, int partkey, dec
// price with integer values: every sum exact in FP64), grouped by custkey.
int run_groups(int argc, char** argv) {
    if (argc < 7) return 2;
    const uint64_t nD = strtoull(argv[2], 0, 10), nF = strtoull(argv[3], 0, 10);
    const int T = atoi(argv[4]);
    const bool gpu = std::string(argv[5]) == "gpu";
    const bool global = atoi(argv[6]) != 0;
    Schema ds, fs;
    ds.add(CT_INTEGER);
    ds.add(CT_INTEGER);
    fs.add(CT_INTEGER);
    fs.add(CT_INTEGER);
    fs.add(CT_DECIMAL);
    std::vector<char> D(nD * 8), F(nF * fs.getTupleSize());
    const uint64_t ncust = nD / 10 + 1;
    for (uint64_t i = 0; i < nD; ++i) {
        int32_t k = (int32_t)i + 1, c = (int32_t)(mix(i) % ncust) - (int32_t)(ncust / 2);
        memcpy(&D[i * 8], &k, 4);
        memcpy(&D[i * 8 + 4], &c, 4);
    }
    for (uint64_t j = 0; j < nF; ++j) {
        const uint64_t z = mix(j + 77);
        int32_t k = (int32_t)(z % (nD + nD / 20 + 1)) + 1, p = (int32_t)(z >> 40);
        double v = (double)(1 + (z >> 20) % 1000);
        char* f = &F[j * fs.getTupleSize()];
        memcpy(f, &k, 4);
        memcpy(f + 4, &p, 4);
        memcpy(f + 8, &v, 8);
    }
    Config cfg;
    base_config(cfg, nF * fs.getTupleSize(), T);
    const char* proj[] = {"B$1", "P$2"};
    Setting& jn = join_node(cfg, T, proj, 2);
    Setting& an = cfg.getRoot().add("agg", Setting::TypeGroup);
    an.add("field", Setting::TypeInt) = 0;
    an.add("sumfield", Setting::TypeInt) = 1;
    Setting& ah = an.add("hash", Setting::TypeGroup);
    ah.add("fn", Setting::TypeString) = "modulo";
    ah.add("buckets", Setting::TypeInt) = 4096;
    ah.add("field", Setting::TypeInt) = 0;
    if (global) {
        an.add("global", Setting::TypeBoolean) = true;
        an.add("threads", Setting::TypeInt) = T;
    }
    Setting& mn = cfg.getRoot().add("merge", Setting::TypeGroup);
    mn.add("threads", Setting::TypeInt) = T;
    Setting& sc = cfg.getRoot().add("scan", Setting::TypeGroup);
    MemScan scanD(&D[0], nD, ds), scanF(&F[0], nF, fs);
    JoinOp* join = gpu ? (JoinOp*)new GpuHashJoinOp() : (JoinOp*)new HashJoinOp();
    SingleInputOp* agg = gpu ? (SingleInputOp*)new GpuAggregateSum() : (SingleInputOp*)new AggregateSum();
    MergeOp merge;
    Query q;
    q.tree = &merge;
    merge.nextOp = agg;
    agg->nextOp = join;
    join->buildOp = &scanD;
    join->probeOp = &scanF;
    scanD.init(cfg, sc);
    scanF.init(cfg, sc);
    join->init(cfg, jn);
    agg->init(cfg, an);
    merge.init(cfg, mn);
    FoldGroups f;
    f.s = &q.getOutSchema();
    double secs = 0;
    if (drive(q, f, &secs)) return 1;
    q.destroynofree();
    uint64_t sum = 0, x = 0;
    for (std::map<int32_t, double>::const_iterator it = f.sums.begin(); it != f.sums.end(); ++it) {
        char b[12];
        memcpy(b, &it->first, 4);
        memcpy(b + 4, &it->second, 8);
        const uint64_t h = row_digest(b, 12);
        sum += h;
        x ^= h;
    }
    printf("{\"mode\": \"groups\", \"ops\": \"%s\", \"groups\": %llu, \"sum\": %llu, \"xor\": %llu, "
           "\"seconds\": %.6f}\n",
           gpu ? "gpu" : "cpu", (unsigned long long)f.sums.size(), (unsigned long long)sum, (unsigned long long)x,
           secs);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    int rc = 2;
    try {
        if (argc >= 2 && std::string(argv[1]) == "c1") rc = run_c1(argc, argv);
        if (argc >= 2 && std::string(argv[1]) == "rows") rc = run_rows(argc, argv);
        if (argc >= 2 && std::string(argv[1]) == "groups") rc = run_groups(argc, argv);
    } catch (...) {
        fprintf(stderr, "refgpujoin: the plan threw\n");
        return 1;
    }
    if (rc == 2)
        fprintf(stderr,
                "usage: refgpujoin c1 <log2 R> <log2 S> <threads> <cpu|gpu> <sum|count|probesum>\n"
                "       refgpujoin rows <R rows> <S rows> <threads> <cpu|gpu> [int|dec|mixed|char|charw]\n"
                "       refgpujoin groups <D rows> <F rows> <threads> <cpu|gpu> <global 0|1>\n");
    return rc;
}
