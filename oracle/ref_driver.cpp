// TEST INFRASTRUCTURE ONLY.  Built by oracle/ref.mk against the reference
// engine's headers and objects (compiled from /root/reference); writes only
// oracle/_ref/refjoin.
//
// Runs the reference's own operators on generated inputs with the plan shape
// of unit_tests/queryhashjoin.cpp: an in-memory page source for each side ->
// HashJoinOp (one thread group of T threads, modulo hash) -> partial
// AggregateSum (or AggregateCount) per worker -> MergeOp(T); the consumer
// folds the per-worker partials in the accumulator's type, as the final
// aggregate of that plan does.  (A worker whose first probe read finds the
// shared scan already drained returns Finished from HashJoinOp::scanStart,
// operators/join.cpp:451-459; the merge then reports Finished from
// scanStart, so a final aggregate above it would skip its input.)  Timed from scanStart to scanStop with CLOCK_MONOTONIC as
// drivers/executequery.cpp:38-59 does (with rdtsc there).
//
// usage: refjoin <c1|c2> <log2 R> <log2 S> <threads> <zipf s> <buckets> <tuplesperbucket> <sum|count|probesum> [runs]
//        refjoin c4 <first join> <joins> <threads> <tuplesperbucket> <sum|probesum> [runs]
//          (the batched-join family: one plan per join, buckets = 2^ceil(log2 n))
//        refjoin hash <int|long|chars> <hash node> <value>...
//          (hash values through the reference's TupleHasher::create)
//        refjoin htwrite|htread <buckets> <tuples per bucket> <values> <path>
//          (HashTable::serialize / deserialize of int tuples)
//        refjoin c3 <dimension rows> <fact rows> <threads> [runs]
//          (join + group-by: hashjoin(D, F) -> aggregate_sum(field 0) per worker ->
//           merge; the consumer folds the per-worker groups)
// prints one JSON line {"value": ..., "seconds": ..., ...}.  With runs > 1 the
// plan is built and run that many times over the same inputs (a fresh plan
// each time, as one executequery process per run would); "seconds" is the
// median and "runs" lists every run.
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <string>
#include <vector>

#include "operators/operators.h"
#include "query.h"
#include "visitors/allvisitors.h"
#include "pgpu_gen.h"
#include "ref_memscan.h"

using namespace libconfig;

// Runs the plan once over packed rows; *value gets the folded aggregate (an
// int32 sum wraps in 32 bits, as AggregateSum's int accumulator does).
static int run_plan(char* R, uint64_t nR, char* S, uint64_t nS, bool is32, bool probe_pay, int T, int buckets,
                    int tpb, const std::string& agg, long long* value_out, double* secs) {
    Schema rs, ss;
    rs.add(is32 ? CT_INTEGER : CT_LONG);
    rs.add(is32 ? CT_INTEGER : CT_LONG);
    ss.add(is32 ? CT_INTEGER : CT_LONG);
    if (probe_pay) ss.add(is32 ? CT_INTEGER : CT_LONG);
    const uint64_t prow = ss.getTupleSize();

    Config cfg;
    cfg.getRoot().add("path", Setting::TypeString) = "./";
    // Many probe pages per worker: a worker whose first probe read finds the
    // shared scan drained makes HashJoinOp::scanStart return Finished
    // (operators/join.cpp:451-459), which ends the plan early.
    uint64_t bs = nS * prow / (64 * (uint64_t)T);
    if (bs > (1u << 20)) bs = 1u << 20;
    if (bs < 4096) bs = 4096;
    bs &= ~(uint64_t)15;
    cfg.getRoot().add("buffsize", Setting::TypeInt) = (int)bs;
    Setting& jn = cfg.getRoot().add("join", Setting::TypeGroup);
    Setting& hn = jn.add("hash", Setting::TypeGroup);
    hn.add("fn", Setting::TypeString) = "modulo";
    hn.add("buckets", Setting::TypeInt) = buckets;
    Setting& tg = jn.add("threadgroups", Setting::TypeList);
    Setting& g0 = tg.add(Setting::TypeArray);
    for (int i = 0; i < T; ++i) g0.add(Setting::TypeInt) = i;
    jn.add("tuplesperbucket", Setting::TypeInt) = tpb;
    jn.add("allocpolicy", Setting::TypeString) = "local";
    jn.add("buildjattr", Setting::TypeInt) = 0;
    jn.add("probejattr", Setting::TypeInt) = 0;
    Setting& pj = jn.add("projection", Setting::TypeList);
    pj.add(Setting::TypeString) = agg == "probesum" ? "P$1" : "B$1";
    Setting& an = cfg.getRoot().add("agg", Setting::TypeGroup);
    an.add("fields", Setting::TypeList);
    an.add("sumfield", Setting::TypeInt) = 0;
    Setting& mn = cfg.getRoot().add("merge", Setting::TypeGroup);
    mn.add("threads", Setting::TypeInt) = T;
    Setting& sc = cfg.getRoot().add("scan", Setting::TypeGroup);

    MemScan scanR(R, nR, rs), scanS(S, nS, ss);
    HashJoinOp join;
    GenericAggregate* partial;
    if (agg == "count")
        partial = new AggregateCount();
    else
        partial = new AggregateSum();
    MergeOp merge;
    Query q;
    q.tree = &merge;
    merge.nextOp = partial;
    partial->nextOp = &join;
    join.buildOp = &scanR;
    join.probeOp = &scanS;
    scanR.init(cfg, sc);
    scanS.init(cfg, sc);
    join.init(cfg, jn);
    partial->init(cfg, an);
    merge.init(cfg, mn);

    q.threadInit();
    double t0 = now();
    if (q.scanStart() == Operator::Error) {
        fprintf(stderr, "scanStart failed\n");
        return 1;
    }
    Operator::GetNextResultT r;
    r.first = Operator::Ready;
    long long value = 0;
    int32_t value32 = 0;
    while (r.first == Operator::Ready) {
        r = q.getNext();
        if (r.first == Operator::Error) {
            fprintf(stderr, "getNext failed\n");
            return 1;
        }
        void* t;
        Operator::Page::Iterator it = r.second->createIterator();
        while ((t = it.next())) {
            if (is32 && agg != "count")
                value32 = (int32_t)((uint32_t)value32 + (uint32_t)q.getOutSchema().asInt(t, 0));
            else
                value = (long long)((unsigned long long)value + (unsigned long long)q.getOutSchema().asLong(t, 0));
        }
    }
    if (is32 && agg != "count") value = value32;
    q.scanStop();
    double t1 = now();
    q.threadClose();
    q.destroynofree();
    delete partial;
    *value_out = value;
    *secs = t1 - t0;
    return 0;
}

static double median(std::vector<double> v) {
    std::sort(v.begin(), v.end());
    const size_t n = v.size();
    return n % 2 ? v[n / 2] : 0.5 * (v[n / 2 - 1] + v[n / 2]);
}

static std::string runs_json(const std::vector<double>& v) {
    std::string out;
    char buf[32];
    for (size_t i = 0; i < v.size(); ++i) {
        snprintf(buf, sizeof buf, "%s%.6f", i ? ", " : "", v[i]);
        out += buf;
    }
    return out;
}

static int parse_runs(int argc, char** argv, int at) {
    const int r = argc > at ? atoi(argv[at]) : 1;
    return r < 1 ? 1 : r;
}

static void usage() {
    fprintf(stderr,
            "usage: refjoin <c1|c2> <log2R> <log2S> <threads> <zipf> <buckets> <tpb> <sum|count|probesum> [runs]\n"
            "       refjoin c4 <first> <joins> <threads> <tpb> <sum|probesum> [runs]\n"
            "       refjoin c3 <dimension rows> <fact rows> <threads> [runs]\n"
            "       refjoin hash <int|long|chars> <hash node> <value>...\n"
            "       refjoin htwrite|htread <buckets> <tuples per bucket> <values> <path>\n");
}

// The batched family: joins first..first+count-1, each its own plan; inputs
// are generated outside the clock; prints the summed plan time and each
// join's aggregate.
static int run_c4(int argc, char** argv) {
    if (argc < 7) {
        usage();
        return 2;
    }
    const uint64_t first = strtoull(argv[2], NULL, 10), count = strtoull(argv[3], NULL, 10);
    const int T = atoi(argv[4]), tpb = atoi(argv[5]);
    const std::string agg = argv[6];
    const int runs = parse_runs(argc, argv, 7);
    const uint64_t seed = PGPU_DEFAULT_SEED;
    std::vector<double> total(runs, 0.0);
    uint64_t tuples = 0;
    std::string values;
    for (uint64_t j = first; j < first + count; ++j) {
        const uint64_t n = pgpu_loguniform_size(seed, 4, j, 10.0, 22.0), m = 8 * n;
        std::vector<char> R(n * 8), S(m * 8);
        for (uint64_t i = 0; i < n; ++i) {
            int32_t k = pgpu_c4_rkey(seed, j, i, n), p = pgpu_c4_rpay(seed, j, i);
            memcpy(&R[i * 8], &k, 4);
            memcpy(&R[i * 8 + 4], &p, 4);
        }
        for (uint64_t i = 0; i < m; ++i) {
            int32_t k = pgpu_c4_sfk(seed, j, i, n), p = pgpu_c4_spay(seed, j, i);
            memcpy(&S[i * 8], &k, 4);
            memcpy(&S[i * 8 + 4], &p, 4);
        }
        int buckets = 1;
        while ((uint64_t)buckets < n) buckets <<= 1;
        long long v = 0;
        for (int r = 0; r < runs; ++r) {
            double secs = 0;
            if (run_plan(&R[0], n, &S[0], m, true, true, T, buckets, tpb, agg, &v, &secs)) return 1;
            total[r] += secs;
        }
        tuples += n + m;
        char buf[32];
        snprintf(buf, sizeof buf, "%s%lld", values.empty() ? "" : ", ", v);
        values += buf;
    }
    printf("{\"workload\": \"c4\", \"first\": %llu, \"joins\": %llu, \"threads\": %d, \"agg\": \"%s\", "
           "\"tuples\": %llu, \"seconds\": %.6f, \"runs\": [%s], \"values\": [%s]}\n",
           (unsigned long long)first, (unsigned long long)count, T, agg.c_str(), (unsigned long long)tuples,
           median(total), runs_json(total).c_str(), values.c_str());
    return 0;
}

static int run_c3_plan(char* D, uint64_t nD, char* F, uint64_t nF, int T, uint64_t* groups_out, double* total_out,
                       double* secs);

// C3: SUM(F.price) GROUP BY D.custkey over F JOIN D ON orderkey, with the
// packed rows the reference would scan (D: int orderkey, int custkey;
// F: int orderkey, int partkey, int quantity, dec price = 20 bytes).
static int run_c3(int argc, char** argv) {
    if (argc < 5) {
        usage();
        return 2;
    }
    const uint64_t nD = strtoull(argv[2], NULL, 10), nF = strtoull(argv[3], NULL, 10);
    const int T = atoi(argv[4]);
    const int runs = parse_runs(argc, argv, 5);
    const uint64_t seed = PGPU_DEFAULT_SEED, norders = nD;
    std::vector<char> D(nD * 8), F(nF * 20);
    for (uint64_t i = 0; i < nD; ++i) {
        int32_t k = pgpu_c3_d_orderkey(seed, i, nD), c = pgpu_c3_d_custkey(seed, i, PGPU_C3_CUSTS);
        memcpy(&D[i * 8], &k, 4);
        memcpy(&D[i * 8 + 4], &c, 4);
    }
    for (uint64_t j = 0; j < nF; ++j) {
        int32_t k = pgpu_c3_f_orderkey(seed, j, norders), pk = pgpu_c3_f_partkey(seed, j),
                q = pgpu_c3_f_quantity(seed, j);
        double pr = pgpu_c3_f_price(seed, j);
        memcpy(&F[j * 20], &k, 4);
        memcpy(&F[j * 20 + 4], &pk, 4);
        memcpy(&F[j * 20 + 8], &q, 4);
        memcpy(&F[j * 20 + 12], &pr, 8);
    }
    std::vector<double> secs(runs);
    uint64_t groups = 0;
    double total = 0;
    for (int r = 0; r < runs; ++r)
        if (run_c3_plan(&D[0], nD, &F[0], nF, T, &groups, &total, &secs[r])) return 1;
    printf("{\"workload\": \"c3\", \"build\": %llu, \"probe\": %llu, \"threads\": %d, \"groups\": %llu, "
           "\"total\": %.17g, \"seconds\": %.6f, \"runs\": [%s]}\n",
           (unsigned long long)nD, (unsigned long long)nF, T, (unsigned long long)groups, total, median(secs),
           runs_json(secs).c_str());
    return 0;
}

// One C3 plan over the packed rows: D (int orderkey, int custkey) and F
// (int orderkey, int partkey, int quantity, dec price).
static int run_c3_plan(char* D, uint64_t nD, char* F, uint64_t nF, int T, uint64_t* groups_out, double* total_out,
                       double* secs) {
    Schema ds, fs;
    ds.add(CT_INTEGER);
    ds.add(CT_INTEGER);
    fs.add(CT_INTEGER);
    fs.add(CT_INTEGER);
    fs.add(CT_INTEGER);
    fs.add(CT_DECIMAL);
    Config cfg;
    cfg.getRoot().add("path", Setting::TypeString) = "./";
    uint64_t bs = nF * 20 / (64 * (uint64_t)T);
    if (bs > (1u << 20)) bs = 1u << 20;
    if (bs < 4096) bs = 4096;
    bs &= ~(uint64_t)15;
    cfg.getRoot().add("buffsize", Setting::TypeInt) = (int)bs;
    Setting& jn = cfg.getRoot().add("join", Setting::TypeGroup);
    Setting& hn = jn.add("hash", Setting::TypeGroup);
    hn.add("fn", Setting::TypeString) = "modulo";
    hn.add("buckets", Setting::TypeInt) = 1 << 24;
    Setting& tg = jn.add("threadgroups", Setting::TypeList);
    Setting& g0 = tg.add(Setting::TypeArray);
    for (int i = 0; i < T; ++i) g0.add(Setting::TypeInt) = i;
    jn.add("tuplesperbucket", Setting::TypeInt) = 4;
    jn.add("allocpolicy", Setting::TypeString) = "local";
    jn.add("buildjattr", Setting::TypeInt) = 0;
    jn.add("probejattr", Setting::TypeInt) = 0;
    Setting& pj = jn.add("projection", Setting::TypeList);
    pj.add(Setting::TypeString) = "B$1";
    pj.add(Setting::TypeString) = "P$3";
    Setting& an = cfg.getRoot().add("agg", Setting::TypeGroup);
    an.add("field", Setting::TypeInt) = 0;
    an.add("sumfield", Setting::TypeInt) = 1;
    Setting& ah = an.add("hash", Setting::TypeGroup);
    ah.add("fn", Setting::TypeString) = "modulo";
    ah.add("buckets", Setting::TypeInt) = 1 << 21;
    ah.add("field", Setting::TypeInt) = 0;
    Setting& mn = cfg.getRoot().add("merge", Setting::TypeGroup);
    mn.add("threads", Setting::TypeInt) = T;
    Setting& sc = cfg.getRoot().add("scan", Setting::TypeGroup);

    MemScan scanD(D, nD, ds), scanF(F, nF, fs);
    HashJoinOp join;
    AggregateSum partial;
    MergeOp merge;
    Query q;
    q.tree = &merge;
    merge.nextOp = &partial;
    partial.nextOp = &join;
    join.buildOp = &scanD;
    join.probeOp = &scanF;
    scanD.init(cfg, sc);
    scanF.init(cfg, sc);
    join.init(cfg, jn);
    partial.init(cfg, an);
    merge.init(cfg, mn);

    q.threadInit();
    std::vector<double> sums(PGPU_C3_CUSTS + 1, 0.0);
    std::vector<char> seen(PGPU_C3_CUSTS + 1, 0);
    double t0 = now();
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
        void* t;
        Operator::Page::Iterator it = r.second->createIterator();
        while ((t = it.next())) {
            const int c = q.getOutSchema().asInt(t, 0);
            if (c < 0 || c > (int)PGPU_C3_CUSTS) continue;
            sums[c] += q.getOutSchema().asDecimal(t, 1);  // the final aggregate over the partials
            seen[c] = 1;
        }
    }
    q.scanStop();
    double t1 = now();
    q.threadClose();
    q.destroynofree();
    uint64_t groups = 0;
    double total = 0;
    for (uint64_t c = 0; c <= PGPU_C3_CUSTS; ++c) {
        groups += seen[c];
        total += sums[c];
    }
    *groups_out = groups;
    *total_out = total;
    *secs = t1 - t0;
    return 0;
}

// refjoin hash <int|long|chars> <hash node> <value>...: builds the hasher
// with the reference's own TupleHasher::create (hash.cpp:99-234) over a
// one-column int or long schema, or two char(1) columns (`chars`, hashed as
// fieldrange [0, 1]), and prints its bucket count and the hash of a tuple
// holding each value (for `chars`, the value's four low bytes).
static int run_hash(int argc, char** argv) {
    if (argc < 4) {
        usage();
        return 2;
    }
    const std::string ty = argv[2];
    Schema s;
    if (ty == "int")
        s.add(CT_INTEGER);
    else if (ty == "long")
        s.add(CT_LONG);
    else {
        s.add(CT_CHAR, 2);  // char(1): one byte and its terminator
        s.add(CT_CHAR, 2);
    }
    Config cfg;  // libconfig 1.2 reads streams only
    std::string text = std::string("h = ") + argv[3] + ";\n";
    FILE* f = fmemopen(&text[0], text.size(), "r");
    if (!f) return 1;
    cfg.read(f);
    fclose(f);
    TupleHasher th = TupleHasher::create(s, cfg.lookup("h"));
    printf("{\"buckets\": %u, \"hashes\": [", th.buckets());
    for (int i = 4; i < argc; ++i) {
        const long long v = strtoll(argv[i], NULL, 10);
        char tuple[16];
        memset(tuple, 0, sizeof tuple);
        if (ty == "int") {
            const int32_t x = (int32_t)v;
            memcpy(tuple, &x, 4);
        } else if (ty == "long") {
            memcpy(tuple, &v, 8);
        } else {
            const uint32_t x = (uint32_t)v;
            memcpy(tuple, &x, 4);
        }
        printf("%s%u", i > 4 ? ", " : "", th.hash(tuple));
    }
    printf("]}\n");
    th.destroy();
    return 0;
}

// refjoin htwrite <buckets> <tuples per bucket> <values> <path>
//   the reference's HashTable (util/hashtable.cpp:57-97) of int tuples, value
//   v in bucket v % buckets (testhashtable.cpp:87-135 with more than one
//   tuple per bucket), serialized to <path> (hashtable.cpp:99-121)
// refjoin htread <buckets> <tuples per bucket> <values> <path>
//   a table of that geometry deserialized from <path> (hashtable.cpp:178-201)
//   and iterated: prints how many values appear exactly once, in their bucket
static int run_ht(int argc, char** argv, bool write) {
    if (argc < 6) {
        usage();
        return 2;
    }
    const unsigned nb = (unsigned)atoi(argv[2]), tpb = (unsigned)atoi(argv[3]);
    const int nv = atoi(argv[4]);
    HashTable ht;
    ht.init(nb, tpb * sizeof(int), sizeof(int), std::vector<char>(), 0);
    ht.bucketclear(0, 1);
    if (write) {
        for (int v = 0; v < nv; ++v) *(int*)ht.allocate(v % nb, 0) = v;
        if (ht.statSpills() != 0) {
            fprintf(stderr, "spills\n");
            return 1;
        }
        ht.serialize(argv[5], 0);
        printf("{\"written\": %d}\n", nv);
    } else {
        ht.deserialize(argv[5], 0);
        std::vector<int> seen(nv, 0);
        int bad = 0;
        HashTable::Iterator it = ht.createIterator();
        for (unsigned b = 0; b < nb; ++b) {
            ht.placeIterator(it, b);
            void* tup;
            while ((tup = it.next())) {
                const int v = *(int*)tup;
                if (v < 0 || v >= nv || (unsigned)v % nb != b)
                    ++bad;
                else
                    ++seen[v];
            }
        }
        int once = 0;
        for (int v = 0; v < nv; ++v) once += seen[v] == 1;
        printf("{\"once\": %d, \"bad\": %d}\n", once, bad);
    }
    ht.destroy();
    return 0;
}

int main(int argc, char** argv) {
    if (argc >= 2 && std::string(argv[1]) == "hash") return run_hash(argc, argv);
    if (argc >= 2 && std::string(argv[1]) == "htwrite") return run_ht(argc, argv, true);
    if (argc >= 2 && std::string(argv[1]) == "htread") return run_ht(argc, argv, false);
    if (argc >= 2 && std::string(argv[1]) == "c4") return run_c4(argc, argv);
    if (argc >= 2 && std::string(argv[1]) == "c3") return run_c3(argc, argv);
    if (argc < 9) {
        usage();
        return 2;
    }
    const std::string wl = argv[1];
    const uint64_t nR = 1ULL << atoi(argv[2]), nS = 1ULL << atoi(argv[3]);
    const int T = atoi(argv[4]);
    const double zipf = atof(argv[5]);
    const int buckets = atoi(argv[6]), tpb = atoi(argv[7]);
    const std::string agg = argv[8];
    const int runs = parse_runs(argc, argv, 9);
    const uint64_t seed = PGPU_DEFAULT_SEED;
    const bool is32 = wl == "c1";
    const unsigned kb = is32 ? 4 : 8;

    // build rows (key, payload); probe rows (fk[, payload])
    const unsigned prow = 8;
    std::vector<char> R(nR * 2 * kb), S(nS * prow);
    std::vector<double> cdf;
    if (zipf > 0) {
        cdf.resize(nR);
        pgpu_zipf_cdf_fill(&cdf[0], nR, zipf);
    }
    for (uint64_t i = 0; i < nR; ++i) {
        if (is32) {
            int32_t k = pgpu_c1_rkey(seed, i, nR), p = pgpu_c1_rpay(seed, i);
            memcpy(&R[i * 8], &k, 4);
            memcpy(&R[i * 8 + 4], &p, 4);
        } else {
            int64_t k = pgpu_c2_rkey(seed, 0, i), p = pgpu_c2_rpay(seed, 0, i);
            memcpy(&R[i * 16], &k, 8);
            memcpy(&R[i * 16 + 8], &p, 8);
        }
    }
    for (uint64_t j = 0; j < nS; ++j) {
        if (is32) {
            int32_t k = pgpu_c1_sfk(seed, j, nR), p = pgpu_c1_spay(seed, j);
            memcpy(&S[j * 8], &k, 4);
            memcpy(&S[j * 8 + 4], &p, 4);
        } else {
            uint64_t row = zipf > 0 ? pgpu_c2_srow_zipf(seed, 0, j, nR, &cdf[0]) : pgpu_c2_srow_uniform(seed, 0, j, nR);
            int64_t k = pgpu_c2_rkey(seed, 0, row);
            memcpy(&S[j * 8], &k, 8);
        }
    }
    long long value = 0;
    std::vector<double> secs(runs);
    for (int r = 0; r < runs; ++r)
        if (run_plan(&R[0], nR, &S[0], nS, is32, is32, T, buckets, tpb, agg, &value, &secs[r])) return 1;
    printf("{\"workload\": \"%s\", \"build\": %llu, \"probe\": %llu, \"threads\": %d, \"agg\": \"%s\", "
           "\"value\": %lld, \"seconds\": %.6f, \"runs\": [%s]}\n",
           wl.c_str(), (unsigned long long)nR, (unsigned long long)nS, T, agg.c_str(), value, median(secs),
           runs_json(secs).c_str());
    return 0;
}
