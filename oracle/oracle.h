/*
 * oracle.h - TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference engine's hash-join / hash-aggregation
 * path (sblanas/pythia), used as the checker of the GPU library and as the
 * timed CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product path never calls it.
 *
 * Each function cites the reference file:line it restates.  Parity of the
 * restatement itself is pinned by tests/test_oracle_*.py against the
 * reference's own unit-test known answers, closed forms, brute force on tiny
 * inputs, and (where it builds) the compiled reference under oracle/_ref.
 */
#ifndef PYTHIA_ORACLE_H
#define PYTHIA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- column types (schema.h:52-63) ------------------------ */
enum orc_ctype { ORC_INT = 0, ORC_LONG = 1, ORC_DEC = 2, ORC_CHAR = 3, ORC_DATE = 4 };

typedef struct orc_schema {
    int ncols;
    int type[32];
    int size[32];
    int offset[32]; /* packed, no padding (schema.cpp:82-117) */
    int tuplesize;
} orc_schema;

/* Appends a column; size is only read for ORC_CHAR. */
void orc_schema_add(orc_schema* s, int type, int size);

/* ---------------- hash functions (hash.h / hash.cpp) ------------------- */
enum orc_hashfn {
    ORC_H_MODULO = 0,
    ORC_H_PARAMMODULO,
    ORC_H_KNUTH,
    ORC_H_RANGE,
    ORC_H_EXACTRANGE,
    ORC_H_BYTES,
    ORC_H_ALWAYSZERO,
    ORC_H_WILLIS,
    ORC_H_TPCHORDERKEY,
    ORC_H_TPCHQ1MAGIC
};

typedef struct orc_hasher {
    int fn;
    int64_t k;        /* log2 buckets, or the mask for modulo-family hashers */
    int64_t min, max; /* range / exactrange / parammodulo offset in min      */
    int64_t bucketrange;
    int skipbits;
    unsigned buckets; /* domain size */
    int offset, size; /* hashed field span inside the tuple */
} orc_hasher;

unsigned orc_getlogarithm(unsigned k);
/* Builds a hasher as TupleHasher::create does (hash.cpp:99-234). */
int orc_hasher_init(orc_hasher* h, int fn, unsigned buckets, int64_t rmin, int64_t rmax, int skipbits,
                    const orc_schema* s, int fieldmin, int fieldmax);
unsigned orc_hash_value(const orc_hasher* h, int64_t value);
unsigned orc_hash_tuple(const orc_hasher* h, const void* tuple);
int64_t orc_exactrange_minimum(const orc_hasher* h, unsigned bucket);
/* ParameterizedModuloValueHasher::generate (hash.cpp:237-267): per pass the
 * bucket count and skipbits. Returns the number of passes written. */
int orc_generate(unsigned buckets, int skipbits, unsigned passes, unsigned* out_buckets, int* out_skip);

/* ---------------- chained-bucket hash table (util/hashtable.*) --------- */
typedef struct orc_htable orc_htable;
orc_htable* orc_ht_create(unsigned nbuckets, unsigned bucksize, unsigned tuplesize);
void orc_ht_destroy(orc_htable* t);
void* orc_ht_allocate(orc_htable* t, unsigned bucket); /* util/hashtable.cpp:230-273 */
/* Visits bucket tuples in chain order; returns the count visited. */
unsigned orc_ht_bucket_tuples(orc_htable* t, unsigned bucket, void** out, unsigned cap);
uint64_t orc_ht_spills(const orc_htable* t);
/* HashTable::serialize / deserialize (util/hashtable.cpp:99-201): the bucket
 * array image of a spill-free table in the reference's byte layout. */
int orc_ht_serialize(orc_htable* t, const char* path);
int orc_ht_deserialize(orc_htable* t, const char* path);

/* ---------------- hash join (operators/join.cpp) ----------------------- */
typedef struct orc_join_desc {
    orc_schema build, probe;
    int buildjattr, probejattr;
    int nproj;
    int proj_side[32]; /* 0 = build ("B$i"), 1 = probe ("P$i") */
    int proj_col[32];
    int hashfn;        /* join hash (modulo unless stated)             */
    unsigned buckets;
    int64_t rmin, rmax;
    int skipbits;
    unsigned tuplesperbucket;
    int threads;       /* worker threads sharing one table (1 group)   */
} orc_join_desc;

typedef struct orc_join_stats {
    uint64_t matches;
    uint64_t spills;
    double build_s, probe_s;
} orc_join_stats;

/* Output schema = projected columns in order (operators/join.cpp:242-259). */
void orc_join_outschema(const orc_join_desc* d, orc_schema* out);
/* Materialising join: writes every output row (probe order x chain order
 * per thread) into out (capacity cap rows). Returns rows written, or -1 when
 * cap is too small. */
int64_t orc_hashjoin(const orc_join_desc* d, const void* build, uint64_t nbuild, const void* probe,
                     uint64_t nprobe, void* out, uint64_t cap, orc_join_stats* st);
/* Join feeding count + sum of one output column (count = AggregateCount,
 * sum accumulated in the column's type as AggregateSum does: int wraps mod
 * 2^32, long mod 2^64, dec in double, in arrival order per thread). */
typedef struct orc_count_sum {
    uint64_t count;
    int64_t isum;   /* int/long sums (int32 sums are sign-extended wraps)  */
    double dsum;
} orc_count_sum;
int orc_hashjoin_count_sum(const orc_join_desc* d, const void* build, uint64_t nbuild, const void* probe,
                           uint64_t nprobe, int sumcol, orc_count_sum* out, orc_join_stats* st);

/* ---------------- aggregation (genericaggregate.cpp, aggregatesum.cpp) - */
typedef struct orc_agg_desc {
    orc_schema in;
    int nfields;      /* 0 = no group-by (alwayszero hash)                  */
    int fields[8];
    int sumfield;     /* AggregateSum column; -1 = AggregateCount           */
    int hashfn;
    unsigned buckets;
    int64_t rmin, rmax;
    int global;       /* Global mode: one shared table                      */
    int threads;
} orc_agg_desc;
void orc_agg_outschema(const orc_agg_desc* d, orc_schema* out);
/* Aggregates nrows input rows split into `threads` contiguous slices, one
 * per worker; emits the groups as GenericAggregate::getNext does per thread
 * (ThreadLocal: one row per (thread, group); Global: each group once).
 * Returns rows written or -1 when cap is too small. */
int64_t orc_aggregate(const orc_agg_desc* d, const void* rows, uint64_t nrows, void* out, uint64_t cap);

/* ---------------- partition (operators/partition.cpp) ------------------ */
/* exactrange partitioning of nrows rows on column attr over [rmin, rmax]
 * into `buckets` partitions; writes partition-ordered rows to out and the
 * partition sizes to counts[buckets]. */
int orc_partition(const orc_schema* s, int attr, int64_t rmin, int64_t rmax, unsigned buckets, const void* rows,
                  uint64_t nrows, void* out, uint64_t* counts);

/* ---------------- closed forms over the generators (pgpu_gen.h) -------- */
typedef struct orc_cs64 {
    uint64_t count;
    uint64_t sum_build;  /* modulo 2^64 */
    uint64_t sum_probe;  /* modulo 2^64 */
} orc_cs64;
/* C1: count and sums of R.payload / S.payload over R join S. */
void orc_c1_closed(uint64_t seed, uint64_t nR, uint64_t nS, int threads, orc_cs64* out);
/* C2: uniform (zipf_s == 0) or Zipf(s) FK over nR keys; sum of R.payload. */
void orc_c2_closed(uint64_t seed, uint64_t sub, uint64_t nR, uint64_t nS, double zipf_s, int threads,
                   orc_cs64* out);
/* C3: exact integer cents of sum(price) per custkey (1..ncust), written to
 * cents[ncust + 1]; returns the number of joined fact rows. */
uint64_t orc_c3_closed(uint64_t seed, uint64_t nD, uint64_t nF, uint64_t norders, uint64_t ncust, int threads,
                       int64_t* cents);
/* C4: count, sum of R.payload and sum of S.payload over the matches of join
 * `join` (build n rows, probe nprobe rows). */
void orc_c4_closed(uint64_t seed, uint64_t join, uint64_t n, uint64_t nprobe, orc_cs64* out);
/* The same for joins first..first+count-1 with build sizes n[i] and probe
 * sizes m[i], spread over `threads` threads. */
void orc_c4_closed_many(uint64_t seed, uint64_t first, uint64_t count, const uint64_t* n, const uint64_t* m,
                        int threads, orc_cs64* out);

/* C5: probe rows first..first+m-1 of the uniform FK column of table `table`
 * (n rows): count and sum of R.payload over the matches. */
void orc_c5_closed(uint64_t seed, uint64_t table, uint64_t n, uint64_t first, uint64_t m, int threads, orc_cs64* out);

/* ---------------- generator wrappers (pgpu_gen.h) ---------------------- */
void orc_gen_c4(uint64_t seed, uint64_t join, uint64_t n, uint64_t m, int32_t* rkey, int32_t* rpay, int32_t* sfk,
                int32_t* spay);
void orc_gen_c1(uint64_t seed, uint64_t nR, uint64_t nS, int32_t* rkey, int32_t* rpay, int32_t* sfk,
                int32_t* spay);
void orc_gen_c2(uint64_t seed, uint64_t sub, uint64_t nR, uint64_t nS, double zipf_s, int64_t* rkey,
                int64_t* rpay, int64_t* sfk);
void orc_gen_c3(uint64_t seed, uint64_t nD, uint64_t nF, uint64_t norders, int32_t* d_orderkey,
                int32_t* d_custkey, int32_t* f_orderkey, double* f_price);
void orc_zipf_cdf(uint64_t n, double s, double* cdf);

/* ---------------- timed multithreaded baseline ------------------------- */
/* The reference plan queryhashjoin shape: T threads build one shared
 * chained table under bucket locks, then probe their slice and fold
 * count + sum of the build payload. Key/payload of width kb (4 or 8).
 * Returns wall seconds from scan start to the last finished thread. */
double orc_baseline_join(int kb, const void* rkey, const void* rpay, uint64_t nR, const void* skey, uint64_t nS,
                         unsigned buckets, unsigned tuplesperbucket, int threads, orc_cs64* out);

#ifdef __cplusplus
}
#endif

#endif
