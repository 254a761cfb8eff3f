/*
 * pgpu.h - C ABI of libpgpu.so, the MI355X (gfx950) hash-join and
 * hash-aggregation engine behind the `gpuhashjoin` / `gpuaggregate_sum`
 * plan operators.
 *
 * Rules of the boundary:
 *   - plain C types only: opaque handles, pointers, sizes; no C++ types,
 *     exceptions or STL cross it;
 *   - every call returns an int status (PGPU_OK = 0, negative = error); the
 *     message of the last failing call on the calling thread is returned by
 *     pgpu_last_error();
 *   - pointers named d_* are device (HBM) pointers, obtained from pgpu_alloc;
 *     every other pointer is host memory owned by the caller;
 *   - calls on distinct handles may run concurrently from different host
 *     threads; calls on one handle are serialised by its caller.
 *
 * The reference engine (sblanas/pythia) has no foreign-function interface:
 * its hash join and aggregation are C++ classes called through the Operator
 * virtuals (operators/operators.h:79-186).  Each declaration below names the
 * reference code it stands in for.
 */
#ifndef PGPU_H
#define PGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Status codes                                                        */
/* ------------------------------------------------------------------ */
#define PGPU_OK 0
#define PGPU_EINVAL -1       /* bad argument                                  */
#define PGPU_ENOMEM -2       /* device allocation failed                      */
#define PGPU_EBUDGET -3      /* HBM budget of the context exceeded            */
#define PGPU_EDEVICE -4      /* HIP runtime error                             */
#define PGPU_EOVERFLOW -5    /* a device-side capacity check tripped          */
#define PGPU_ESTATE -6       /* call out of lifecycle order                   */
#define PGPU_EUNSUPPORTED -7 /* column type / width not handled               */

/* ------------------------------------------------------------------ */
/* Context: one device, one HBM budget, the streams work is issued on   */
/* Replaces numaallocate_onnode / numaallocate_local and their byte     */
/* accounting (util/numaallocate.cpp:60-181, :402-457).                 */
/* ------------------------------------------------------------------ */
typedef struct pgpu_ctx pgpu_ctx;

const char* pgpu_version(void);
const char* pgpu_last_error(void);
int pgpu_device_count(int* n);
/* hbm_budget_bytes = 0 means "all free device memory". */
int pgpu_ctx_create(int device, uint64_t hbm_budget_bytes, pgpu_ctx** out);
int pgpu_ctx_destroy(pgpu_ctx* ctx);
/* The hipStream_t the context issues work on (for events / interop). */
int pgpu_ctx_stream(pgpu_ctx* ctx, void** hip_stream);
int pgpu_alloc(pgpu_ctx* ctx, uint64_t bytes, void** d_ptr);
int pgpu_free(pgpu_ctx* ctx, void* d_ptr);
int pgpu_hbm_usage(pgpu_ctx* ctx, uint64_t* used, uint64_t* budget, uint64_t* peak);
int pgpu_h2d(pgpu_ctx* ctx, void* d_dst, const void* src, uint64_t bytes);
int pgpu_d2h(pgpu_ctx* ctx, void* dst, const void* d_src, uint64_t bytes);
int pgpu_sync(pgpu_ctx* ctx);
/* Pinned (page-locked) host memory, outside the HBM budget: device copies
 * into it are direct DMA.  Stands in for the page buffers the operators
 * allocate with numaallocate_local (util/numaallocate.cpp:60-181). */
int pgpu_host_alloc(pgpu_ctx* ctx, uint64_t bytes, void** p);
int pgpu_host_free(pgpu_ctx* ctx, void* p);
/* Roofline calibration (no reference counterpart): sustained HBM GB/s of a
 * streaming read and of a read + write copy over `bytes`, `iters` times. */
int pgpu_membw(pgpu_ctx* ctx, uint64_t bytes, int iters, double* read_gbps, double* copy_gbps);

/* ------------------------------------------------------------------ */
/* Synthetic inputs generated in HBM (the formulas of pgpu_gen.h).      */
/* No reference counterpart: the reference reads text/binary files      */
/* (operators/loaders/table.cpp); these fill device columns directly.   */
/* ------------------------------------------------------------------ */
enum pgpu_gen_column {
    PGPU_GEN_C1_RKEY = 1,  /* int32 */
    PGPU_GEN_C1_RPAY,      /* int32 */
    PGPU_GEN_C1_SFK,       /* int32 */
    PGPU_GEN_C1_SPAY,      /* int32 */
    PGPU_GEN_C2_RKEY,      /* int64, sub = table id */
    PGPU_GEN_C2_RPAY,      /* int64 */
    PGPU_GEN_C2_SFK_UNIFORM, /* int64 */
    PGPU_GEN_C2_SFK_ZIPF,  /* int64, needs d_cdf */
    PGPU_GEN_C3_D_ORDERKEY, /* int32 */
    PGPU_GEN_C3_D_CUSTKEY, /* int32 */
    PGPU_GEN_C3_F_ORDERKEY, /* int32 */
    PGPU_GEN_C3_F_PARTKEY, /* int32 */
    PGPU_GEN_C3_F_QUANTITY, /* int32 */
    PGPU_GEN_C3_F_PRICE,   /* double */
    PGPU_GEN_C4_RKEY,      /* int32, sub = join id */
    PGPU_GEN_C4_RPAY,      /* int32 */
    PGPU_GEN_C4_SFK,       /* int32 */
    PGPU_GEN_C4_SPAY       /* int32 */
};

typedef struct pgpu_gen_spec {
    uint64_t seed;      /* base seed (42)                                 */
    uint64_t sub;       /* table / join index within a family             */
    uint64_t n_build;   /* build cardinality the column refers to         */
    uint64_t offset;    /* first row index generated                      */
    const double* d_cdf;/* Zipf CDF in HBM (n_build doubles) or NULL      */
} pgpu_gen_spec;

int pgpu_gen(pgpu_ctx* ctx, int column, const pgpu_gen_spec* spec, uint64_t n, void* d_out);
/* Host-side helpers of the same generators: the Zipf(s) CDF over n ranks
 * (upload it and pass it as d_cdf), and `count` log-uniform sizes
 * round(2^U(lo, hi)) of family `sub` starting at item `first` (C4 join
 * sizes use sub 4, C5 table sizes 5 and probe batch sizes 6). */
int pgpu_zipf_cdf(uint64_t n, double s, double* cdf);
int pgpu_sizes(uint64_t seed, uint64_t sub, uint64_t first, uint64_t count, double lo, double hi, uint64_t* out);

/* ------------------------------------------------------------------ */
/* Equi-join with fused count + sum consumer, device-resident columns.  */
/* Stands in for HashJoinOp::scanStart (build, operators/join.cpp:      */
/* 396-464, :641-667) + HashJoinOp::getNext (probe, :503-562) feeding   */
/* AggregateCount / AggregateSum over the join output                   */
/* (operators/aggregatecount.cpp:37-52, operators/aggregatesum.cpp:     */
/* 37-107) with no group-by (`fields = ()`).                            */
/* ------------------------------------------------------------------ */
enum pgpu_join_algo {
    PGPU_ALGO_AUTO = 0,   /* RADIX1 for builds of 2^16 to 2^22 rows (plans    */
                          /* of at most 2^22 build rows), GLOBAL below, else  */
                          /* RADIX                                            */
    PGPU_ALGO_RADIX = 1,  /* 2-pass radix partitioning + LDS tables         */
    PGPU_ALGO_GLOBAL = 2, /* one open-addressing table in HBM (no partitions) */
    PGPU_ALGO_RADIX1 = 3  /* one radix pass of <= 10 bits (partitions of     */
                          /* ~4096 build rows) + the batched-join LDS kernels */
};

typedef struct pgpu_join_cfg {
    uint32_t key_bytes;     /* 4 (int32) or 8 (int64); payloads have the same width */
    uint32_t radix_bits;    /* total radix bits (2..14) over both passes; 0 = choose */
    uint32_t probe_payload; /* 1: probe rows carry a payload summed per match        */
    uint32_t algo;          /* enum pgpu_join_algo                                   */
    uint32_t join_variant;  /* LDS join shape (slots x threads, probe rows in flight  */
                            /* per thread): 0/6 = 4096 x 512 x 6, 1 = 4096 x 512 x 8,*/
                            /* 2 = 4096 x 256 x 32, 3 = 2048 x 256 x 16, 4 = 4096 x  */
                            /* 512 x 16, 5 = 4096 x 512 x 8 at one workgroup per CU, */
                            /* 7 = 4096 x 512 x 6 with 2560-row pieces, 8 = 8192 x  */
                            /* 1024 x 6 with 5120-row pieces; the result does not    */
                            /* depend on it                                          */
} pgpu_join_cfg;

typedef struct pgpu_join_result {
    uint64_t matches;   /* |{(r, s) : r.key == s.key}|                         */
    uint64_t sum_build; /* sum of r.payload over the matches, modulo 2^64      */
    uint64_t sum_probe; /* sum of s.payload over the matches, modulo 2^64      */
    uint32_t radix_bits;/* bits actually used (0 for the global table)         */
    uint32_t lds_passes;/* max build pieces one partition needed (1 = fit LDS) */
} pgpu_join_result;

typedef struct pgpu_join_plan pgpu_join_plan;

/* Allocates every buffer a join of up to max_build x max_probe rows needs,
 * so later calls allocate nothing. */
int pgpu_join_plan_create(pgpu_ctx* ctx, const pgpu_join_cfg* cfg, uint64_t max_build,
                          uint64_t max_probe, pgpu_join_plan** out);
int pgpu_join_plan_destroy(pgpu_join_plan* plan);
/* Enqueue the whole join on the context stream; returns without waiting. */
int pgpu_join_count_sum_async(pgpu_join_plan* plan, const void* d_rkey, const void* d_rpay,
                              uint64_t n_build, const void* d_skey, const void* d_spay,
                              uint64_t n_probe);
/* Wait for the last enqueued join and copy its result. */
int pgpu_join_plan_result(pgpu_join_plan* plan, pgpu_join_result* out);
/* async + result. */
int pgpu_join_count_sum(pgpu_join_plan* plan, const void* d_rkey, const void* d_rpay,
                        uint64_t n_build, const void* d_skey, const void* d_spay,
                        uint64_t n_probe, pgpu_join_result* out);

/* The same join (count + sums, the global-table algorithm of
 * PGPU_ALGO_GLOBAL) with no allocation at all: no reference counterpart; it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Its hash
 * table, partial results and result words live in a 32 MiB device array of
 * the library, one per device, kept until the process ends; the inputs stay
 * the caller's.  Builds of at most 2^20 rows (4-byte keys) or 2^19 rows
 * (8-byte keys) are accepted, larger ones fail with PGPU_EINVAL.  Every call
 * on a device shares that device's array: calls that overlap on one device
 * (from several host threads or streams) race; each call clears what it
 * needs and waits for its result on the context's stream. */
int pgpu_join_count_sum_noalloc(pgpu_ctx* ctx, uint32_t key_bytes, uint32_t probe_payload, const void* d_rkey,
                                const void* d_rpay, uint64_t n_build, const void* d_skey, const void* d_spay,
                                uint64_t n_probe, pgpu_join_result* out);

/* Per-kernel timing with HIP events on the stream the kernels run on.
 * enable != 0 starts recording (and clears totals); pgpu_join_plan_times
 * returns, per kernel, its name, launch count and summed milliseconds. */
typedef struct pgpu_kernel_time {
    char name[48];
    uint64_t launches;
    double total_ms;
} pgpu_kernel_time;
int pgpu_join_plan_timing(pgpu_join_plan* plan, int enable);
int pgpu_join_plan_times(pgpu_join_plan* plan, pgpu_kernel_time* out, int cap, int* n);

/* ------------------------------------------------------------------ */
/* Batched joins: many independent joins (count + sums, as above) of     */
/* uneven sizes bound into one plan and run as one hipGraph: segmented   */
/* radix partitioning of every join at once, then one persistent kernel  */
/* pulling (join, partition) items from a work queue.                    */
/* Stands in for running the queryhashjoin plan (HashJoinOp ->           */
/* AggregateCount / AggregateSum, operators/join.cpp:396-562,            */
/* aggregatesum.cpp:81-107) once per join; the reference has no batched  */
/* form.                                                                 */
/* ------------------------------------------------------------------ */
typedef struct pgpu_batch_join {
    const void* d_rkey; /* build keys (key_bytes wide)                     */
    const void* d_rpay; /* build payloads (same width)                     */
    uint64_t n_build;   /* < 2^32                                          */
    const void* d_skey; /* probe keys                                      */
    const void* d_spay; /* probe payloads (read when probe_payload != 0)   */
    uint64_t n_probe;   /* < 2^32                                          */
} pgpu_batch_join;

typedef struct pgpu_batch_cfg {
    uint32_t key_bytes;     /* 4 or 8                                               */
    uint32_t probe_payload; /* 1: sum probe payloads of matches too                 */
    uint32_t variant;       /* join workgroup shape: 1 = 4096-slot LDS table x 512   */
                            /* threads, 2 = 8192 slots x 1024 threads, 3 (4-byte    */
                            /* keys) = 8192 8-byte (key, payload) slots x 512       */
                            /* threads, two per CU, for partitions of unique build  */
                            /* keys, the others handed to variant 2's kernel; 0 = 3 */
                            /* for 4-byte keys, else 2                              */
    uint32_t part_rows;     /* target build rows per partition; 0 = variant default */
    uint32_t graph;         /* 1: capture the launch sequence into a hipGraph       */
} pgpu_batch_cfg;

typedef struct pgpu_batch_info_t {
    pgpu_batch_cfg cfg;     /* the configuration in force (defaults resolved) */
    uint32_t items;         /* work items of the persistent join             */
    uint32_t grid;          /* persistent workgroups                          */
    uint64_t build_rows, probe_rows, chunks;
    uint32_t graph;         /* 1 if replayed as a hipGraph                    */
} pgpu_batch_info_t;

typedef struct pgpu_tune_entry {
    uint32_t variant, part_rows;
    double ms;              /* mean batch time of this candidate */
} pgpu_tune_entry;

typedef struct pgpu_batch_plan pgpu_batch_plan;
/* The plan keeps the join descriptors' device pointers: the columns must
 * stay allocated and keep their sizes for the plan's lifetime. */
int pgpu_batch_plan_create(pgpu_ctx* ctx, const pgpu_batch_cfg* cfg, const pgpu_batch_join* joins, uint32_t n,
                           pgpu_batch_plan** out);
int pgpu_batch_plan_destroy(pgpu_batch_plan* plan);
/* Enqueues the whole batch on the context stream (a graph launch). */
int pgpu_batch_run_async(pgpu_batch_plan* plan);
/* Waits for the batch; out (may be NULL) receives n results in join order. */
int pgpu_batch_wait(pgpu_batch_plan* plan, pgpu_join_result* out);
int pgpu_batch_info(pgpu_batch_plan* plan, pgpu_batch_info_t* info);
/* Per-kernel HIP-event timing (runs the launches directly, not the graph). */
int pgpu_batch_timing(pgpu_batch_plan* plan, int enable);
int pgpu_batch_times(pgpu_batch_plan* plan, pgpu_kernel_time* out, int cap, int* n);
/* Autotuner: times every (variant, part_rows) candidate on this batch
 * (reps graph replays each after a warm-up) and returns the fastest in
 * *best (key_bytes / probe_payload / graph taken from base). */
int pgpu_batch_autotune(pgpu_ctx* ctx, const pgpu_batch_cfg* base, const pgpu_batch_join* joins, uint32_t n,
                        int reps, pgpu_batch_cfg* best, pgpu_tune_entry* entries, int cap, int* n_entries);

/* ------------------------------------------------------------------ */
/* Serving: join + sum queries against build tables kept resident in    */
/* HBM, admitted FIFO under an HBM budget and continuously batched onto */
/* overlapping streams (PGPU_SERVE_BATCHED), or run one at a time       */
/* through the per-query join plan that rebuilds the table              */
/* (PGPU_SERVE_PLAIN).  Stands in for executing the queryhashjoin plan  */
/* per query (drivers/executequery.cpp:38-59 over operators/join.cpp:   */
/* 396-562 and aggregatesum.cpp:81-107), which rebuilds the HashTable   */
/* (util/hashtable.cpp:57-97) every time; the reference has no server.  */
/* Keys and payloads are int64; a query's probe batch is device memory  */
/* owned by the caller and held until its result is polled.             */
/* ------------------------------------------------------------------ */
enum pgpu_serve_mode { PGPU_SERVE_BATCHED = 0, PGPU_SERVE_PLAIN = 1 };

typedef struct pgpu_server_cfg {
    uint64_t hbm_budget;  /* resident tables + admitted probe batches; 0 = the context's */
    uint32_t key_bytes;   /* 8                                                          */
    uint32_t mode;        /* enum pgpu_serve_mode                                       */
    uint32_t slots;       /* batches in flight at once (streams); 0 = 3                 */
    uint32_t max_batch;   /* queries per launch; 0 = 256                                */
    uint32_t max_items;   /* random-probe work items per launch, counted as 16K-row    */
                          /* items at admission (smaller ones when they fit); 0 = 32768 */
    uint32_t part_min_rows; /* probe batches of at least this many rows (and a quarter */
                            /* of their table) take the partitioned path; 0 = 2^20,    */
                            /* 1 = every batch, UINT32_MAX = none                      */
} pgpu_server_cfg;

typedef struct pgpu_served_result {
    uint64_t tag;           /* the caller's tag                          */
    uint32_t table;
    uint32_t batch_size;    /* queries that shared its launch            */
    uint64_t m;             /* probe rows                                */
    uint64_t matches;
    uint64_t sum_build;     /* sum of build payloads over the matches    */
    double t_arrival;       /* seconds on the server clock               */
    double t_admit;
    double t_done;
    uint64_t batch;         /* launch sequence number                    */
} pgpu_served_result;

typedef struct pgpu_server_stats {
    uint64_t tables, resident_bytes, reserved_bytes, budget;
    uint64_t submitted, completed, batches, queued;
    uint64_t partitioned;   /* queries that took the partitioned path */
} pgpu_server_stats;

typedef struct pgpu_server pgpu_server;
int pgpu_server_create(pgpu_ctx* ctx, const pgpu_server_cfg* cfg, pgpu_server** out);
int pgpu_server_destroy(pgpu_server* s);
/* Registers a table (int64 key / payload columns in HBM, kept allocated by
 * the caller); batched mode hashes it into a resident table now.  Tables are
 * registered while no query is outstanding. */
int pgpu_server_add_table(pgpu_server* s, const void* d_key, const void* d_pay, uint64_t n, uint32_t* table_id);
/* Plain mode: sizes the per-query join plan for probes up to max_probe rows. */
int pgpu_server_prepare_plain(pgpu_server* s, uint64_t max_probe);
/* Queues a query (thread-safe, returns at once).  t_arrival: seconds on the
 * server clock (pgpu_server_now), 0 = now; latency is t_done - t_arrival. */
int pgpu_server_submit(pgpu_server* s, uint32_t table, const void* d_probe, uint64_t m, uint64_t tag,
                       double t_arrival);
/* Waits until every submitted query has completed. */
int pgpu_server_drain(pgpu_server* s);
/* Takes up to cap completed results (completion order). */
int pgpu_server_poll(pgpu_server* s, pgpu_served_result* out, int cap, int* n);
int pgpu_server_now(pgpu_server* s, double* t);
/* Load generator: submits query i at offsets[i] seconds after the call
 * (busy-waiting for each arrival), drains, and returns result i in out[i]. */
int pgpu_server_replay(pgpu_server* s, const uint32_t* tables, const void* const* probes, const uint64_t* ms,
                       const double* offsets, uint64_t n, pgpu_served_result* out);
int pgpu_server_stats_get(pgpu_server* s, pgpu_server_stats* st);
/* Persisting resident tables (batched mode): a table's radix-partitioned
 * (key, payload) rows and their bounds saved as one memory segment at
 * `path` (/dev/shm/<name> through shm_open, any other path through open)
 * and attached by a server without the table's columns and without the
 * partition pass (a warm start: the rows are uploaded and the hash slots
 * derived from them by one kernel).  Stand in for HashTable::serialize /
 * deserialize (util/hashtable.cpp:99-201, segments made by
 * creatememsegment / attachmemsegment) and the segment writer
 * (operators/memsegmentwriter.cpp:154-462).  Loading follows the rules of
 * pgpu_server_add_table (before any query; the HBM budget applies). */
int pgpu_server_save_table(pgpu_server* s, uint32_t table, const char* path);
int pgpu_server_load_table(pgpu_server* s, const char* path, uint32_t* table_id);
/* The C5 query trace: query q's table (uniform over ntables), probe rows
 * round(2^U(lo, hi)) and Poisson arrival offset (seconds, rate per second). */
int pgpu_c5_trace(uint64_t seed, uint64_t count, uint32_t ntables, double lo, double hi, double rate,
                  uint32_t* tables, uint64_t* ms, double* offsets);

/* ------------------------------------------------------------------ */
/* Row-format hash join with materialised output (gpuhashjoin).         */
/* Replaces HashJoinOp's build (buildFromPage into util/hashtable,       */
/* operators/join.cpp:641-667) and probe/projection loop                 */
/* (getNext + constructOutputTuple, operators/join.cpp:475-562).         */
/* Build rows are the reference's `sbuild` layout (join key then the     */
/* projected build columns, packed); probe rows are the probe child's    */
/* tuples.  Output rows are the projection, packed, in the output        */
/* schema's order.  Keys compare as the reference's Comparator does:    */
/* integers of 4 or 8 bytes by value, dec keys as doubles, char keys by   */
/* strncmp (see pgpu_hj_desc).                                            */
/* ------------------------------------------------------------------ */
/* Column types where a kernel interprets a column (join keys, sort keys). */
enum pgpu_col_type { PGPU_COL_INT32 = 1, PGPU_COL_INT64 = 2, PGPU_COL_F64 = 3, PGPU_COL_CHAR = 4 };

typedef struct pgpu_proj_col {
    uint32_t side;       /* 0: build (sbuild) row, 1: probe row */
    uint32_t src_offset; /* byte offset in that row              */
    uint32_t bytes;      /* column width                         */
} pgpu_proj_col;

#define PGPU_MAX_PROJ 32

typedef struct pgpu_hj_desc {
    uint32_t build_row_bytes, build_key_offset, build_key_bytes;
    uint32_t probe_row_bytes, probe_key_offset, probe_key_bytes;
    uint32_t nproj;
    pgpu_proj_col proj[PGPU_MAX_PROJ];
    uint32_t slots;      /* independent probe streams (worker threads), >= 1 */
    /* Key types (enum pgpu_col_type; 0 = a signed integer of key_bytes).
     * Equality follows the reference's Comparator (comparator.cpp:43-413,
     * rawcompfns.cpp): integer keys of either width by value; a dec (FP64)
     * key against a dec or integer key as doubles (so +0 equals -0, a NaN
     * key matches nothing); char keys against char keys as strncmp over the
     * smaller key width (bytes up to the first NUL).  Other pairs are
     * refused (PGPU_EUNSUPPORTED). */
    uint32_t build_key_type, probe_key_type;
} pgpu_hj_desc;

typedef struct pgpu_hj_stats {
    uint64_t build_rows, probe_rows, output_rows;
    uint64_t table_slots;
    uint64_t h2d_bytes, d2h_bytes;
    uint64_t probe_reruns;   /* output buffer regrowths */
} pgpu_hj_stats;

typedef struct pgpu_hj pgpu_hj;
int pgpu_hj_create(pgpu_ctx* ctx, const pgpu_hj_desc* desc, pgpu_hj** out);
int pgpu_hj_destroy(pgpu_hj* h);
/* Thread-safe: any worker may append build rows (copied before return). */
int pgpu_hj_build_append(pgpu_hj* h, const void* rows, uint64_t n);
/* Uploads the staged build rows and builds the device table. */
int pgpu_hj_build_finish(pgpu_hj* h);
/* Joins n probe rows on stream `slot`; the output stays in HBM and is read
 * with pgpu_hj_pull.  *n_out receives the number of output rows. */
int pgpu_hj_probe(pgpu_hj* h, uint32_t slot, const void* rows, uint64_t n, uint64_t* n_out);
/* Copies up to cap_rows further output rows of `slot` into dst. */
int pgpu_hj_pull(pgpu_hj* h, uint32_t slot, void* dst, uint64_t cap_rows, uint64_t* n_rows);
/* Device-resident inputs and outputs (tables kept in HBM across queries, an
 * aggregate fused over the join): no reference counterpart — the reference
 * moves every tuple through host pages.  build_device builds the table from
 * n rows already in device memory (they must stay valid until the handle is
 * destroyed; exclusive with build_append); probe_device joins n device rows;
 * output hands out the device address of the slot's not yet pulled output
 * rows (valid until the slot's next probe) and marks them pulled. */
int pgpu_hj_build_device(pgpu_hj* h, const void* d_rows, uint64_t n);
int pgpu_hj_probe_device(pgpu_hj* h, uint32_t slot, const void* d_rows, uint64_t n, uint64_t* n_out);
int pgpu_hj_output(pgpu_hj* h, uint32_t slot, const void** d_rows, uint64_t* n_rows);
int pgpu_hj_stats_get(pgpu_hj* h, pgpu_hj_stats* out);

/* One 4- or 8-byte field of n packed rows in device memory (row_bytes apart,
 * at `offset`) copied into a dense device column d_out (aligned to the field
 * width).  Stands in for reading a column through Schema::calcOffset /
 * asInt / asLong / asDecimal (schema.inl:141-202) on every tuple: the
 * operators transpose resident rows into the columns the partitioned
 * kernels stream. */
int pgpu_rows_column(pgpu_ctx* ctx, const void* d_rows, uint64_t n, uint32_t row_bytes, uint32_t offset,
                     uint32_t bytes, void* d_out);

/* ------------------------------------------------------------------ */
/* Row-format hash aggregation (gpuaggregate_sum / gpuaggregate_count). */
/* Replaces GenericAggregate::remember + AggregateSum/AggregateCount    */
/* fold (operators/genericaggregate.cpp:240-282, aggregatesum.cpp:      */
/* 53-107, aggregatecount.cpp:44-52) and the group emission of          */
/* GenericAggregate::getNext (:419-487).  Group keys compare bytewise;   */
/* the accumulator has the summed column's type (int wraps mod 2^32,    */
/* long mod 2^64, dec is an FP64 sum); count is a long.                 */
/* ------------------------------------------------------------------ */
enum pgpu_agg_kind { PGPU_AGG_SUM_I32 = 1, PGPU_AGG_SUM_I64, PGPU_AGG_SUM_F64, PGPU_AGG_COUNT };

typedef struct pgpu_agg_desc {
    uint32_t row_bytes;
    uint32_t nkeys;            /* 0 = one group (no GROUP BY)      */
    uint32_t key_offset[8];
    uint32_t key_bytes[8];     /* total key bytes <= 64            */
    uint32_t kind;             /* enum pgpu_agg_kind               */
    uint32_t val_offset;       /* summed column (unused for count) */
    uint64_t groups_hint;      /* initial table sizing, 0 = 64K    */
} pgpu_agg_desc;

typedef struct pgpu_agg pgpu_agg;
int pgpu_agg_create(pgpu_ctx* ctx, const pgpu_agg_desc* desc, pgpu_agg** out);
int pgpu_agg_destroy(pgpu_agg* a);
/* Thread-safe; rows are copied to HBM and folded before return. */
int pgpu_agg_append(pgpu_agg* a, const void* rows, uint64_t n);
/* As pgpu_agg_append for rows already in device memory (for example a join's
 * pgpu_hj_output); no reference counterpart.  Each aggregate folds on a
 * stream of its own, ordered after work already queued on the context's
 * stream (such as pgpu_gen writing the rows). */
int pgpu_agg_append_device(pgpu_agg* a, const void* d_rows, uint64_t n);
/* Compacts the groups; output rows = key columns then the accumulator. */
int pgpu_agg_finish(pgpu_agg* a, uint64_t* n_groups, uint32_t* out_row_bytes);
/* Copies output rows [first, first + n) into dst. */
int pgpu_agg_pull(pgpu_agg* a, uint64_t first, uint64_t n, void* dst);

/* ------------------------------------------------------------------ */
/* Row-format range partitioning (gpupartition).  Stands in for          */
/* PartitionOp (operators/partition.cpp:44-404) with its exactrange hash  */
/* (ExactRangeValueHasher, hash.h:444-479):                              */
/*   bucket = (key - min) / ((max - min + buckets) / buckets)            */
/* in signed 64-bit arithmetic truncating toward zero, key = the signed  */
/* 4- or 8-byte column at key_offset (an 8-byte dec column is taken by   */
/* its bits, as numericalize does).  Output bucket b = every input row   */
/* of bucket b in input order, source 0's rows first (the reference's    */
/* idxstart order over worker threads), optionally sorted by a column    */
/* (sortAllInPage; ties keep input order).                               */
/* ------------------------------------------------------------------ */
#define PGPU_PART_MAX_BUCKETS 127 /* PartitionOp: buckets < MAX_THREADS (128) */
#define PGPU_PART_MAX_SOURCES 256

typedef struct pgpu_part_desc {
    uint32_t row_bytes;
    uint32_t key_offset, key_bytes; /* key_bytes 4 or 8                     */
    uint32_t buckets;               /* 1..PGPU_PART_MAX_BUCKETS             */
    int64_t range_min, range_max;
    uint32_t sort;                  /* 0: keep input order in a bucket      */
    uint32_t sort_offset, sort_type; /* enum pgpu_col_type                  */
    uint32_t pad;
} pgpu_part_desc;

typedef struct pgpu_part pgpu_part;
int pgpu_part_create(pgpu_ctx* ctx, const pgpu_part_desc* desc, pgpu_part** out);
int pgpu_part_destroy(pgpu_part* p);
/* Thread-safe: stages n rows of input stream `source` (copied before
 * return).  Streams are concatenated in source order. */
int pgpu_part_append(pgpu_part* p, uint32_t source, const void* rows, uint64_t n);
/* Uploads the staged rows and partitions (and sorts) them on the device;
 * counts (optional) receives the rows per bucket.  A key outside the
 * buckets (the reference indexes past its histogram there) fails with
 * PGPU_EOVERFLOW. */
int pgpu_part_finish(pgpu_part* p, uint64_t* counts);
/* Copies rows [first, first + n) of bucket b into dst. */
int pgpu_part_pull(pgpu_part* p, uint32_t bucket, uint64_t first, uint64_t n, void* dst);
/* Drops the staged input and the result; the handle accepts new rows. */
int pgpu_part_reset(pgpu_part* p);

/* ------------------------------------------------------------------ */
/* Join + group-by FP64 sum over device-resident columns (C3 shape):    */
/*   SUM(F.val) GROUP BY D.grp over F JOIN D ON F.key = D.key            */
/* Stands in for the plan hashjoin(build D, probe F, projection          */
/* [B$grp, P$val]) -> aggregate_sum(field 0, sumfield 1): HashJoinOp's   */
/* probe loop (operators/join.cpp:503-562) feeding                       */
/* GenericAggregate::remember / AggregateSum::fold                       */
/* (operators/genericaggregate.cpp:240-282, aggregatesum.cpp:81-107).    */
/* Keys and groups are int32; sums are FP64 (order of additions is not   */
/* fixed: results agree with exact sums within rounding).                */
/* ------------------------------------------------------------------ */
typedef struct pgpu_groupjoin pgpu_groupjoin;
int pgpu_groupjoin_create(pgpu_ctx* ctx, uint64_t max_dim, uint64_t max_fact, uint64_t max_groups,
                          pgpu_groupjoin** out);
int pgpu_groupjoin_destroy(pgpu_groupjoin* p);
/* Runs the whole query on the context stream and waits; *n_groups receives
 * the number of result groups, *n_joined (optional) the joined fact rows. */
int pgpu_groupjoin_run(pgpu_groupjoin* p, const int32_t* d_dkey, const int32_t* d_dgrp, uint64_t n_dim,
                       const int32_t* d_fkey, const double* d_fval, uint64_t n_fact, uint64_t* n_groups,
                       uint64_t* n_joined);
/* Copies the first n result groups (any of the outputs may be NULL). */
int pgpu_groupjoin_fetch(pgpu_groupjoin* p, int32_t* keys, double* sums, uint64_t* counts, uint64_t n);
/* The algorithm the last run used: PGPU_GJ_PARTITIONED (unique dimension
 * keys within a range of 2^25 values, group keys within 2^22: two radix
 * passes, LDS lookups and LDS sums) or PGPU_GJ_GENERAL (HBM hash tables and
 * FP64 atomics: any keys, duplicate dimension keys). */
#define PGPU_GJ_PARTITIONED 1
#define PGPU_GJ_GENERAL 2
int pgpu_groupjoin_path(pgpu_groupjoin* p, int* path);
/* Result rows [first, first + n) of the last run as the aggregate over the
 * join emits them (GenericAggregate's output: group column then the
 * accumulator, genericaggregate.cpp:96-104): the int32 group key followed by
 * the FP64 sum (PGPU_GJ_ROW_SUM, AggregateSum over a dec column,
 * aggregatesum.cpp:46-49) or the int64 count (PGPU_GJ_ROW_COUNT,
 * aggregatecount.cpp:37-42), packed into 12-byte rows, into host memory
 * dst.  When dst lies in a block from pgpu_host_alloc and is 16-byte
 * aligned the rows are written there by the kernel that packs them;
 * otherwise they are packed in HBM and copied.  Returns when they are in
 * place. */
#define PGPU_GJ_ROW_SUM 1
#define PGPU_GJ_ROW_COUNT 2
int pgpu_groupjoin_rows(pgpu_groupjoin* p, int acc, uint64_t first, uint64_t n, void* dst);
int pgpu_groupjoin_timing(pgpu_groupjoin* p, int enable);
int pgpu_groupjoin_times(pgpu_groupjoin* p, pgpu_kernel_time* out, int cap, int* n);

#ifdef __cplusplus
}
#endif

#endif /* PGPU_H */
