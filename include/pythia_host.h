/*
 * pythia_host.h - C entry of libpythia_host.so, the plan engine that hosts
 * the `gpuhashjoin` / `gpuaggregate_sum` / `gpuaggregate_count` operators.
 *
 * pythia_run_query is the executequery driver (drivers/executequery.cpp:
 * 32-101) as a call: it parses a plan in libconfig syntax, builds the tree by
 * type string (query.cpp:62-233), runs threadInit / scanStart / getNext* /
 * scanStop / threadClose / destroy and returns every output row.
 */
#ifndef PYTHIA_HOST_H
#define PYTHIA_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pythia_result {
    void* rows;          /* packed output rows (malloc'ed)            */
    uint64_t nrows;
    uint32_t row_bytes;
    char schema[256];    /* e.g. "long,dec"                           */
    double seconds;      /* scanStart .. scanStop wall time           */
    uint64_t getnext_calls;
    /* How the GPU operators ran (no reference counterpart): the OR of the
     * fused forms the plan's aggregates took over resident joins (1: join
     * rows folded in HBM, 2: count / sum fused into the radix join, 4: group
     * sum fused into the partitioned group join), and the join output bytes
     * copied to host pages. */
    uint32_t gpu_fused;
    uint64_t join_d2h_bytes;
} pythia_result;

const char* pythia_last_error(void);
/* Makes rows available to plans as `type = "memscan"; table = "<name>";`. */
int pythia_register_table(const char* name, const void* rows, uint64_t nrows, uint32_t row_bytes);
int pythia_clear_tables(void);
/* Frees the device copies that plans with gpu { resident = true; } keep in
 * HBM across queries (no reference counterpart: the reference re-reads its
 * inputs into pages on every query). */
int pythia_clear_resident(void);
int pythia_run_query(const char* plan_text, pythia_result* out);
void pythia_free_result(pythia_result* r);

#ifdef __cplusplus
}
#endif

#endif
