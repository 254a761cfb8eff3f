// Process-wide registries of the host engine: named in-memory tables (the
// `memscan` source) and one shared pgpu context per device.
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "pgpu.h"

namespace pythia {

void register_table(const std::string& name, const void* rows, uint64_t nrows, unsigned row_bytes);
void clear_tables();
// `generation` (optional) receives a number that changes whenever the name is
// registered again, so a device copy of an older registration is not reused.
bool lookup_table(const std::string& name, unsigned row_bytes, std::vector<char>& out, uint64_t* generation = nullptr);

// Device copies of scan sources kept in HBM across queries.  `key` names the
// source's contents (file paths with sizes and modification times, or a
// registered table with its generation); a miss copies `bytes` from `rows`.
// The entry holds a reference on the device context until clear_resident().
const void* resident_rows(int device, uint64_t hbm_budget, const std::string& key, const char* rows, uint64_t bytes);
// A dense device column holding one 4- or 8-byte field of every row of a
// resident source (`d_rows`, the pointer resident_rows returned for `key`),
// extracted on first use and kept with the source's copy.
const void* resident_column(int device, uint64_t hbm_budget, const std::string& key, const void* d_rows, uint64_t n,
                            unsigned row_bytes, unsigned offset, unsigned bytes);
void clear_resident();

// Plans of the fused join consumers, kept between queries (creating one
// allocates its partition buffers): `fn` runs with the device's plan of
// this kind, at least as large as asked, held exclusively.  Released by
// clear_resident().
void with_join_plan(int device, uint64_t hbm_budget, const pgpu_join_cfg& cfg, uint64_t max_build,
                    uint64_t max_probe, const std::function<void(pgpu_join_plan*)>& fn);
void with_groupjoin(int device, uint64_t hbm_budget, uint64_t max_dim, uint64_t max_fact, uint64_t max_groups,
                    const std::function<void(pgpu_groupjoin*)>& fn);
// Pinned host buffers of at least `bytes` from a per-device pool (kept for
// reuse until clear_resident()).
void* acquire_pinned(int device, uint64_t hbm_budget, uint64_t bytes);
void release_pinned(void* p);

// Shared device context; the first acquirer's budget (0 = all free HBM)
// applies.  It is kept when its last user releases it (the next query reuses
// its pools) unless the next acquirer asks for another budget.  Asking for
// another budget while the context is held (by a running query or by
// resident device copies) throws.  Throws QueryExecutionError with the
// library's message on failure.
pgpu_ctx* acquire_device(int device, uint64_t hbm_budget);
void release_device(pgpu_ctx* ctx);

// Throws QueryExecutionError("<what>: <pgpu_last_error>") when rc != PGPU_OK.
void check(int rc, const char* what);

}  // namespace pythia
