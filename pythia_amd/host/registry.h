// Process-wide registries of the host engine: named in-memory tables (the
// `memscan` source) and one shared pgpu context per device.
#pragma once

#include <cstdint>
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
void clear_resident();

// Shared device context; the first acquirer's budget (0 = all free HBM)
// applies.  It is kept when its last user releases it (the next query reuses
// its pools) unless the next acquirer asks for another budget.  Throws
// QueryExecutionError with the library's message on failure.
pgpu_ctx* acquire_device(int device, uint64_t hbm_budget);
void release_device(pgpu_ctx* ctx);

// Throws QueryExecutionError("<what>: <pgpu_last_error>") when rc != PGPU_OK.
void check(int rc, const char* what);

}  // namespace pythia
