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
bool lookup_table(const std::string& name, unsigned row_bytes, std::vector<char>& out);

// Shared device context; the first acquirer's budget (0 = all free HBM)
// applies.  Throws QueryExecutionError with the library's message on failure.
pgpu_ctx* acquire_device(int device, uint64_t hbm_budget);
void release_device(pgpu_ctx* ctx);

// Throws QueryExecutionError("<what>: <pgpu_last_error>") when rc != PGPU_OK.
void check(int rc, const char* what);

}  // namespace pythia
