// Host launchers of the triangle-counting kernels.
#pragma once

#include <cstddef>

#include "runtime.hpp"

namespace indigo {

size_t tc_workspace_bytes(int nodes, int edges);
// Enqueues TC on ctx->stream.
int tc_run(indigo_ctx* ctx, int nodes, int edges, const int* d_nindex, const int* d_nlist, int* launches);
// Waits and reads the count (INDIGO_ERR_UNSORTED for unsorted / self-loop input).
int tc_collect(indigo_ctx* ctx, long long* count);

}  // namespace indigo
