// Device-side helpers of the graph plumbing (graphio.hip).
#pragma once

#include "runtime.hpp"

namespace indigo {

// Enqueues the edge-source array d_sp[j] = v for every CSR entry j of vertex
// v on ctx->stream (no timing; indigo_build_sp is the timed C entry).
int build_sp_async(indigo_ctx* ctx, int nodes, int edges, const int* d_nindex, int* d_sp);

}  // namespace indigo
