// Host launchers of the SSSP kernel.
#pragma once

#include <cstddef>

#include "runtime.hpp"

namespace indigo {

size_t sssp_workspace_bytes(int nodes);
// Enqueues SSSP from src on ctx->stream (weights d_eweight, int32).
int sssp_run(indigo_ctx* ctx, int nodes, int edges, const int* d_nindex, const int* d_nlist, const int* d_eweight,
             int src, int* d_dist);
int sssp_collect(indigo_ctx* ctx, int* iterations, int* epochs);

}  // namespace indigo
