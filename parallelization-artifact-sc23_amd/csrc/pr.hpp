// Host launchers of the PageRank kernel.
#pragma once

#include <cstddef>

#include "runtime.hpp"

namespace indigo {

size_t pr_workspace_bytes(int nodes, int edges);
// Enqueues PageRank on ctx->stream.  d_scores: initial scores when
// init_uniform == 0, else overwritten with 1/nodes first; result on exit.
// d_degree may be NULL (degrees from nindex).
int pr_run(indigo_ctx* ctx, int nodes, int edges, const int* d_nindex, const int* d_nlist, const int* d_degree,
           float* d_scores, int init_uniform);
int pr_collect(indigo_ctx* ctx, int* iterations, double* last_diff);

}  // namespace indigo
