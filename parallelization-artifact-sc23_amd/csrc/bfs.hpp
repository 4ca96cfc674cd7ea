// Host launchers of the BFS kernels.
#pragma once

#include <cstddef>

#include "runtime.hpp"

namespace indigo {

struct BfsStats {
  int launches = 0;
  int grid = 0;
  int levels = 0;
  int bottom_up_levels = 0;
  int small_levels = 0;
};

size_t bfs_workspace_bytes(int nodes);
// Enqueues the BFS from `src` on ctx->stream (workspace must be sized).
int bfs_run(indigo_ctx* ctx, int nodes, int edges, const int* d_nindex, const int* d_nlist, int src, int* d_dist,
            BfsStats* stats);
// Waits for the BFS and reads its control block (levels, status).
int bfs_collect(indigo_ctx* ctx, BfsStats* stats);

}  // namespace indigo
