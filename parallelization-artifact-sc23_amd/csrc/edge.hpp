// Host launchers of the edge-centric, topology-driven variants (edge.hip).
#pragma once

#include <cstddef>

#include "runtime.hpp"

namespace indigo {

enum EdgeAlg { kEdgeBfs = 0, kEdgeSssp = 1, kEdgeCc = 2, kEdgeMis = 3 };

// Workgroup size of the edge kernels for ctx (INDIGO_HINT_BLOCK_SIZE: 256,
// 512 or 1024; anything else 512).
int edge_block(const indigo_ctx* ctx);
size_t edge_workspace_bytes(int nodes, int edges, bool build_sp);
// Runs `alg` edge-centrically and waits for it: d_out receives BFS / SSSP
// distances, CC labels or MIS status.  d_sp is the caller's device edge-source
// array, or NULL to build one in the workspace (counted in the call's time).
// eweight is read by SSSP only; src by BFS / SSSP; signed_priority by MIS.
int edge_device(indigo_ctx* ctx, int alg, int nodes, int edges, const int* d_nindex, const int* d_nlist,
                const int* d_eweight, const int* d_sp, int src, int signed_priority, int* d_out, double* seconds);

}  // namespace indigo
