// Host launchers of the connected-components kernels.
#pragma once

#include <cstddef>

#include "runtime.hpp"

namespace indigo {

size_t cc_workspace_bytes(int nodes);
// Enqueues CC on ctx->stream; d_label[v] = min vertex id of v's component.
int cc_run(indigo_ctx* ctx, int nodes, int edges, const int* d_nindex, const int* d_nlist, int* d_label,
           int* launches);

}  // namespace indigo
