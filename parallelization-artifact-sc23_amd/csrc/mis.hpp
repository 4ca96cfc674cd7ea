// Host launchers of the MIS kernels.
#pragma once

#include <cstddef>

#include "runtime.hpp"

namespace indigo {

size_t mis_workspace_bytes(int nodes);
// Enqueues MIS on ctx->stream; d_status[v] = 1 (included) or 2 (excluded).
int mis_run(indigo_ctx* ctx, int nodes, const int* d_nindex, const int* d_nlist, int signed_priority,
            int* d_status);
// Waits and reads the round count / status.
int mis_collect(indigo_ctx* ctx, int* rounds);

}  // namespace indigo
