// Host launchers of the connected-components kernels.
#pragma once

#include <cstddef>

#include "runtime.hpp"

namespace indigo {

size_t cc_workspace_bytes(int nodes);
// Enqueues CC on ctx->stream; d_label[v] = min vertex id of v's component.
int cc_run(indigo_ctx* ctx, int nodes, int edges, const int* d_nindex, const int* d_nlist, int* d_label,
           int* launches);
// The same with the workspace at `ws` (cc_workspace_bytes(nodes) bytes).
int cc_run_in(indigo_ctx* ctx, void* ws, int nodes, int edges, const int* d_nindex, const int* d_nlist, int* d_label,
              int* launches);

// Largest graph of the allocation-free entry: its scratch holds
// cc_workspace_bytes(kCcScratchMaxNodes), about 64 MiB per device.
constexpr int kCcScratchMaxNodes = 1 << 22;
// CC on the file-scope scratch array instead of the context workspace:
// no device allocation; waits for the result (indigo_cc_noalloc).
int cc_noalloc_device(indigo_ctx* ctx, int nodes, int edges, const int* d_nindex, const int* d_nlist, int* d_label);

}  // namespace indigo
