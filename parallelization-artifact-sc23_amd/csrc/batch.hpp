// Host launcher of the batched multi-graph / multi-source engine.
#pragma once

#include "runtime.hpp"

namespace indigo {

// Runs every query of the batch in one persistent launch on ctx->stream and
// waits for it.  Validates the batch first (INDIGO_ERR_INVALID_ARG).
int batch_run(indigo_ctx* ctx, const indigo_graph_ref* graphs, int ngraphs, const indigo_query* queries,
              int nqueries, int max_workgroups);

}  // namespace indigo
