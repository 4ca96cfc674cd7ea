// Host launcher of the batched multi-graph / multi-source engine.
#pragma once

#include "runtime.hpp"

namespace indigo {

// Runs every query of the batch in one persistent launch on ctx->stream and
// waits for it.  Validates the batch first (INDIGO_ERR_INVALID_ARG).
int batch_run(indigo_ctx* ctx, const indigo_graph_ref* graphs, int ngraphs, const indigo_query* queries,
              int nqueries, int max_workgroups);
// The same in two halves: batch_launch validates, uploads the task table and
// launches without waiting (one batch in flight per context); batch_ready
// polls; batch_wait waits, checks the run and relaunches a grid that did not
// assemble.
int batch_launch(indigo_ctx* ctx, const indigo_graph_ref* graphs, int ngraphs, const indigo_query* queries,
                 int nqueries, int max_workgroups);
int batch_ready(indigo_ctx* ctx, int* ready);
int batch_wait(indigo_ctx* ctx);
// Workgroups that any persistent kernel of the library can keep resident on
// the whole device (the base of the serving slots' shares).
int batch_grid_limit(indigo_ctx* ctx, int* limit);
// The engine's gang size for a query (before clamping to the grid).
int batch_gang(int alg, int nodes, int edges);

}  // namespace indigo
