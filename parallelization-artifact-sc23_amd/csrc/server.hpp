// Serving loop (indigo_serve): continuous batching of requests on slots.
#pragma once

#include "runtime.hpp"

namespace indigo {

// Replays reqs in real time (arguments validated by the caller) and fills in
// their admission / completion times.
int serve_run(const indigo_graph_ref* graphs, int ngraphs, indigo_request* reqs, int nreqs,
              const indigo_serve_config* cfg, indigo_serve_stats* stats);

}  // namespace indigo
