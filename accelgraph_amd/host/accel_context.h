/* accel_context.h -- process-wide device and device-graph cache of the drop-in entry points. */
#ifndef ACCEL_CONTEXT_H
#define ACCEL_CONTEXT_H

#include "accelgraph.h"
#include "graphCSR.h"
#include "accelgraph_compat.h"

#ifdef __cplusplus
extern "C" {
#endif

ag_device *accelgraphDevice(void);
ag_dgraph *accelgraphUploadGraph(const struct GraphCSR *graph, int with_weights);
void accelgraphFail(const char *what, int status);

#ifdef __cplusplus
}
#endif

#endif
