/*
 * accelgraph_compat.h -- lifetime controls of the drop-in entry points (pageRank.h, SPMV.h).
 *
 * The drop-in entry points open the GPU once per process and keep the last graph they
 * uploaded resident in HBM (keyed by the GraphCSR pointer, its sizes, its array pointers and
 * a digest of the arrays' contents taken on the host at every call, so a graph rebuilt at the
 * same addresses or changed in place is uploaded again; ACCELGRAPH_TRUST_GRAPH_POINTERS=1
 * drops the digest for callers that release the graph themselves). These calls have no
 * counterpart in the reference, which attached and released its FPGA inside every call
 * (02_capi_graph/src/capi_utils/capienv.c:329-451, 765-782).
 */
#ifndef ACCELGRAPH_COMPAT_H
#define ACCELGRAPH_COMPAT_H

#ifdef __cplusplus
extern "C" {
#endif

/* Drop the cached device copy of the last graph (needed before mutating or freeing a graph
 * only under ACCELGRAPH_TRUST_GRAPH_POINTERS=1). */
void accelgraphReleaseGraph(void);
/* Release the cached graph and close the device. */
void accelgraphShutdown(void);

#ifdef __cplusplus
}
#endif

#endif
