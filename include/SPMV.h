/*
 * SPMV.h -- sparse matrix-vector pull entry points on struct GraphCSR (drop-in).
 *
 * Replaces the CSR part of 02_capi_graph/include/algorithms/capi/SPMV.h:
 *   struct SPMVStats              SPMV.h:19-28   (same fields, same order)
 *   newSPMVStatsGraphCSR          SPMV.h:30      (SPMV.c:51-73)
 *   freeSPMVStats                 SPMV.h:35      (SPMV.c:145-158)
 *   SPMVGraphCSR                  SPMV.h:51      (SPMV.c:594-621)
 *   SPMVPullGraphCSR              SPMV.h:52      (SPMV.c:622-751)
 *   SPMVPullFixedPointGraphCSR    SPMV.h:55      (SPMV.c:849-1016)
 *   SPMVPushGraphCSR              SPMV.h:53      (SPMV.c:752-848)
 *   SPMVPushFixedPointGraphCSR    SPMV.h:56      (SPMV.c:1017-1136)
 *
 * y = A^T-pull of x over the inverse CSR: y[v] = sum over in-edges (u->v) of w(u,v)*x[u].
 * When the graph carries no weights every edge weighs 0.0001f, the reference's
 * unweighted constant (SPMV.c:1206).
 */
#ifndef SPMV_H
#define SPMV_H

#include <stdint.h>

#include "graphConfig.h"
#include "graphCSR.h"

#ifdef __cplusplus
extern "C" {
#endif

struct SPMVStats
{
    uint32_t iterations;
    struct Arguments *arguments;
    uint32_t num_vertices;
    float *vector_output;
    float *vector_input;
    double time_total;
};

struct SPMVStats *newSPMVStatsGraphCSR(struct GraphCSR *graph);
void freeSPMVStats(struct SPMVStats *stats);

struct SPMVStats *SPMVGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct SPMVStats *SPMVPullGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct SPMVStats *SPMVPullFixedPointGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct SPMVStats *SPMVPushGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct SPMVStats *SPMVPushFixedPointGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);

#ifdef __cplusplus
}
#endif

#endif
