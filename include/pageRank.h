/*
 * pageRank.h -- PageRank pull entry points on struct GraphCSR (drop-in for the reference).
 *
 * Replaces the CSR part of 02_capi_graph/include/algorithms/capi/pageRank.h:
 *   #define Damp                       pageRank.h:13
 *   struct PageRankStats               pageRank.h:19-30   (same fields, same order)
 *   newPageRankStatsGraphCSR           pageRank.h:32      (pageRank.c:50-77)
 *   freePageRankStats                  pageRank.h:37      (pageRank.c:170-180)
 *   pageRankGraphCSR                   pageRank.h:69      (pageRank.c:874-939)
 *   pageRankPullGraphCSR               pageRank.h:70      (pageRank.c:942-1109)
 *   pageRankPullFixedPoint64BitGraphCSR pageRank.h:73     (pageRank.c:1276-1453)
 *   pageRankPullFixedPoint32BitGraphCSR pageRank.h:74     (pageRank.c:1455-1633)
 *   pageRankPullFixedPoint16BitGraphCSR pageRank.h:75     (pageRank.c:1635-1811)
 *   pageRankPullFixedPoint8BitGraphCSR  pageRank.h:76     (pageRank.c:1813-1990)
 *   pageRankPullQuant32BitGraphCSR      pageRank.h:79     (pageRank.c:2119-2303)
 *   pageRankPullQuant16BitGraphCSR      pageRank.h:80     (pageRank.c:2305-2489)
 *   pageRankPullQuant8BitGraphCSR       pageRank.h:81     (pageRank.c:2491-2676)
 *   pageRankPushGraphCSR                pageRank.h:71     (pageRank.c:1111-1272)
 *   pageRankPushFixedPointGraphCSR      pageRank.h:77     (pageRank.c:1992-2117)
 *   pageRankDataDrivenPullGraphCSR      pageRank.h:84     (pageRank.c:2810-2969)
 *
 * The gather and the per-vertex update run on the GPU through the accelgraph C-ABI
 * (accelgraph.h). As in the reference, a device failure prints one diagnostic line to
 * stderr and exits the process with EXIT_FAILURE; callers that need an error code use
 * the ag_* functions directly.
 */
#ifndef PAGERANK_H
#define PAGERANK_H

#include <stdint.h>

#include "graphConfig.h"
#include "graphCSR.h"

#ifdef __cplusplus
extern "C" {
#endif

#define Damp 0.85f

struct PageRankStats
{
    float damp;
    float base_pr;
    uint32_t iterations;
    uint32_t num_vertices;
    uint32_t activeVertices;
    uint32_t *realRanks;
    float *pageRanks;
    double time_total;
    double error_total;
};

struct PageRankStats *newPageRankStatsGraphCSR(struct GraphCSR *graph);
void freePageRankStats(struct PageRankStats *stats);

struct PageRankStats *pageRankGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct PageRankStats *pageRankPullGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct PageRankStats *pageRankPullFixedPoint64BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct PageRankStats *pageRankPullFixedPoint32BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct PageRankStats *pageRankPullFixedPoint16BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct PageRankStats *pageRankPullFixedPoint8BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct PageRankStats *pageRankPullQuant32BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct PageRankStats *pageRankPullQuant16BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct PageRankStats *pageRankPullQuant8BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct PageRankStats *pageRankPushGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct PageRankStats *pageRankPushFixedPointGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct PageRankStats *pageRankDataDrivenPullGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);

#ifdef __cplusplus
}
#endif

#endif
