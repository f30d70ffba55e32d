/*
 * pageRank.c -- drop-in PageRank pull entry points on struct GraphCSR.
 *
 * Each entry point keeps the reference's contract (02_capi_graph/src/algorithms/capi/
 * pageRank.c): the stats record, the printed table, the reference stop rule (no vertex with
 * |delta| >= epsilon), `iterations` as the index of the stopping iteration, ranks divided by
 * N at the end and error_total as the last iteration's L1 delta. The gather, the damping
 * update, the reductions and the next contribution run on the GPU (ag_pr_* in accelgraph.h);
 * this file only drives them and prints.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pageRank.h"
#include "accelgraph.h"
#include "accel_context.h"

/* pageRank.c:50-77 */
struct PageRankStats *newPageRankStatsGraphCSR(struct GraphCSR *graph)
{
    struct PageRankStats *stats = (struct PageRankStats *)calloc(1, sizeof(struct PageRankStats));
    if (!stats)
        accelgraphFail("newPageRankStatsGraphCSR", AG_ERR_OOM);
    stats->damp = Damp;
    stats->base_pr = (1.0f - stats->damp);
    stats->iterations = 0;
    stats->num_vertices = graph->num_vertices;
    stats->time_total = 0.0;
    stats->error_total = 0.0;
    size_t n = graph->num_vertices ? graph->num_vertices : 1;
    stats->realRanks = (uint32_t *)malloc(n * sizeof(uint32_t));
    stats->pageRanks = (float *)malloc(n * sizeof(float));
    if (!stats->realRanks || !stats->pageRanks)
        accelgraphFail("newPageRankStatsGraphCSR", AG_ERR_OOM);
    for (uint32_t v = 0; v < stats->num_vertices; v++) {
        stats->pageRanks[v] = stats->base_pr;
        stats->realRanks[v] = v;
    }
    return stats;
}

/* pageRank.c:170-180 */
void freePageRankStats(struct PageRankStats *stats)
{
    if (stats) {
        free(stats->realRanks);
        free(stats->pageRanks);
        free(stats);
    }
}

static const char *variant_title(uint32_t kind, uint32_t direction)
{
    if (direction == AG_DIR_PUSH)
        return kind == AG_KIND_F32 ? "Starting Page Rank Push (tolerance/epsilon)"
                                   : "Starting Page Rank Push FP (tolerance/epsilon)";
    if (direction == AG_DIR_DATA_DRIVEN)
        return "Starting Page Rank Pull DD (tolerance/epsilon)";
    switch (kind) {
    case AG_KIND_FIXED64: return "Starting Page Rank Pull FP_64 (tolerance/epsilon)";
    case AG_KIND_FIXED32: return "Starting Page Rank Pull FP_32 (tolerance/epsilon)";
    case AG_KIND_FIXED16: return "Starting Page Rank Pull FP_16 (tolerance/epsilon)";
    case AG_KIND_FIXED8: return "Starting Page Rank Pull FP_8 (tolerance/epsilon)";
    case AG_KIND_QUANT32: return "Starting Page Rank Pull Quant_32 (tolerance/epsilon)";
    case AG_KIND_QUANT16: return "Starting Page Rank Pull Quant_16 (tolerance/epsilon)";
    case AG_KIND_QUANT8: return "Starting Page Rank Pull Quant_8 (tolerance/epsilon)";
    default: return "Starting Page Rank Pull (tolerance/epsilon)";
    }
}

/* Shared body of the FP32, fixed-point and quantised pulls, the pushes and the data-driven
 * pull (pageRank.c:942-1109, 1111-1272, 1276-2117, 2119-2676, 2810-2969). */
static struct PageRankStats *pagerank_run(struct Arguments *arguments, struct GraphCSR *graph,
                                          uint32_t kind, uint32_t direction)
{
    struct PageRankStats *stats = newPageRankStatsGraphCSR(graph);

    printf(" -----------------------------------------------------\n");
    printf("| %-51s | \n", "               ---->>> GPU <<<----");
    printf(" -----------------------------------------------------\n");
    printf("| %-51s | \n", variant_title(kind, direction));
    printf(" -----------------------------------------------------\n");
    printf("| %-51.13lf | \n", arguments->epsilon);
    printf(" -----------------------------------------------------\n");
    printf("| %-10s | %-8s | %-15s | %-9s | \n", "Iteration", "Active", "Error", "Time (S)");
    printf(" -----------------------------------------------------\n");

    ag_dgraph *dg = accelgraphUploadGraph(graph, 0);
    ag_pr_options o;
    ag_pr_options_default(&o);
    o.kind = kind;
    o.direction = direction;
    o.max_iterations = arguments->iterations;
    o.epsilon = arguments->epsilon;
    o.stop_rule = AG_STOP_ACTIVE;
    o.init = AG_INIT_REFERENCE;
    o.record_history = 1;

    ag_pr_session *s = NULL;
    int st = ag_pr_begin(dg, &o, &s);
    if (st != AG_OK)
        accelgraphFail("ag_pr_begin", st);
    ag_pr_result r;
    memset(&r, 0, sizeof(r));
    uint32_t enq = 0;
    const uint32_t chunk = 8;
    while (enq < o.max_iterations) {
        uint32_t k = o.max_iterations - enq < chunk ? o.max_iterations - enq : chunk;
        st = ag_pr_step(s, k);
        if (st == AG_OK)
            st = ag_pr_poll(s, &r);
        if (st != AG_OK)
            accelgraphFail("ag_pr_step", st);
        enq += k;
        if (r.executed > r.iterations)
            break; /* the stop rule fired (iterations = index of the stopping iteration) */
    }
    uint32_t rows = 0, cap = r.executed;
    uint32_t *act = (uint32_t *)calloc(cap ? cap : 1, sizeof(uint32_t));
    double *err = (double *)calloc(cap ? cap : 1, sizeof(double));
    float *ms = (float *)calloc(cap ? cap : 1, sizeof(float));
    if (!act || !err || !ms)
        accelgraphFail("history buffers", AG_ERR_OOM);
    st = ag_pr_history(s, cap, act, err, ms, &rows);
    if (st != AG_OK)
        accelgraphFail("ag_pr_history", st);
    for (uint32_t i = 0; i < rows; i++)
        printf("| %-10u | %-8u | %-15.13lf | %-9f | \n", i, act[i], err[i], ms[i] * 1e-3);
    free(act);
    free(err);
    free(ms);

    st = ag_pr_end(s, stats->pageRanks, &r);
    if (st != AG_OK)
        accelgraphFail("ag_pr_end", st);

    stats->iterations = r.iterations;
    stats->activeVertices = r.active_vertices;
    stats->error_total = r.error_total;
    stats->time_total = r.total_ms * 1e-3;

    printf(" -----------------------------------------------------\n");
    printf("| %-10s | %-8s | %-15s | %-9s | \n", "Iterations", "PR Sum", "Error", "Time (S)");
    printf(" -----------------------------------------------------\n");
    printf("| %-10u | %-8lf | %-15.13lf | %-9f | \n", stats->iterations, r.rank_sum,
           stats->error_total, stats->time_total);
    printf(" -----------------------------------------------------\n");
    fflush(stdout);
    return stats;
}

/* pageRank.c:942-1109 */
struct PageRankStats *pageRankPullGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    return pagerank_run(arguments, graph, AG_KIND_F32, AG_DIR_PULL);
}

/* pageRank.c:1276-1453 */
struct PageRankStats *pageRankPullFixedPoint64BitGraphCSR(struct Arguments *arguments,
                                                          struct GraphCSR *graph)
{
    return pagerank_run(arguments, graph, AG_KIND_FIXED64, AG_DIR_PULL);
}

/* pageRank.c:1455-1633 */
struct PageRankStats *pageRankPullFixedPoint32BitGraphCSR(struct Arguments *arguments,
                                                          struct GraphCSR *graph)
{
    return pagerank_run(arguments, graph, AG_KIND_FIXED32, AG_DIR_PULL);
}

/* pageRank.c:1635-1811 */
struct PageRankStats *pageRankPullFixedPoint16BitGraphCSR(struct Arguments *arguments,
                                                          struct GraphCSR *graph)
{
    return pagerank_run(arguments, graph, AG_KIND_FIXED16, AG_DIR_PULL);
}

/* pageRank.c:1813-1990 */
struct PageRankStats *pageRankPullFixedPoint8BitGraphCSR(struct Arguments *arguments,
                                                         struct GraphCSR *graph)
{
    return pagerank_run(arguments, graph, AG_KIND_FIXED8, AG_DIR_PULL);
}

/* pageRank.c:2119-2303 */
struct PageRankStats *pageRankPullQuant32BitGraphCSR(struct Arguments *arguments,
                                                     struct GraphCSR *graph)
{
    return pagerank_run(arguments, graph, AG_KIND_QUANT32, AG_DIR_PULL);
}

/* pageRank.c:2305-2489 */
struct PageRankStats *pageRankPullQuant16BitGraphCSR(struct Arguments *arguments,
                                                     struct GraphCSR *graph)
{
    return pagerank_run(arguments, graph, AG_KIND_QUANT16, AG_DIR_PULL);
}

/* pageRank.c:2491-2676 */
struct PageRankStats *pageRankPullQuant8BitGraphCSR(struct Arguments *arguments,
                                                    struct GraphCSR *graph)
{
    return pagerank_run(arguments, graph, AG_KIND_QUANT8, AG_DIR_PULL);
}

/* pageRank.c:1111-1272 */
struct PageRankStats *pageRankPushGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    return pagerank_run(arguments, graph, AG_KIND_F32, AG_DIR_PUSH);
}

/* pageRank.c:1992-2117 (64-bit fixed-point contributions and sums, as the pull at :1276) */
struct PageRankStats *pageRankPushFixedPointGraphCSR(struct Arguments *arguments,
                                                     struct GraphCSR *graph)
{
    return pagerank_run(arguments, graph, AG_KIND_FIXED64, AG_DIR_PUSH);
}

/* pageRank.c:2810-2969 */
struct PageRankStats *pageRankDataDrivenPullGraphCSR(struct Arguments *arguments,
                                                     struct GraphCSR *graph)
{
    return pagerank_run(arguments, graph, AG_KIND_F32, AG_DIR_DATA_DRIVEN);
}

/* pageRank.c:874-939. Not provided: the quantised push (5), whose zero-point quantiser API
 * (getMinMax, GetScale, GetZeroPoint, quantize; pageRank.c:2725-2733) is absent with
 * OpenGraph, and the data-driven push and pull-push (7, 8), which read and add to shared
 * residuals without ordering and decide activity from the value read (pageRank.c:3068-3078,
 * 3242-3256), so their results depend on the thread schedule and have no fixed answer to
 * reproduce. */
struct PageRankStats *pageRankGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    switch (arguments->pushpull) {
    case 0:
        return pageRankPullGraphCSR(arguments, graph);
    case 2:
        return pageRankPullFixedPoint64BitGraphCSR(arguments, graph);
    case 9:
        return pageRankPullFixedPoint32BitGraphCSR(arguments, graph);
    case 10:
        return pageRankPullFixedPoint16BitGraphCSR(arguments, graph);
    case 11:
        return pageRankPullFixedPoint8BitGraphCSR(arguments, graph);
    case 1:
        return pageRankPushGraphCSR(arguments, graph);
    case 3:
        return pageRankPushFixedPointGraphCSR(arguments, graph);
    case 6:
        return pageRankDataDrivenPullGraphCSR(arguments, graph);
    case 4:
        return pageRankPullQuant32BitGraphCSR(arguments, graph);
    case 12:
        return pageRankPullQuant16BitGraphCSR(arguments, graph);
    case 13:
        return pageRankPullQuant8BitGraphCSR(arguments, graph);
    case 5: case 7: case 8:
        fprintf(stderr, "[accelgraph] pageRankGraphCSR: pushpull=%u (%s) is not provided by "
                        "this library: %s\n", arguments->pushpull,
                arguments->pushpull == 5 ? "quantised push"
                : arguments->pushpull == 7 ? "data-driven push" : "data-driven pull-push",
                arguments->pushpull == 5 ? "its quantiser is absent with OpenGraph"
                                         : "its result depends on the thread schedule");
        exit(EXIT_FAILURE);
    default:
        return pageRankPullGraphCSR(arguments, graph);
    }
}
