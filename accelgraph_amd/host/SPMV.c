/*
 * SPMV.c -- drop-in SpMV pull entry points on struct GraphCSR.
 *
 * Mirrors 02_capi_graph/src/algorithms/capi/SPMV.c:
 *   the input vector x[v] = 1/out_degree[v], or 0.001 for sinks (:694-701);
 *   `iterations` products y = A x, each overwriting y as the FPGA run does (:705-724);
 *   the checksum sum((int)(y*100 + .5) / 100.0) and the printed tables (:687-691,722,
 *   :726-739).
 * The fixed-point variant follows the FPGA datapath contract (x and w in Q16.16, per-edge
 * (u32)((x*w) >> 16), u64 row sums; cu_SPMV/CSR/PULL/FixedPoint/cu/cu_sum_kernel_control.sv:
 * 155-164) and converts the sums back with Fixed32ToFloat. (The reference host passes its
 * float arrays to the FPGA and reads back a buffer the device never writes, SPMV.c:899-901,
 * 979-983; the datapath contract is the defined behaviour.)
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "SPMV.h"
#include "fixedPoint.h"
#include "accelgraph.h"
#include "accel_context.h"

/* SPMV.c:51-73 */
struct SPMVStats *newSPMVStatsGraphCSR(struct GraphCSR *graph)
{
    struct SPMVStats *stats = (struct SPMVStats *)calloc(1, sizeof(struct SPMVStats));
    if (!stats)
        accelgraphFail("newSPMVStatsGraphCSR", AG_ERR_OOM);
    stats->iterations = 0;
    stats->num_vertices = graph->num_vertices;
    stats->time_total = 0.0;
    size_t n = graph->num_vertices ? graph->num_vertices : 1;
    stats->vector_output = (float *)calloc(n, sizeof(float));
    stats->vector_input = (float *)calloc(n, sizeof(float));
    if (!stats->vector_output || !stats->vector_input)
        accelgraphFail("newSPMVStatsGraphCSR", AG_ERR_OOM);
    return stats;
}

/* SPMV.c:145-158 */
void freeSPMVStats(struct SPMVStats *stats)
{
    if (stats) {
        free(stats->vector_output);
        free(stats->vector_input);
        free(stats);
    }
}

static double wall_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + (double)ts.tv_nsec * 1e-9;
}

static void reference_input(struct GraphCSR *graph, float *x)
{
    for (uint32_t v = 0; v < graph->num_vertices; v++) {
        if (graph->vertices->out_degree[v])
            x[v] = (1.0f / graph->vertices->out_degree[v]);
        else
            x[v] = 0.001f;
    }
}

static double checksum(const float *y, uint32_t n)
{
    double sum = 0.0;
    for (uint32_t v = 0; v < n; v++)
        sum += ((int)(y[v] * 100 + .5) / 100.0);
    return sum;
}

static void print_tail(const struct SPMVStats *stats, double sum)
{
    printf(" -----------------------------------------------------\n");
    printf("| %-15s | %-15s | %-15s | \n", "Iterations", "Sum", "Time (S)");
    printf(" -----------------------------------------------------\n");
    printf("| %-15u | %-15lf | %-15f | \n", stats->iterations, sum, stats->time_total);
    printf(" -----------------------------------------------------\n");
    fflush(stdout);
}

/* SPMV.c:622-751 */
struct SPMVStats *SPMVPullGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    struct SPMVStats *stats = newSPMVStatsGraphCSR(graph);
    stats->arguments = arguments;
    ag_dgraph *dg = accelgraphUploadGraph(graph, 1);

    printf(" -----------------------------------------------------\n");
    printf("| %-51s | \n", "Starting SPMV-PULL");
    printf(" -----------------------------------------------------\n");
    printf("| %-21s | %-27s | \n", "Iteration", "Time (S)");
    printf(" -----------------------------------------------------\n");

    reference_input(graph, stats->vector_input);
    double t0 = wall_s();
    double avg_ms = 0.0;
    uint32_t reps = arguments->iterations;
    if (reps) {
        int st = ag_spmv_pull(dg, stats->vector_input, stats->vector_output, reps, &avg_ms);
        if (st != AG_OK)
            accelgraphFail("ag_spmv_pull", st);
    }
    for (stats->iterations = 0; stats->iterations < reps; stats->iterations++)
        printf("| %-21u | %-27f | \n", stats->iterations, avg_ms * 1e-3);
    double sum = checksum(stats->vector_output, graph->num_vertices);
    stats->time_total = wall_s() - t0;
    print_tail(stats, sum);
    return stats;
}

/* SPMV.c:849-1016, following the FPGA fixed-point datapath contract */
struct SPMVStats *SPMVPullFixedPointGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    struct SPMVStats *stats = newSPMVStatsGraphCSR(graph);
    stats->arguments = arguments;
    ag_dgraph *dg = accelgraphUploadGraph(graph, 1);
    uint32_t n = graph->num_vertices;

    printf(" -----------------------------------------------------\n");
    printf("| %-51s | \n", "Starting SPMV-PULL Fixed-Point");
    printf(" -----------------------------------------------------\n");
    printf("| %-21s | %-27s | \n", "Iteration", "Time (S)");
    printf(" -----------------------------------------------------\n");

    reference_input(graph, stats->vector_input);
    uint32_t *xq = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    uint64_t *yq = (uint64_t *)calloc(n ? n : 1, sizeof(uint64_t));
    if (!xq || !yq)
        accelgraphFail("SPMVPullFixedPointGraphCSR buffers", AG_ERR_OOM);
    for (uint32_t v = 0; v < n; v++)
        xq[v] = FloatToFixed32(stats->vector_input[v]);
    double t0 = wall_s();
    double avg_ms = 0.0;
    uint32_t reps = arguments->iterations;
    if (reps) {
        int st = ag_spmv_pull_fixed(dg, xq, yq, reps, &avg_ms);
        if (st != AG_OK)
            accelgraphFail("ag_spmv_pull_fixed", st);
    }
    for (stats->iterations = 0; stats->iterations < reps; stats->iterations++)
        printf("| %-21u | %-27f | \n", stats->iterations, avg_ms * 1e-3);
    for (uint32_t v = 0; v < n; v++)
        stats->vector_output[v] = Fixed32ToFloat(yq[v]);
    double sum = checksum(stats->vector_output, n);
    stats->time_total = wall_s() - t0;
    free(xq);
    free(yq);
    print_tail(stats, sum);
    return stats;
}

/* SPMV.c:752-848: the output accumulates over the iterations (never cleared in between) */
struct SPMVStats *SPMVPushGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    struct SPMVStats *stats = newSPMVStatsGraphCSR(graph);
    stats->arguments = arguments;
    ag_dgraph *dg = accelgraphUploadGraph(graph, 1);

    printf(" -----------------------------------------------------\n");
    printf("| %-51s | \n", "Starting SPMV-PUSH");
    printf(" -----------------------------------------------------\n");
    printf("| %-21s | %-27s | \n", "Iteration", "Time (S)");
    printf(" -----------------------------------------------------\n");

    reference_input(graph, stats->vector_input);
    double t0 = wall_s();
    double avg_ms = 0.0;
    uint32_t reps = arguments->iterations;
    int st = ag_spmv_push(dg, stats->vector_input, stats->vector_output, reps, &avg_ms);
    if (st != AG_OK)
        accelgraphFail("ag_spmv_push", st);
    for (stats->iterations = 0; stats->iterations < reps; stats->iterations++)
        printf("| %-21u | %-27f | \n", stats->iterations, avg_ms * 1e-3);
    double sum = checksum(stats->vector_output, graph->num_vertices);
    stats->time_total = wall_s() - t0;
    print_tail(stats, sum);
    return stats;
}

/* SPMV.c:1017-1136: Q32.32 x and weights, MULFixed64V1 products, u64 sums accumulated over
 * the iterations, converted back with Fixed64ToDouble */
struct SPMVStats *SPMVPushFixedPointGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    struct SPMVStats *stats = newSPMVStatsGraphCSR(graph);
    stats->arguments = arguments;
    ag_dgraph *dg = accelgraphUploadGraph(graph, 1);
    uint32_t n = graph->num_vertices;

    printf(" -----------------------------------------------------\n");
    printf("| %-51s | \n", "Starting SPMV-PUSH Fixed-Point");
    printf(" -----------------------------------------------------\n");
    printf("| %-21s | %-27s | \n", "Iteration", "Time (S)");
    printf(" -----------------------------------------------------\n");

    reference_input(graph, stats->vector_input);
    uint64_t *xq = (uint64_t *)calloc(n ? n : 1, sizeof(uint64_t));
    uint64_t *yq = (uint64_t *)calloc(n ? n : 1, sizeof(uint64_t));
    if (!xq || !yq)
        accelgraphFail("SPMVPushFixedPointGraphCSR buffers", AG_ERR_OOM);
    for (uint32_t v = 0; v < n; v++)
        xq[v] = DoubleToFixed64(stats->vector_input[v]);
    double t0 = wall_s();
    double avg_ms = 0.0;
    uint32_t reps = arguments->iterations;
    int st = ag_spmv_push_fixed(dg, xq, yq, reps, &avg_ms);
    if (st != AG_OK)
        accelgraphFail("ag_spmv_push_fixed", st);
    for (stats->iterations = 0; stats->iterations < reps; stats->iterations++)
        printf("| %-21u | %-27f | \n", stats->iterations, avg_ms * 1e-3);
    for (uint32_t v = 0; v < n; v++)
        stats->vector_output[v] = (float)Fixed64ToDouble(yq[v]);
    double sum = checksum(stats->vector_output, n);
    stats->time_total = wall_s() - t0;
    free(xq);
    free(yq);
    print_tail(stats, sum);
    return stats;
}

/* SPMV.c:594-621 */
struct SPMVStats *SPMVGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    switch (arguments->pushpull) {
    case 0:
        return SPMVPullGraphCSR(arguments, graph);
    case 1:
        return SPMVPushGraphCSR(arguments, graph);
    case 2:
        return SPMVPullFixedPointGraphCSR(arguments, graph);
    case 3:
        return SPMVPushFixedPointGraphCSR(arguments, graph);
    default:
        return SPMVPullGraphCSR(arguments, graph);
    }
}
