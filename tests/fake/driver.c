/* Drop-in caller for the fake-device tests: FP32 and fixed-32 PageRank pulls and the FP32
 * SpMV pull through the reference's entry points on an RMAT-10 graph. Prints the iteration
 * counts and every normalised rank / output bit pattern for the test to check. */
#include <stdio.h>
#include <string.h>

#include "SPMV.h"
#include "accelgraph_compat.h"
#include "graphCSR.h"
#include "pageRank.h"

static void run_pr(struct Arguments *a, struct GraphCSR *g, int code)
{
    a->pushpull = code;
    a->iterations = 100;
    struct PageRankStats *st = pageRankGraphCSR(a, g);
    printf("PR %d %u %u\n", code, g->num_vertices, st->iterations);
    for (uint32_t v = 0; v < g->num_vertices; v++) {
        uint32_t bits;
        memcpy(&bits, &st->pageRanks[v], 4);
        printf("R%d %08x\n", code, bits);
    }
    freePageRankStats(st);
}

int main(void)
{
    struct GraphCSR *g = graphCSRGenerateRMAT(10, 16, 0.57, 0.19, 0.19, 27491095u, 1);
    if (!g)
        return 2;
    struct Arguments a;
    memset(&a, 0, sizeof(a));
    a.epsilon = 1e-4;
    a.ker_numThreads = 1;
    a.afu_config = a.afu_config_2 = a.cu_config = a.cu_config_2 = 1;
    run_pr(&a, g, 0);   /* FP32 pull */
    run_pr(&a, g, 9);   /* fixed-point 32 pull */
    a.pushpull = 0;     /* FP32 SpMV pull, one repetition */
    a.iterations = 1;
    struct SPMVStats *sp = SPMVGraphCSR(&a, g);
    for (uint32_t v = 0; v < g->num_vertices; v++) {
        uint32_t bits;
        memcpy(&bits, &sp->vector_output[v], 4);
        printf("Y %08x\n", bits);
    }
    freeSPMVStats(sp);
    accelgraphShutdown();
    graphCSRFree(g);
    return 0;
}
