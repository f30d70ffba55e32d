/* The planner under AddressSanitizer + UndefinedBehaviorSanitizer (host code only): every plan
 * shape the GPU parity tests use, on RMAT graphs of scale 8-14 and on tiny graphs. Exits
 * non-zero when a valid plan is refused or an invalid one accepted. */
#include <stdio.h>
#include <string.h>
#include "graphCSR.h"
#include "accelgraph.h"
#include "ag_plan.h"

static int failures;

static void run(const struct GraphCSR *g, const ag_plan_params *p, int expect_ok, const char *name)
{
    ag_plan plan;
    const char *err = "";
    const int st = ag_plan_build(g->num_vertices, g->num_edges, g->vertices->out_degree,
                                 g->inverse_vertices->edges_idx, g->inverse_vertices->out_degree,
                                 g->inverse_sorted_edges_array->edges_array_dest, NULL, p, &plan,
                                 &err);
    if ((st == 0) != expect_ok) {
        printf("%s (n=%u): status %d %s\n", name, g->num_vertices, st, err);
        failures++;
    }
    if (st == 0)
        ag_plan_free(&plan);
}

int main(void)
{
    for (uint32_t scale = 8; scale <= 14; scale += 3) {
        struct GraphCSR *g = graphCSRGenerateRMAT(scale, 16, 0.57, 0.19, 0.19, 27491095u, 1);
        ag_plan_params p, s, b;
        ag_plan_params_default(&p);
        run(g, &p, 1, "default");
        s = p;
        s.thread_max_deg = 2; s.group16_max_deg = 6; s.wave_max_deg = 20; s.block_max_deg = 60;
        s.tile_rows = 7; s.tile_nnz = 9; s.segment_nnz = 256; s.order = 0;
        run(g, &s, 1, "small");
        s.order = 1;
        run(g, &s, 1, "small-order1");
        b = p;
        b.src_block_bits = 10; b.blocked_min_bin = 1;
        run(g, &b, 1, "blocked");
        b.src_block_bits = 8; b.blocked_min_bin = 0;
        run(g, &b, scale <= 16, "blocked-all");
        s.order = 0; s.src_block_bits = 9; s.blocked_min_bin = 1;
        run(g, &s, 1, "blocked-small");
        b.src_block_bits = 30;
        run(g, &b, 1, "one block");
        b.tile_nnz = 5000;
        run(g, &b, 0, "tile_nnz above the limit");
        graphCSRFree(g);
    }
    uint32_t src[3] = {1, 1, 1}, dst[3] = {0, 0, 0};
    struct GraphCSR *t = graphCSRFromEdgeList(4, 3, src, dst, NULL, 0);
    ag_plan_params p;
    ag_plan_params_default(&p);
    run(t, &p, 1, "multi-edge star");
    p.src_block_bits = 8; p.blocked_min_bin = 0;
    run(t, &p, 1, "multi-edge star, blocked");
    graphCSRFree(t);
    t = graphCSRFromEdgeList(5, 0, src, dst, NULL, 0);
    ag_plan_params_default(&p);
    run(t, &p, 1, "no edges");
    graphCSRFree(t);
    printf("failures %d\n", failures);
    return failures ? 1 : 0;
}
