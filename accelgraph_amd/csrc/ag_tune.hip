// ag_tune.hip -- plan autotuner: times candidate pull plans on the graph they will run on.
#include <stdlib.h>
#include <string.h>

#include "ag_internal.h"

static int cmp_double(const void *a, const void *b)
{
    const double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

extern "C" int ag_autotune_plan(ag_device *dev, uint32_t n, uint32_t m, const uint32_t *out_degree,
                                const uint32_t *inv_edges_idx, const uint32_t *inv_degree,
                                const uint32_t *inv_src, const ag_plan_params *cand,
                                uint32_t count, uint32_t iterations, uint32_t repeats,
                                double *ms_out, uint32_t *best)
{
    if (!dev || !cand || count == 0 || !best || iterations == 0)
        return ag_set_error(AG_ERR_INVALID, "bad argument");
    if (repeats == 0)
        repeats = 1;
    double *samples = (double *)malloc(repeats * sizeof(double));
    if (!samples)
        return ag_set_error(AG_ERR_OOM, "host allocation failed");
    double best_ms = 0.0;
    *best = 0;
    int st = AG_OK;
    for (uint32_t c = 0; c < count && st == AG_OK; c++) {
        ag_dgraph *g = NULL;
        st = ag_graph_upload(dev, n, m, out_degree, inv_edges_idx, inv_degree, inv_src, NULL,
                             &cand[c], &g);
        if (st != AG_OK)
            break;
        ag_pr_options o;
        ag_pr_options_default(&o);
        o.max_iterations = iterations * (repeats + 1);
        o.stop_rule = AG_STOP_NONE;
        o.init = AG_INIT_UNIFORM;
        ag_pr_session *s = NULL;
        st = ag_pr_begin(g, &o, &s);
        if (st == AG_OK)
            st = ag_pr_step(s, iterations);  // warm-up
        for (uint32_t r = 0; r < repeats && st == AG_OK; r++) {
            st = ag_pr_step(s, iterations);
            uint32_t launches = 0;
            if (st == AG_OK)
                st = ag_pr_last_step_kernel_ms(s, &samples[r], &launches);
        }
        if (s) {
            int st2 = ag_pr_end(s, NULL, NULL);
            if (st == AG_OK)
                st = st2;
        }
        ag_graph_free(g);
        if (st != AG_OK)
            break;
        qsort(samples, repeats, sizeof(double), cmp_double);
        const double med = samples[repeats / 2];
        if (ms_out)
            ms_out[c] = med;
        if (c == 0 || med < best_ms) {
            best_ms = med;
            *best = c;
        }
    }
    free(samples);
    return st;
}
