/* A CPU fake of the device layer (include/accelgraph.h) for tests of the drop-in host code
 * without a GPU: the reference's entry points (host/pageRank.c, host/SPMV.c) and the graph
 * cache (host/accel_context.c) are linked against these functions instead of the HIP
 * library. It computes through the oracle (oracle/ag_oracle.c), so a run without faults has
 * the oracle's results, and it injects the failures the environment variable AG_FAKE_FAIL
 * names, the way the reference's fake libcxl forces open, MMIO, stall and device errors on
 * its host runtime (02_capi_graph tests, test_capienv_verification.c:164-236):
 *   open     ag_device_open        AG_ERR_NO_DEVICE
 *   upload   ag_graph_upload       AG_ERR_OOM
 *   begin    ag_pr_begin           AG_ERR_HIP
 *   device   ag_pr_step            AG_ERR_HIP        (a kernel failed mid-run)
 *   timeout  ag_pr_poll            AG_ERR_TIMEOUT    (the bounded wait expired)
 *   history  ag_pr_history         AG_ERR_HIP
 *   end      ag_pr_end             AG_ERR_HIP
 *   spmv     ag_spmv_pull          AG_ERR_TIMEOUT
 * Test infrastructure only: never built into the library. */
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "accelgraph.h"

/* the oracle's restatements (oracle/ag_oracle.c) */
typedef struct {
    uint32_t iterations, executed, active;
    double error_total, rank_sum;
} or_pr_out;
int or_pagerank_pull(uint32_t n, const uint32_t *outdeg, const uint32_t *idx,
                     const uint32_t *deg, const uint32_t *src, uint32_t kind, uint32_t max_iter,
                     double eps, uint32_t stop_rule, double l1_tol, uint32_t init,
                     const float *init_ranks, int threads, float *ranks_out, float *scaled_out,
                     uint32_t *hist_active, double *hist_err, or_pr_out *out);
void or_spmv_f32(uint32_t n, const uint32_t *idx, const uint32_t *deg, const uint32_t *src,
                 const float *w, const float *x, float *y, int threads);

struct ag_device {
    int live;
};
struct ag_dgraph {
    ag_device *dev;
    uint32_t n, m;
    uint32_t *outdeg, *idx, *deg, *src;
    float *w;
};
struct ag_pr_session {
    ag_dgraph *g;
    ag_pr_options o;
    uint32_t enqueued;
    float *ranks;
    uint32_t *hist_active;
    double *hist_err;
    or_pr_out out;
};

static char g_msg[256];

static int fail_is(const char *what)
{
    const char *e = getenv("AG_FAKE_FAIL");
    return e && strcmp(e, what) == 0;
}

static int set_error(int st, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
    va_end(ap);
    return st;
}

const char *ag_last_error(void) { return g_msg; }

const char *ag_status_string(int status)
{
    switch (status) {
    case AG_OK: return "AG_OK";
    case AG_ERR_INVALID: return "AG_ERR_INVALID";
    case AG_ERR_NO_DEVICE: return "AG_ERR_NO_DEVICE";
    case AG_ERR_HIP: return "AG_ERR_HIP";
    case AG_ERR_OOM: return "AG_ERR_OOM";
    case AG_ERR_TIMEOUT: return "AG_ERR_TIMEOUT";
    case AG_ERR_UNSUPPORTED: return "AG_ERR_UNSUPPORTED";
    default: return "unknown status";
    }
}

int ag_device_open(int ordinal, ag_device **out)
{
    (void)ordinal;
    if (fail_is("open"))
        return set_error(AG_ERR_NO_DEVICE, "fake: no usable device");
    *out = (ag_device *)calloc(1, sizeof(ag_device));
    return *out ? AG_OK : set_error(AG_ERR_OOM, "fake: host allocation");
}

int ag_device_close(ag_device *dev)
{
    if (dev && dev->live)
        return set_error(AG_ERR_INVALID, "fake: graphs still alive");
    free(dev);
    return AG_OK;
}

static void *dup_array(const void *p, size_t bytes)
{
    void *q = malloc(bytes ? bytes : 1);
    if (q && bytes)
        memcpy(q, p, bytes);
    return q;
}

int ag_graph_upload(ag_device *dev, uint32_t n, uint32_t m, const uint32_t *out_degree,
                    const uint32_t *inv_edges_idx, const uint32_t *inv_degree,
                    const uint32_t *inv_src, const float *inv_weight,
                    const ag_plan_params *params, ag_dgraph **out)
{
    (void)params;
    if (!dev || !out)
        return set_error(AG_ERR_INVALID, "fake: null argument");
    if (fail_is("upload"))
        return set_error(AG_ERR_OOM, "fake: device allocation failed");
    ag_dgraph *g = (ag_dgraph *)calloc(1, sizeof(ag_dgraph));
    if (!g)
        return set_error(AG_ERR_OOM, "fake: host allocation");
    g->dev = dev;
    g->n = n;
    g->m = m;
    g->outdeg = (uint32_t *)dup_array(out_degree, (size_t)n * 4);
    g->idx = (uint32_t *)dup_array(inv_edges_idx, (size_t)n * 4);
    g->deg = (uint32_t *)dup_array(inv_degree, (size_t)n * 4);
    g->src = (uint32_t *)dup_array(inv_src, (size_t)m * 4);
    g->w = inv_weight ? (float *)dup_array(inv_weight, (size_t)m * 4) : NULL;
    dev->live++;
    *out = g;
    return AG_OK;
}

int ag_graph_free(ag_dgraph *g)
{
    if (!g)
        return AG_OK;
    g->dev->live--;
    free(g->outdeg);
    free(g->idx);
    free(g->deg);
    free(g->src);
    free(g->w);
    free(g);
    return AG_OK;
}

void ag_pr_options_default(ag_pr_options *o)
{
    memset(o, 0, sizeof(*o));
    o->kind = AG_KIND_F32;
    o->max_iterations = 20;
    o->epsilon = 1e-4;
    o->stop_rule = AG_STOP_ACTIVE;
    o->l1_tolerance = 1e-6;
    o->init = AG_INIT_REFERENCE;
    o->storage = AG_STORE_F32;
    o->direction = AG_DIR_PULL;
}

int ag_pr_begin(ag_dgraph *g, const ag_pr_options *opts, ag_pr_session **out)
{
    if (!g || !opts || !out)
        return set_error(AG_ERR_INVALID, "fake: null argument");
    if (fail_is("begin"))
        return set_error(AG_ERR_HIP, "fake: hipMalloc failed");
    if (opts->direction != AG_DIR_PULL || opts->storage != AG_STORE_F32)
        return set_error(AG_ERR_UNSUPPORTED, "fake: pull with f32 storage only");
    ag_pr_session *s = (ag_pr_session *)calloc(1, sizeof(ag_pr_session));
    const uint32_t cap = opts->max_iterations ? opts->max_iterations : 1;
    if (s) {
        s->ranks = (float *)malloc((size_t)(g->n ? g->n : 1) * sizeof(float));
        s->hist_active = (uint32_t *)calloc(cap, sizeof(uint32_t));
        s->hist_err = (double *)calloc(cap, sizeof(double));
    }
    if (!s || !s->ranks || !s->hist_active || !s->hist_err) {
        if (s) {
            free(s->ranks);
            free(s->hist_active);
            free(s->hist_err);
        }
        free(s);
        return set_error(AG_ERR_OOM, "fake: host allocation");
    }
    s->g = g;
    s->o = *opts;
    if (or_pagerank_pull(g->n, g->outdeg, g->idx, g->deg, g->src, opts->kind,
                         opts->max_iterations, opts->epsilon, opts->stop_rule,
                         opts->l1_tolerance, opts->init, opts->init_ranks, 0, s->ranks, NULL,
                         s->hist_active, s->hist_err, &s->out) != 0) {
        free(s->ranks);
        free(s->hist_active);
        free(s->hist_err);
        free(s);
        return set_error(AG_ERR_UNSUPPORTED, "fake: kind %u", opts->kind);
    }
    *out = s;
    return AG_OK;
}

/* the solve ran in ag_pr_begin; the session reveals it as iterations are enqueued */
static void fill_result(const ag_pr_session *s, ag_pr_result *r)
{
    memset(r, 0, sizeof(*r));
    const uint32_t shown = s->enqueued < s->out.executed ? s->enqueued : s->out.executed;
    r->executed = shown;
    r->iterations = shown == s->out.executed ? s->out.iterations : shown;
    r->active_vertices = shown ? s->hist_active[shown - 1] : 0;
    r->error_total = shown ? s->hist_err[shown - 1] : 0.0;
    r->rank_sum = s->out.rank_sum;
    r->refine_iteration = -1;
}

int ag_pr_step(ag_pr_session *s, uint32_t iterations)
{
    if (!s)
        return set_error(AG_ERR_INVALID, "fake: null session");
    if (fail_is("device"))
        return set_error(AG_ERR_HIP, "fake: kernel launch failed");
    s->enqueued += iterations;
    return AG_OK;
}

int ag_pr_poll(ag_pr_session *s, ag_pr_result *res)
{
    if (!s || !res)
        return set_error(AG_ERR_INVALID, "fake: null argument");
    if (fail_is("timeout"))
        return set_error(AG_ERR_TIMEOUT, "fake: device wait exceeded its deadline");
    fill_result(s, res);
    return AG_OK;
}

int ag_pr_history(ag_pr_session *s, uint32_t cap, uint32_t *active, double *error,
                  float *iter_ms, uint32_t *rows)
{
    if (!s || !rows)
        return set_error(AG_ERR_INVALID, "fake: null argument");
    if (fail_is("history"))
        return set_error(AG_ERR_HIP, "fake: hipMemcpy failed");
    uint32_t k = s->out.executed < cap ? s->out.executed : cap;
    for (uint32_t i = 0; i < k; i++) {
        if (active)
            active[i] = s->hist_active[i];
        if (error)
            error[i] = s->hist_err[i];
        if (iter_ms)
            iter_ms[i] = 0.0f;
    }
    *rows = k;
    return AG_OK;
}

int ag_pr_end(ag_pr_session *s, float *ranks_out, ag_pr_result *res)
{
    if (!s)
        return set_error(AG_ERR_INVALID, "fake: null session");
    int st = AG_OK;
    if (fail_is("end")) {
        st = set_error(AG_ERR_HIP, "fake: hipMemcpy failed");
    } else {
        if (ranks_out)
            memcpy(ranks_out, s->ranks, (size_t)s->g->n * sizeof(float));
        if (res) {
            s->enqueued = s->out.executed;
            fill_result(s, res);
        }
    }
    free(s->ranks);
    free(s->hist_active);
    free(s->hist_err);
    free(s);
    return st;
}

int ag_spmv_pull(ag_dgraph *g, const float *x, float *y, uint32_t repetitions,
                 double *avg_kernel_ms)
{
    if (!g || !x || !y)
        return set_error(AG_ERR_INVALID, "fake: null argument");
    if (fail_is("spmv"))
        return set_error(AG_ERR_TIMEOUT, "fake: device wait exceeded its deadline");
    for (uint32_t r = 0; r < repetitions; r++)
        or_spmv_f32(g->n, g->idx, g->deg, g->src, g->w, x, y, 0);
    if (avg_kernel_ms)
        *avg_kernel_ms = 0.0;
    return AG_OK;
}

int ag_spmv_pull_fixed(ag_dgraph *g, const uint32_t *x_q, uint64_t *y_q, uint32_t repetitions,
                       double *avg_kernel_ms)
{
    (void)g; (void)x_q; (void)y_q; (void)repetitions; (void)avg_kernel_ms;
    return set_error(AG_ERR_UNSUPPORTED, "fake: fixed-point SpMV not provided");
}

int ag_spmv_push(ag_dgraph *g, const float *x, float *y, uint32_t repetitions,
                 double *avg_kernel_ms)
{
    (void)g; (void)x; (void)y; (void)repetitions; (void)avg_kernel_ms;
    return set_error(AG_ERR_UNSUPPORTED, "fake: push not provided");
}

int ag_spmv_push_fixed(ag_dgraph *g, const uint64_t *x_q, uint64_t *y_q, uint32_t repetitions,
                       double *avg_kernel_ms)
{
    (void)g; (void)x_q; (void)y_q; (void)repetitions; (void)avg_kernel_ms;
    return set_error(AG_ERR_UNSUPPORTED, "fake: push not provided");
}
