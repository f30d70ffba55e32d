// ag_batch.hip -- many independent small graphs solved by one persistent launch.
//
// Each workgroup takes the next graph from a list sorted by decreasing edge count (an atomic
// ticket per graph, 512 tickets for 512 graphs) and runs the whole PageRank-pull iteration of
// pageRank.c:942-1109 for it: light rows (<= AG_BATCH_HEAVY_DEG in-edges) are summed by one
// lane each in the reference's stream order, heavy rows by a wavefront; the damping update,
// the FP64 L1 / active reductions and the next contribution are fused as in the single-graph
// kernel; the workgroup barrier separates iterations. The launch (a tiny ticket-reset kernel
// plus the solver) is captured once into a hipGraph and replayed; options live in a device
// buffer written before each replay, so the graph never needs re-capturing.
#include <stdlib.h>
#include <string.h>

#include "ag_internal.h"
#include "ag_pull_kernel.h"

// 1024 lanes (16 wavefronts) per graph: the largest graphs of a batch set the makespan, and
// their iterations are latency-bound chains of gathers that more wavefronts overlap (512 lanes
// measured 1.9x slower on 512 graphs of scale 10-16)
#define AG_BATCH_WG 1024

struct BatchGraph {
    uint32_t v0, n, e0, m, ob, h0, nh, pad;
};

struct BatchCfg {
    uint32_t max_iter, stop_rule, init, num_graphs;
    double eps, l1_tol;
};

struct BatchArgs {
    const BatchGraph *graphs;
    const uint32_t *order;
    const uint32_t *off;      // per graph: n+1 local offsets, graph g starts at ob
    const uint32_t *src;      // graph-local source ids
    const uint32_t *outdeg;
    const uint32_t *heavy;    // graph-local heavy row ids
    float *rank, *c0, *c1, *out;
    uint32_t *iters;
    double *err;
    uint32_t *ticket;
    const BatchCfg *cfg;
};

struct ag_batch {
    ag_device *dev;
    uint32_t G, total_n, total_m, total_h;
    BatchArgs a;
    void *allocs[16];
    int nalloc;
    BatchCfg *cfg_host;   // pinned
    hipGraph_t graph;
    hipGraphExec_t exec;
    int captured;
};

namespace {

__device__ __forceinline__ void batch_finish(uint32_t r, float sum, float *rank, float *cout,
                                             const uint32_t *od, float base, float damp,
                                             double dn, double eps, Red &red)
{
    const float nw = __fadd_rn(base, __fmul_rn(damp, sum));
    const float old = rank[r];
    rank[r] = nw;
    const double e = (double)fabsf(__fsub_rn(nw, old));
    red.err += __ddiv_rn(e, dn);
    red.active += (e >= eps) ? 1u : 0u;
    const uint32_t d = od[r];
    cout[r] = d ? __fdiv_rn(nw, __uint2float_rn(d)) : 0.0f;
}

__global__ void batch_reset_kernel(uint32_t *ticket) { *ticket = 0u; }

__global__ __launch_bounds__(AG_BATCH_WG) void batch_kernel(BatchArgs A)
{
    __shared__ uint32_t s_g;
    __shared__ uint32_t s_stop;
    __shared__ double s_e[AG_BATCH_WG / AG_WAVE];
    __shared__ uint32_t s_a[AG_BATCH_WG / AG_WAVE];
    const uint32_t tid = threadIdx.x;
    const BatchCfg cfg = *A.cfg;
    const float damp = 0.85f;
    const float base = __fsub_rn(1.0f, damp);
    for (;;) {
        if (tid == 0)
            s_g = atomicAdd(A.ticket, 1u);
        __syncthreads();
        const uint32_t k = s_g;
        __syncthreads();
        if (k >= cfg.num_graphs)
            return;
        const uint32_t gi = A.order[k];
        const BatchGraph gr = A.graphs[gi];
        const uint32_t n = gr.n;
        const uint32_t *off = A.off + gr.ob;
        const uint32_t *src = A.src + gr.e0;
        const uint32_t *od = A.outdeg + gr.v0;
        const uint32_t *heavy = A.heavy + gr.h0;
        float *rank = A.rank + gr.v0;
        float *c0 = A.c0 + gr.v0;
        float *c1 = A.c1 + gr.v0;
        const double dn = (double)n;
        for (uint32_t r = tid; r < n; r += AG_BATCH_WG) {
            const float x = cfg.init == AG_INIT_REFERENCE ? base : 1.0f;
            rank[r] = x;
            const uint32_t d = od[r];
            c0[r] = d ? __fdiv_rn(x, __uint2float_rn(d)) : 0.0f;
        }
        __syncthreads();
        uint32_t it = 0, stopped = 0, act = 0;
        double E = 0.0;
        for (it = 0; it < cfg.max_iter; it++) {
            const float *cin = (it & 1u) ? c1 : c0;
            float *cout = (it & 1u) ? c0 : c1;
            Red red{};
            for (uint32_t r = tid; r < n; r += AG_BATCH_WG) {
                const uint32_t b = off[r], e = off[r + 1];
                if (e - b > AG_BATCH_HEAVY_DEG)
                    continue;
                float s = 0.0f;
                if (e > b) {
                    s = cin[src[b]];
                    uint32_t j = b + 1;
                    for (; j + 3 < e; j += 4) {
                        const uint32_t u0 = src[j], u1 = src[j + 1], u2 = src[j + 2], u3 = src[j + 3];
                        const float x0 = cin[u0], x1 = cin[u1], x2 = cin[u2], x3 = cin[u3];
                        s = s + x0;
                        s = s + x1;
                        s = s + x2;
                        s = s + x3;
                    }
                    for (; j < e; j++)
                        s = s + cin[src[j]];
                }
                batch_finish(r, s, rank, cout, od, base, damp, dn, cfg.eps, red);
            }
            const uint32_t wave = tid / AG_WAVE, lane = tid % AG_WAVE;
            for (uint32_t h = wave; h < gr.nh; h += AG_BATCH_WG / AG_WAVE) {
                const uint32_t r = heavy[h];
                const uint32_t b = off[r], e = off[r + 1];
                float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
                uint32_t j = b + lane;
                for (; j + 3 * AG_WAVE < e; j += 4 * AG_WAVE) {
                    const uint32_t u0 = src[j], u1 = src[j + AG_WAVE], u2 = src[j + 2 * AG_WAVE],
                                   u3 = src[j + 3 * AG_WAVE];
                    a0 = a0 + cin[u0];
                    a1 = a1 + cin[u1];
                    a2 = a2 + cin[u2];
                    a3 = a3 + cin[u3];
                }
                for (; j < e; j += AG_WAVE)
                    a0 = a0 + cin[src[j]];
                float s = group_reduce<float, AG_WAVE>((a0 + a1) + (a2 + a3));
                if (lane == 0)
                    batch_finish(r, s, rank, cout, od, base, damp, dn, cfg.eps, red);
            }
            const Red t = block_reduce_red<AG_BATCH_WG>(red, s_e, s_a);
            if (tid == 0) {
                E = t.err;
                act = t.active;
                uint32_t stop = 0;
                if (cfg.stop_rule == AG_STOP_ACTIVE)
                    stop = (act == 0);
                else if (cfg.stop_rule == AG_STOP_L1)
                    stop = (E <= cfg.l1_tol);
                s_stop = stop;
            }
            __syncthreads();
            if (s_stop) {
                stopped = 1;
                break;
            }
        }
        const float nf = __uint2float_rn(n);
        for (uint32_t r = tid; r < n; r += AG_BATCH_WG)
            A.out[gr.v0 + r] = __fdiv_rn(rank[r], nf);
        if (tid == 0) {
            A.iters[gi] = stopped ? it : cfg.max_iter;
            A.err[gi] = E;
        }
        __syncthreads();
    }
}

} // namespace

extern "C" void ag_batch_options_default(ag_batch_options *o)
{
    memset(o, 0, sizeof(*o));
    o->max_iterations = 100;
    o->stop_rule = AG_STOP_L1;
    o->epsilon = 1e-4;
    o->l1_tolerance = 1e-6;
    o->init = AG_INIT_UNIFORM;
    o->use_hip_graph = 1;
}

extern "C" int ag_batch_free(ag_batch *b)
{
    if (!b)
        return AG_OK;
    (void)hipSetDevice(b->dev->ordinal);
    (void)hipStreamSynchronize(b->dev->stream);
    if (b->captured) {
        (void)hipGraphExecDestroy(b->exec);
        (void)hipGraphDestroy(b->graph);
    }
    for (int i = 0; i < b->nalloc; i++)
        (void)hipFree(b->allocs[i]);
    if (b->cfg_host)
        (void)hipHostFree(b->cfg_host);
    if (b->dev->live_graphs)
        b->dev->live_graphs--;
    free(b);
    return AG_OK;
}

template <typename T>
static int balloc(ag_batch *b, T **p, size_t count, const T *src)
{
    size_t bytes = (count ? count : 1) * sizeof(T);
    AG_HIP(hipMalloc((void **)p, bytes));
    b->allocs[b->nalloc++] = *p;
    if (src && count)
        AG_HIP(hipMemcpyAsync(*p, src, count * sizeof(T), hipMemcpyHostToDevice, b->dev->stream));
    return AG_OK;
}

extern "C" int ag_batch_upload(ag_device *dev, uint32_t G, const uint32_t *vert_off,
                               const uint32_t *edge_off, const uint32_t *out_degree,
                               const uint32_t *inv_edges_idx, const uint32_t *inv_degree,
                               const uint32_t *inv_src, ag_batch **out)
{
    if (!dev || !out || G == 0 || !vert_off || !edge_off)
        return ag_set_error(AG_ERR_INVALID, "null argument or empty batch");
    *out = NULL;
    const uint32_t N = vert_off[G], M = edge_off[G];
    if (vert_off[0] != 0 || edge_off[0] != 0)
        return ag_set_error(AG_ERR_INVALID, "offsets must start at 0");
    if ((N && (!out_degree || !inv_edges_idx || !inv_degree)) || (M && !inv_src))
        return ag_set_error(AG_ERR_INVALID, "null array");
    BatchGraph *hg = (BatchGraph *)calloc(G, sizeof(BatchGraph));
    uint32_t *hoff = (uint32_t *)malloc(((size_t)N + G) * sizeof(uint32_t));
    uint32_t *hsrc = (uint32_t *)malloc((size_t)(M ? M : 1) * sizeof(uint32_t));
    uint32_t *horder = (uint32_t *)malloc((size_t)G * sizeof(uint32_t));
    uint64_t *keys = (uint64_t *)malloc((size_t)G * sizeof(uint64_t));
    uint32_t *hheavy = (uint32_t *)malloc((size_t)(N ? N : 1) * sizeof(uint32_t));
    int st = AG_OK;
    if (!hg || !hoff || !hsrc || !horder || !keys || !hheavy)
        st = ag_set_error(AG_ERR_OOM, "host allocation failed");
    uint32_t nh_total = 0;
    for (uint32_t g = 0; g < G && st == AG_OK; g++) {
        const uint32_t v0 = vert_off[g], v1 = vert_off[g + 1], e0 = edge_off[g], e1 = edge_off[g + 1];
        if (v1 < v0 || e1 < e0) {
            st = ag_set_error(AG_ERR_INVALID, "offsets must be non-decreasing (graph %u)", g);
            break;
        }
        const uint32_t n = v1 - v0, m = e1 - e0;
        BatchGraph b;
        b.v0 = v0;
        b.n = n;
        b.e0 = e0;
        b.ob = v0 + g;
        b.h0 = nh_total;
        b.nh = 0;
        b.pad = 0;
        uint64_t acc = 0;
        for (uint32_t r = 0; r < n && st == AG_OK; r++) {
            const uint32_t idx = inv_edges_idx[v0 + r], d = inv_degree[v0 + r];
            if ((uint64_t)idx + d > m) {
                st = ag_set_error(AG_ERR_INVALID, "graph %u row %u runs past its edges", g, r);
                break;
            }
            hoff[b.ob + r] = (uint32_t)acc;
            for (uint32_t k = 0; k < d; k++) {
                const uint32_t s = inv_src[e0 + idx + k];
                if (s >= n) {
                    st = ag_set_error(AG_ERR_INVALID, "graph %u source id out of range", g);
                    break;
                }
                hsrc[e0 + acc + k] = s;
            }
            acc += d;
            if (d > AG_BATCH_HEAVY_DEG)
                hheavy[nh_total + b.nh++] = r;
        }
        if (st != AG_OK)
            break;
        if (acc > m) {
            st = ag_set_error(AG_ERR_INVALID, "graph %u degrees exceed its edge range", g);
            break;
        }
        hoff[b.ob + n] = (uint32_t)acc;
        b.m = (uint32_t)acc;
        nh_total += b.nh;
        hg[g] = b;
        keys[g] = ((uint64_t)(0xFFFFFFFFu - b.m) << 32) | g;  // largest first, then by index
    }
    if (st != AG_OK) {
        free(hg); free(hoff); free(hsrc); free(horder); free(keys); free(hheavy);
        return st;
    }
    qsort(keys, G, sizeof(uint64_t), [](const void *a, const void *b) {
        uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
        return (x > y) - (x < y);
    });
    for (uint32_t g = 0; g < G; g++)
        horder[g] = (uint32_t)keys[g];

    hipError_t he = hipSetDevice(dev->ordinal);
    ag_batch *b = (ag_batch *)calloc(1, sizeof(ag_batch));
    if (he != hipSuccess || !b) {
        free(hg); free(hoff); free(hsrc); free(horder); free(keys); free(hheavy); free(b);
        return he != hipSuccess ? ag_hip_fail(he, "hipSetDevice", __FILE__, __LINE__)
                                : ag_set_error(AG_ERR_OOM, "host allocation failed");
    }
    b->dev = dev;
    b->G = G;
    b->total_n = N;
    b->total_m = M;
    b->total_h = nh_total;
    BatchGraph *dgraphs = NULL;
    uint32_t *dorder = NULL, *doff = NULL, *dsrc = NULL, *dod = NULL, *dheavy = NULL, *dit = NULL,
             *dticket = NULL;
    float *drank = NULL, *dc0 = NULL, *dc1 = NULL, *dout = NULL;
    double *derr = NULL;
    BatchCfg *dcfg = NULL;
    if (st == AG_OK) st = balloc(b, &dgraphs, G, hg);
    if (st == AG_OK) st = balloc(b, &dorder, G, horder);
    if (st == AG_OK) st = balloc(b, &doff, (size_t)N + G, hoff);
    if (st == AG_OK) st = balloc(b, &dsrc, M, hsrc);
    if (st == AG_OK) st = balloc(b, &dod, N, out_degree);
    if (st == AG_OK) st = balloc(b, &dheavy, nh_total, hheavy);
    if (st == AG_OK) st = balloc<float>(b, &drank, N, NULL);
    if (st == AG_OK) st = balloc<float>(b, &dc0, N, NULL);
    if (st == AG_OK) st = balloc<float>(b, &dc1, N, NULL);
    if (st == AG_OK) st = balloc<float>(b, &dout, N, NULL);
    if (st == AG_OK) st = balloc<uint32_t>(b, &dit, G, NULL);
    if (st == AG_OK) st = balloc<double>(b, &derr, G, NULL);
    if (st == AG_OK) st = balloc<uint32_t>(b, &dticket, 1, NULL);
    if (st == AG_OK) st = balloc<BatchCfg>(b, &dcfg, 1, NULL);
    if (st == AG_OK) {
        he = hipHostMalloc((void **)&b->cfg_host, sizeof(BatchCfg), hipHostMallocDefault);
        if (he != hipSuccess)
            st = ag_hip_fail(he, "hipHostMalloc", __FILE__, __LINE__);
    }
    if (st == AG_OK)
        st = ag_wait_stream(dev);  // host staging buffers must outlive the copies
    free(hg); free(hoff); free(hsrc); free(horder); free(keys); free(hheavy);
    dev->live_graphs++;
    if (st != AG_OK) {
        ag_batch_free(b);
        return st;
    }
    b->a = BatchArgs{dgraphs, dorder, doff, dsrc, dod, dheavy, drank, dc0, dc1, dout,
                     dit, derr, dticket, dcfg};
    *out = b;
    return AG_OK;
}

static int batch_enqueue(ag_batch *b)
{
    hipStream_t s = b->dev->stream;
    static const uint32_t per_cu = ag_resident_per_cu(batch_kernel, AG_BATCH_WG);
    uint32_t grid = b->dev->num_cus * per_cu;
    if (grid > b->G)
        grid = b->G;
    batch_reset_kernel<<<1, 1, 0, s>>>(b->a.ticket);
    batch_kernel<<<grid, AG_BATCH_WG, 0, s>>>(b->a);
    AG_HIP(hipGetLastError());
    return AG_OK;
}

extern "C" int ag_batch_pagerank(ag_batch *b, const ag_batch_options *opts, float *ranks_out,
                                 uint32_t *iterations, double *error, double *device_ms)
{
    if (!b)
        return ag_set_error(AG_ERR_INVALID, "batch is null");
    ag_batch_options o;
    if (opts)
        o = *opts;
    else
        ag_batch_options_default(&o);
    if (o.stop_rule > AG_STOP_NONE || o.init > AG_INIT_UNIFORM)
        return ag_set_error(AG_ERR_INVALID, "unsupported stop rule or init mode");
    AG_HIP(hipSetDevice(b->dev->ordinal));
    hipStream_t s = b->dev->stream;
    // options go through a device buffer so a captured graph stays valid
    int st = ag_wait_stream(b->dev);  // the pinned staging buffer may still feed a copy
    if (st != AG_OK)
        return st;
    BatchCfg c;
    c.max_iter = o.max_iterations;
    c.stop_rule = o.stop_rule;
    c.init = o.init;
    c.num_graphs = b->G;
    c.eps = o.epsilon;
    c.l1_tol = o.l1_tolerance;
    *b->cfg_host = c;
    AG_HIP(hipMemcpyAsync((void *)b->a.cfg, b->cfg_host, sizeof(BatchCfg), hipMemcpyHostToDevice, s));
    hipEvent_t e0, e1;
    AG_HIP(hipEventCreate(&e0));
    AG_HIP(hipEventCreate(&e1));
    hipError_t he = hipEventRecord(e0, s);
    if (he == hipSuccess) {
        if (o.use_hip_graph) {
            if (!b->captured) {
                he = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
                if (he == hipSuccess) {
                    int est = batch_enqueue(b);
                    hipError_t ce = hipStreamEndCapture(s, &b->graph);
                    if (est != AG_OK) {
                        if (ce == hipSuccess)
                            (void)hipGraphDestroy(b->graph);
                        (void)hipEventDestroy(e0);
                        (void)hipEventDestroy(e1);
                        return est;
                    }
                    he = ce;
                }
                if (he == hipSuccess)
                    he = hipGraphInstantiate(&b->exec, b->graph, NULL, NULL, 0);
                if (he == hipSuccess)
                    b->captured = 1;
            }
            if (he == hipSuccess)
                he = hipGraphLaunch(b->exec, s);
        } else {
            int est = batch_enqueue(b);
            if (est != AG_OK) {
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                return est;
            }
        }
    }
    if (he == hipSuccess)
        he = hipEventRecord(e1, s);
    if (he == hipSuccess && ranks_out && b->total_n)
        he = hipMemcpyAsync(ranks_out, b->a.out, (size_t)b->total_n * sizeof(float),
                            hipMemcpyDeviceToHost, s);
    if (he == hipSuccess && iterations)
        he = hipMemcpyAsync(iterations, b->a.iters, (size_t)b->G * sizeof(uint32_t),
                            hipMemcpyDeviceToHost, s);
    if (he == hipSuccess && error)
        he = hipMemcpyAsync(error, b->a.err, (size_t)b->G * sizeof(double), hipMemcpyDeviceToHost, s);
    st = (he == hipSuccess) ? ag_wait_stream(b->dev) : ag_hip_fail(he, "batch launch", __FILE__, __LINE__);
    if (st == AG_OK && device_ms) {
        float ms = 0.0f;
        he = hipEventElapsedTime(&ms, e0, e1);
        if (he != hipSuccess)
            st = ag_hip_fail(he, "hipEventElapsedTime", __FILE__, __LINE__);
        *device_ms = ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return st;
}
