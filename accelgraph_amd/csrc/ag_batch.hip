// ag_batch.hip -- many independent graphs solved together.
//
// Every graph is cut into parts: runs of consecutive rows with at most part_edges in-edges (a
// row with more is a part by itself) and at most AG_BATCH_PART_ROWS rows; the parts of the
// largest graphs come first. One iteration of pageRank.c:942-1109 for every graph still
// iterating is two launches:
//   batch_part_kernel   every part of every unstopped graph: light rows (<= heavy_deg
//                       in-edges) summed by one lane in the reference's stream order, heavy
//                       rows by a wavefront (four strided accumulators, fixed tree); the fused
//                       damping update and next contribution; the part's FP64 L1 and active
//                       partial (fixed-order workgroup reduction)
//   batch_fin_kernel    per graph: its parts' partials added in part order, the stop rule
//                       (active count or L1), the iteration count, the graph's done flag
// The kernel boundary is the only synchronisation between the parts of a graph. A stopped
// graph's parts are skipped; once every graph has stopped, both launches return at once. The
// solve -- state reset, init, max_iterations x (parts, finish), normalised output -- is
// captured once per iteration budget and part layout into a hipGraph and replayed as one
// launch; options live in a device buffer written before each replay (no memset nodes).
#include <stdlib.h>
#include <string.h>

#include "ag_internal.h"
#include "ag_pull_kernel.h"

#define AG_BATCH_PART_ROWS 2048u

struct BatchGraph {
    uint32_t v0, n, e0, m, ob, pad0, pad1, pad2;
};

struct BatchPart {
    uint32_t g;         // graph index
    uint32_t r0, r1;    // graph-local rows [r0, r1)
    uint32_t h0, h1;    // the part's heavy rows: heavy[h0 .. h1) (graph-local row ids)
};

struct BatchCfg {
    uint32_t max_iter, stop_rule, init, num_graphs;
    double eps, l1_tol;
};

struct BatchArgs {
    const BatchGraph *graphs;
    const BatchPart *parts;
    const uint32_t *gpart;    // [2G] first and end part of each graph
    const uint32_t *off;      // per graph: n+1 local offsets, graph g's start at ob
    const uint32_t *src;      // graph-local source ids
    const uint32_t *outdeg;
    const uint32_t *heavy;    // graph-local heavy row ids, graph by graph, ascending
    float *rank, *c0, *c1, *out;
    uint32_t *iters;
    double *err;
    uint32_t *done;           // [G]
    uint32_t *ndone;          // graphs stopped so far
    double *perr;             // [2 * nparts] part L1 partials by iteration parity
    uint32_t *pact;           // [2 * nparts]
    const BatchCfg *cfg;
    uint32_t nparts, heavy_deg, total_n;
};

struct ag_batch {
    ag_device *dev;
    uint32_t G, total_n, total_m;
    BatchArgs a;
    void *allocs[24];
    int nalloc;
    // part layout (rebuilt when part_edges / heavy_deg change)
    uint32_t part_edges, heavy_deg;
    BatchPart *d_parts;
    uint32_t *d_gpart, *d_heavy;
    double *d_perr;
    uint32_t *d_pact;
    // host copies kept for re-partitioning
    BatchGraph *h_graphs;
    uint32_t *h_off;
    uint32_t *h_order;        // graphs by decreasing edge count
    BatchCfg *cfg_host;       // pinned
    uint32_t *ndone_host;     // pinned
    hipGraph_t graph;
    hipGraphExec_t exec;
    uint32_t exec_maxit;      // iterations of the captured solve (0: none)
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

__global__ void batch_reset_kernel(BatchArgs A)
{
    const uint32_t G = A.cfg->num_graphs, maxit = A.cfg->max_iter;
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += gridDim.x * blockDim.x) {
        A.done[g] = 0u;
        A.iters[g] = maxit;
        A.err[g] = 0.0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        *A.ndone = 0u;
}

__global__ void batch_init_kernel(BatchArgs A)
{
    const float damp = 0.85f;
    const float base = __fsub_rn(1.0f, damp);
    const float x = A.cfg->init == AG_INIT_REFERENCE ? base : 1.0f;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < A.total_n;
         v += gridDim.x * blockDim.x) {
        A.rank[v] = x;
        const uint32_t d = A.outdeg[v];
        A.c0[v] = d ? __fdiv_rn(x, __uint2float_rn(d)) : 0.0f;
    }
}

__global__ __launch_bounds__(AG_WG) void batch_part_kernel(BatchArgs A, uint32_t it)
{
    __shared__ double s_e[AG_WG / AG_WAVE];
    __shared__ uint32_t s_a[AG_WG / AG_WAVE];
    const BatchCfg cfg = *A.cfg;
    if (*(volatile const uint32_t *)A.ndone >= cfg.num_graphs)
        return;
    const uint32_t tid = threadIdx.x, wave = tid / AG_WAVE, lane = tid % AG_WAVE;
    const float damp = 0.85f;
    const float base = __fsub_rn(1.0f, damp);
    for (uint32_t p = blockIdx.x; p < A.nparts; p += gridDim.x) {
        const BatchPart pt = A.parts[p];
        if (A.done[pt.g])        // the same for every lane: written only by earlier launches
            continue;
        const BatchGraph gr = A.graphs[pt.g];
        const uint32_t *off = A.off + gr.ob;
        const uint32_t *src = A.src + gr.e0;
        const uint32_t *od = A.outdeg + gr.v0;
        float *rank = A.rank + gr.v0;
        const float *cin = ((it & 1u) ? A.c1 : A.c0) + gr.v0;
        float *cout = ((it & 1u) ? A.c0 : A.c1) + gr.v0;
        const double dn = (double)gr.n;
        Red red{};
        for (uint32_t r = pt.r0 + tid; r < pt.r1; r += AG_WG) {
            const uint32_t b = off[r], e = off[r + 1];
            if (e - b > A.heavy_deg)
                continue;
            float s = 0.0f;
            if (e > b) {
                s = cin[src[b]];   // the first element starts the sum, as the FPGA's does
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
        for (uint32_t h = pt.h0 + wave; h < pt.h1; h += AG_WG / AG_WAVE) {
            const uint32_t r = A.heavy[h];
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
            const float s = group_reduce<float, AG_WAVE>((a0 + a1) + (a2 + a3));
            if (lane == 0)
                batch_finish(r, s, rank, cout, od, base, damp, dn, cfg.eps, red);
        }
        const Red t = block_reduce_red<AG_WG>(red, s_e, s_a);
        if (tid == 0) {
            A.perr[(size_t)(it & 1u) * A.nparts + p] = t.err;
            A.pact[(size_t)(it & 1u) * A.nparts + p] = t.active;
        }
    }
}

// One thread per graph: the graph's L1 delta and active count of iteration `it` (its parts'
// partials in part order), then the stop rule.
__global__ void batch_fin_kernel(BatchArgs A, uint32_t it)
{
    const BatchCfg cfg = *A.cfg;
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= cfg.num_graphs || A.done[g])
        return;
    double E = 0.0;
    uint32_t act = 0;
    const size_t pb = (size_t)(it & 1u) * A.nparts;
    const uint32_t p1 = A.gpart[2 * g + 1];
    uint32_t p = A.gpart[2 * g];
    // part order, eight loads in flight
    for (; p + 8u <= p1; p += 8u) {
        double e[8];
        uint32_t a[8];
#pragma unroll
        for (uint32_t u = 0; u < 8u; u++) {
            e[u] = A.perr[pb + p + u];
            a[u] = A.pact[pb + p + u];
        }
#pragma unroll
        for (uint32_t u = 0; u < 8u; u++) {
            E += e[u];
            act += a[u];
        }
    }
    for (; p < p1; p++) {
        E += A.perr[pb + p];
        act += A.pact[pb + p];
    }
    A.err[g] = E;
    const bool stop = (cfg.stop_rule == AG_STOP_ACTIVE && act == 0) ||
                      (cfg.stop_rule == AG_STOP_L1 && E <= cfg.l1_tol);
    if (stop) {
        A.done[g] = 1u;
        A.iters[g] = it;
        atomicAdd(A.ndone, 1u);
    }
}

// Normalised ranks (scaled / n of the vertex's graph); a workgroup per graph.
__global__ void batch_out_kernel(BatchArgs A)
{
    const uint32_t G = A.cfg->num_graphs;
    for (uint32_t g = blockIdx.x; g < G; g += gridDim.x) {
        const BatchGraph gr = A.graphs[g];
        const float nf = __uint2float_rn(gr.n);
        for (uint32_t r = threadIdx.x; r < gr.n; r += blockDim.x)
            A.out[gr.v0 + r] = __fdiv_rn(A.rank[gr.v0 + r], nf);
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
    o->part_edges = AG_BATCH_PART_EDGES;
    o->heavy_deg = AG_BATCH_HEAVY_DEG;
}

static void drop_capture(ag_batch *b)
{
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        (void)hipGraphDestroy(b->graph);
        b->exec = NULL;
        b->graph = NULL;
    }
    b->exec_maxit = 0;
}

static void free_layout(ag_batch *b)
{
    void *p[] = {b->d_parts, b->d_gpart, b->d_heavy, b->d_perr, b->d_pact};
    for (void *q : p)
        if (q)
            (void)hipFree(q);
    b->d_parts = NULL;
    b->d_gpart = b->d_heavy = b->d_pact = NULL;
    b->d_perr = NULL;
}

extern "C" int ag_batch_free(ag_batch *b)
{
    if (!b)
        return AG_OK;
    (void)hipSetDevice(b->dev->ordinal);
    (void)hipStreamSynchronize(b->dev->stream);
    drop_capture(b);
    free_layout(b);
    for (int i = 0; i < b->nalloc; i++)
        (void)hipFree(b->allocs[i]);
    if (b->cfg_host)
        (void)hipHostFree(b->cfg_host);
    if (b->ndone_host)
        (void)hipHostFree(b->ndone_host);
    free(b->h_graphs);
    free(b->h_off);
    free(b->h_order);
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

// Builds and uploads the parts, the per-graph part ranges and the heavy-row lists for
// (part_edges, heavy_deg), replacing any earlier layout (and the captured solve).
static int build_layout(ag_batch *b, uint32_t part_edges, uint32_t heavy_deg)
{
    const uint32_t G = b->G;
    // heavy rows (graph by graph, ascending) and parts (graphs by decreasing edge count)
    size_t nh = 0, np = 0;
    for (uint32_t g = 0; g < G; g++) {
        const BatchGraph &gr = b->h_graphs[g];
        const uint32_t *off = b->h_off + gr.ob;
        uint32_t r = 0;
        while (r < gr.n) {
            uint32_t rows = 0, e = 0;
            while (r < gr.n && rows < AG_BATCH_PART_ROWS) {
                const uint32_t d = off[r + 1] - off[r];
                if (rows > 0 && (uint64_t)e + d > part_edges)
                    break;
                e += d;
                rows++;
                r++;
            }
            np++;
        }
        for (uint32_t q = 0; q < gr.n; q++)
            nh += (off[q + 1] - off[q] > heavy_deg);
    }
    BatchPart *parts = (BatchPart *)malloc((np ? np : 1) * sizeof(BatchPart));
    uint32_t *gpart = (uint32_t *)malloc((size_t)G * 2 * sizeof(uint32_t));
    uint32_t *heavy = (uint32_t *)malloc((nh ? nh : 1) * sizeof(uint32_t));
    uint32_t *h0 = (uint32_t *)malloc((size_t)(G + 1) * sizeof(uint32_t));
    if (!parts || !gpart || !heavy || !h0) {
        free(parts); free(gpart); free(heavy); free(h0);
        return ag_set_error(AG_ERR_OOM, "host allocation failed");
    }
    size_t k = 0;
    for (uint32_t g = 0; g < G; g++) {
        const BatchGraph &gr = b->h_graphs[g];
        const uint32_t *off = b->h_off + gr.ob;
        h0[g] = (uint32_t)k;
        for (uint32_t q = 0; q < gr.n; q++)
            if (off[q + 1] - off[q] > heavy_deg)
                heavy[k++] = q;
    }
    h0[G] = (uint32_t)k;
    size_t p = 0;
    for (uint32_t i = 0; i < G; i++) {
        const uint32_t g = b->h_order[i];
        const BatchGraph &gr = b->h_graphs[g];
        const uint32_t *off = b->h_off + gr.ob;
        gpart[2 * g] = (uint32_t)p;
        uint32_t r = 0, h = h0[g];
        while (r < gr.n) {
            const uint32_t r0 = r, hs = h;
            uint32_t rows = 0, e = 0;
            while (r < gr.n && rows < AG_BATCH_PART_ROWS) {
                const uint32_t d = off[r + 1] - off[r];
                if (rows > 0 && (uint64_t)e + d > part_edges)
                    break;
                e += d;
                rows++;
                r++;
            }
            while (h < h0[g + 1] && heavy[h] < r)
                h++;
            parts[p++] = BatchPart{g, r0, r, hs, h};
        }
        gpart[2 * g + 1] = (uint32_t)p;
    }
    free(h0);
    drop_capture(b);
    free_layout(b);
    int st = AG_OK;
    hipStream_t s = b->dev->stream;
    hipError_t e = hipMalloc((void **)&b->d_parts, (np ? np : 1) * sizeof(BatchPart));
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_gpart, (size_t)G * 2 * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_heavy, (nh ? nh : 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_perr, (np ? np : 1) * 2 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_pact, (np ? np : 1) * 2 * sizeof(uint32_t));
    if (e == hipSuccess && np)
        e = hipMemcpyAsync(b->d_parts, parts, np * sizeof(BatchPart), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->d_gpart, gpart, (size_t)G * 2 * sizeof(uint32_t), hipMemcpyHostToDevice, s);
    if (e == hipSuccess && nh)
        e = hipMemcpyAsync(b->d_heavy, heavy, nh * sizeof(uint32_t), hipMemcpyHostToDevice, s);
    st = (e == hipSuccess) ? ag_wait_stream(b->dev)   // the host arrays must outlive the copies
                           : ag_hip_fail(e, "batch layout upload", __FILE__, __LINE__);
    free(parts);
    free(gpart);
    free(heavy);
    if (st != AG_OK) {
        free_layout(b);
        return st;
    }
    b->part_edges = part_edges;
    b->heavy_deg = heavy_deg;
    b->a.parts = b->d_parts;
    b->a.gpart = b->d_gpart;
    b->a.heavy = b->d_heavy;
    b->a.perr = b->d_perr;
    b->a.pact = b->d_pact;
    b->a.nparts = (uint32_t)np;
    b->a.heavy_deg = heavy_deg;
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
    if ((uint64_t)N + G > 0xFFFF0000ull)
        return ag_set_error(AG_ERR_INVALID, "too many vertices in one batch");
    BatchGraph *hg = (BatchGraph *)calloc(G, sizeof(BatchGraph));
    uint32_t *hoff = (uint32_t *)malloc(((size_t)N + G) * sizeof(uint32_t));
    uint32_t *hsrc = (uint32_t *)malloc((size_t)(M ? M : 1) * sizeof(uint32_t));
    uint32_t *horder = (uint32_t *)malloc((size_t)G * sizeof(uint32_t));
    uint64_t *keys = (uint64_t *)malloc((size_t)G * sizeof(uint64_t));
    int st = AG_OK;
    if (!hg || !hoff || !hsrc || !horder || !keys)
        st = ag_set_error(AG_ERR_OOM, "host allocation failed");
    for (uint32_t g = 0; g < G && st == AG_OK; g++) {
        const uint32_t v0 = vert_off[g], v1 = vert_off[g + 1], e0 = edge_off[g], e1 = edge_off[g + 1];
        if (v1 < v0 || e1 < e0) {
            st = ag_set_error(AG_ERR_INVALID, "offsets must be non-decreasing (graph %u)", g);
            break;
        }
        const uint32_t n = v1 - v0, m = e1 - e0;
        BatchGraph bg;
        memset(&bg, 0, sizeof(bg));
        bg.v0 = v0;
        bg.n = n;
        bg.e0 = e0;
        bg.ob = v0 + g;
        uint64_t acc = 0;
        for (uint32_t r = 0; r < n && st == AG_OK; r++) {
            const uint32_t idx = inv_edges_idx[v0 + r], d = inv_degree[v0 + r];
            if ((uint64_t)idx + d > m || acc + d > m) {
                st = ag_set_error(AG_ERR_INVALID, "graph %u row %u runs past its edges", g, r);
                break;
            }
            hoff[bg.ob + r] = (uint32_t)acc;
            for (uint32_t k = 0; k < d; k++) {
                const uint32_t s = inv_src[e0 + idx + k];
                if (s >= n) {
                    st = ag_set_error(AG_ERR_INVALID, "graph %u source id out of range", g);
                    break;
                }
                hsrc[e0 + acc + k] = s;
            }
            acc += d;
        }
        if (st != AG_OK)
            break;
        hoff[bg.ob + n] = (uint32_t)acc;
        bg.m = (uint32_t)acc;
        hg[g] = bg;
        keys[g] = ((uint64_t)(0xFFFFFFFFu - bg.m) << 32) | g;  // largest first, then by index
    }
    if (st != AG_OK) {
        free(hg); free(hoff); free(hsrc); free(horder); free(keys);
        return st;
    }
    qsort(keys, G, sizeof(uint64_t), [](const void *a, const void *b) {
        uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
        return (x > y) - (x < y);
    });
    for (uint32_t g = 0; g < G; g++)
        horder[g] = (uint32_t)keys[g];
    free(keys);

    hipError_t he = hipSetDevice(dev->ordinal);
    ag_batch *b = (ag_batch *)calloc(1, sizeof(ag_batch));
    if (he != hipSuccess || !b) {
        free(hg); free(hoff); free(hsrc); free(horder); free(b);
        return he != hipSuccess ? ag_hip_fail(he, "hipSetDevice", __FILE__, __LINE__)
                                : ag_set_error(AG_ERR_OOM, "host allocation failed");
    }
    b->dev = dev;
    b->G = G;
    b->total_n = N;
    b->total_m = M;
    b->h_graphs = hg;
    b->h_off = hoff;
    b->h_order = horder;
    dev->live_graphs++;
    BatchGraph *dgraphs = NULL;
    uint32_t *doff = NULL, *dsrc = NULL, *dod = NULL, *dit = NULL, *ddone = NULL, *dnd = NULL;
    float *drank = NULL, *dc0 = NULL, *dc1 = NULL, *dout = NULL;
    double *derr = NULL;
    BatchCfg *dcfg = NULL;
    if (st == AG_OK) st = balloc(b, &dgraphs, G, hg);
    if (st == AG_OK) st = balloc(b, &doff, (size_t)N + G, hoff);
    if (st == AG_OK) st = balloc(b, &dsrc, M, hsrc);
    if (st == AG_OK) st = balloc(b, &dod, N, out_degree);
    if (st == AG_OK) st = balloc<float>(b, &drank, N, NULL);
    if (st == AG_OK) st = balloc<float>(b, &dc0, N, NULL);
    if (st == AG_OK) st = balloc<float>(b, &dc1, N, NULL);
    if (st == AG_OK) st = balloc<float>(b, &dout, N, NULL);
    if (st == AG_OK) st = balloc<uint32_t>(b, &dit, G, NULL);
    if (st == AG_OK) st = balloc<double>(b, &derr, G, NULL);
    if (st == AG_OK) st = balloc<uint32_t>(b, &ddone, G, NULL);
    if (st == AG_OK) st = balloc<uint32_t>(b, &dnd, 1, NULL);
    if (st == AG_OK) st = balloc<BatchCfg>(b, &dcfg, 1, NULL);
    if (st == AG_OK) {
        he = hipHostMalloc((void **)&b->cfg_host, sizeof(BatchCfg), hipHostMallocDefault);
        if (he == hipSuccess)
            he = hipHostMalloc((void **)&b->ndone_host, sizeof(uint32_t), hipHostMallocDefault);
        if (he != hipSuccess)
            st = ag_hip_fail(he, "hipHostMalloc", __FILE__, __LINE__);
    }
    if (st == AG_OK)
        st = ag_wait_stream(dev);  // host staging buffers must outlive the copies
    free(hsrc);
    if (st == AG_OK) {
        memset(&b->a, 0, sizeof(b->a));
        b->a.graphs = dgraphs;
        b->a.off = doff;
        b->a.src = dsrc;
        b->a.outdeg = dod;
        b->a.rank = drank;
        b->a.c0 = dc0;
        b->a.c1 = dc1;
        b->a.out = dout;
        b->a.iters = dit;
        b->a.err = derr;
        b->a.done = ddone;
        b->a.ndone = dnd;
        b->a.cfg = dcfg;
        b->a.total_n = N;
        st = build_layout(b, AG_BATCH_PART_EDGES, AG_BATCH_HEAVY_DEG);
    }
    if (st != AG_OK) {
        ag_batch_free(b);
        return st;
    }
    *out = b;
    return AG_OK;
}

static uint32_t part_grid(const ag_batch *b)
{
    static const uint32_t per_cu = ag_resident_per_cu(batch_part_kernel, AG_WG);
    uint32_t grid = b->dev->num_cus * (per_cu < 8u ? per_cu : 8u);
    if (grid > b->a.nparts)
        grid = b->a.nparts;
    return grid ? grid : 1u;
}

static hipError_t enqueue_head(ag_batch *b, hipStream_t s)
{
    const uint32_t g1 = (b->G + 255) / 256;
    batch_reset_kernel<<<g1, 256, 0, s>>>(b->a);
    const uint32_t gv = (b->total_n + 255) / 256 < 4096u ? (b->total_n + 255) / 256 : 4096u;
    batch_init_kernel<<<gv ? gv : 1u, 256, 0, s>>>(b->a);
    return hipGetLastError();
}

static hipError_t enqueue_iterations(ag_batch *b, uint32_t it0, uint32_t it1, hipStream_t s)
{
    const uint32_t grid = part_grid(b), gf = (b->G + 255) / 256;
    for (uint32_t it = it0; it < it1; it++) {
        batch_part_kernel<<<grid, AG_WG, 0, s>>>(b->a, it);
        batch_fin_kernel<<<gf, 256, 0, s>>>(b->a, it);
    }
    return hipGetLastError();
}

static hipError_t enqueue_tail(ag_batch *b, hipStream_t s)
{
    batch_out_kernel<<<b->G < 4096u ? b->G : 4096u, 256, 0, s>>>(b->a);
    return hipGetLastError();
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
    const uint32_t part_edges = o.part_edges ? o.part_edges : AG_BATCH_PART_EDGES;
    const uint32_t heavy_deg = o.heavy_deg ? o.heavy_deg : AG_BATCH_HEAVY_DEG;
    const uint32_t maxit = o.max_iterations;
    AG_HIP(hipSetDevice(b->dev->ordinal));
    hipStream_t s = b->dev->stream;
    int st = ag_wait_stream(b->dev);  // the pinned staging buffers may still feed a copy
    if (st != AG_OK)
        return st;
    if (part_edges != b->part_edges || heavy_deg != b->heavy_deg) {
        st = build_layout(b, part_edges, heavy_deg);
        if (st != AG_OK)
            return st;
    }
    // options go through a device buffer so a captured solve stays valid
    BatchCfg c;
    c.max_iter = maxit;
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
    if (he == hipSuccess && o.use_hip_graph) {
        if (b->exec && b->exec_maxit != maxit)
            drop_capture(b);
        if (!b->exec) {
            he = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
            if (he == hipSuccess) {
                hipError_t qe = enqueue_head(b, s);
                if (qe == hipSuccess)
                    qe = enqueue_iterations(b, 0, maxit, s);
                if (qe == hipSuccess)
                    qe = enqueue_tail(b, s);
                hipError_t ce = hipStreamEndCapture(s, &b->graph);
                he = qe != hipSuccess ? qe : ce;
                if (qe != hipSuccess && ce == hipSuccess) {
                    (void)hipGraphDestroy(b->graph);
                    b->graph = NULL;
                }
            }
            if (he == hipSuccess)
                he = hipGraphInstantiate(&b->exec, b->graph, NULL, NULL, 0);
            if (he == hipSuccess)
                b->exec_maxit = maxit;
        }
        if (he == hipSuccess)
            he = hipGraphLaunch(b->exec, s);
    } else if (he == hipSuccess) {
        // eager: iterations queued in chunks, polling the stopped-graph count in between
        he = enqueue_head(b, s);
        const uint32_t chunk = 16;
        for (uint32_t it = 0; he == hipSuccess && st == AG_OK && it < maxit; it += chunk) {
            he = enqueue_iterations(b, it, it + chunk < maxit ? it + chunk : maxit, s);
            if (he == hipSuccess)
                he = hipMemcpyAsync(b->ndone_host, b->a.ndone, sizeof(uint32_t),
                                    hipMemcpyDeviceToHost, s);
            if (he == hipSuccess)
                st = ag_wait_stream(b->dev);
            if (st == AG_OK && he == hipSuccess && *b->ndone_host >= b->G)
                break;
        }
        if (he == hipSuccess && st == AG_OK)
            he = enqueue_tail(b, s);
    }
    if (he == hipSuccess && st == AG_OK)
        he = hipEventRecord(e1, s);
    if (he == hipSuccess && st == AG_OK && ranks_out && b->total_n)
        he = hipMemcpyAsync(ranks_out, b->a.out, (size_t)b->total_n * sizeof(float),
                            hipMemcpyDeviceToHost, s);
    if (he == hipSuccess && st == AG_OK && iterations)
        he = hipMemcpyAsync(iterations, b->a.iters, (size_t)b->G * sizeof(uint32_t),
                            hipMemcpyDeviceToHost, s);
    if (he == hipSuccess && st == AG_OK && error)
        he = hipMemcpyAsync(error, b->a.err, (size_t)b->G * sizeof(double), hipMemcpyDeviceToHost, s);
    if (st == AG_OK)
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
