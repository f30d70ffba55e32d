/*
 * accel_context.c -- process-wide device context and device-graph cache used by the drop-in
 * entry points (pageRank.c, SPMV.c).
 *
 * The reference opens and attaches the FPGA for every call (setupAFUGraphCSR/startAFU,
 * 02_capi_graph/src/capi_utils/capienv.c:329-451) and releases it before returning
 * (releaseAFU, :765-782); its failures print one line and exit (capienv.c:80-132). Here the
 * GPU is opened once (ordinal from ACCELGRAPH_DEVICE, default 0) and the last uploaded graph
 * stays resident in HBM. The cache key is the GraphCSR pointer, its sizes, its array pointers
 * and a 64-bit digest of the contents of every uploaded array, recomputed on the host at each
 * call: the reference's callers free a graph and build another one of the same size, often at
 * the same addresses (02_capi_graph/src/tests/sweep_order-PR-performance-graph.c:205-233), or
 * rewrite arrays in place, and never release anything, so pointers alone would hand them the
 * previous graph's ranks. The digest reads the upload set once (RMAT-25: 2.5 GB, tens of ms on
 * the host cores, against about 2 s to upload and plan). ACCELGRAPH_TRUST_GRAPH_POINTERS=1
 * skips it for callers that promise to call accelgraphReleaseGraph before any such change.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "accel_context.h"

static ag_device *g_dev = NULL;
static ag_dgraph *g_cached = NULL;
static struct {
    const struct GraphCSR *graph;
    uint32_t n, m;
    const void *a, *b, *c, *d, *w;
    uint64_t digest;
    int ref_order;
} g_key;

#define DIGEST_CHUNK (1u << 20) /* u32 elements per independently hashed chunk */

static inline uint64_t mix64(uint64_t h)
{
    /* splitmix64 finaliser (Steele, Lea, Flood 2014) */
    h ^= h >> 30;
    h *= 0xbf58476d1ce4e5b9ull;
    h ^= h >> 27;
    h *= 0x94d049bb133111ebull;
    h ^= h >> 31;
    return h;
}

/* Digest of n 32-bit words: chunks of DIGEST_CHUNK words are hashed in parallel, then the
 * chunk hashes are folded in chunk order, so the value does not depend on the thread count. */
static uint64_t digest_u32(const void *p, size_t n, uint64_t seed)
{
    if (!p || n == 0)
        return mix64(seed ^ 0x9e3779b97f4a7c15ull);
    const uint32_t *a = (const uint32_t *)p;
    const size_t chunks = (n + DIGEST_CHUNK - 1) / DIGEST_CHUNK;
    uint64_t *part = (uint64_t *)malloc(chunks * sizeof(uint64_t));
    if (!part)
        accelgraphFail("graph digest", AG_ERR_OOM);
    #pragma omp parallel for schedule(static)
    for (size_t c = 0; c < chunks; c++) {
        const size_t lo = c * DIGEST_CHUNK, hi = lo + DIGEST_CHUNK < n ? lo + DIGEST_CHUNK : n;
        uint64_t h0 = 0x243f6a8885a308d3ull ^ c, h1 = 0x13198a2e03707344ull, h2 = 0xa4093822299f31d0ull,
                 h3 = 0x082efa98ec4e6c89ull;
        size_t i = lo;
        for (; i + 8 <= hi; i += 8) {
            uint64_t w0, w1, w2, w3;
            memcpy(&w0, a + i, 8);
            memcpy(&w1, a + i + 2, 8);
            memcpy(&w2, a + i + 4, 8);
            memcpy(&w3, a + i + 6, 8);
            h0 = (h0 ^ w0) * 0x9e3779b97f4a7c15ull;
            h1 = (h1 ^ w1) * 0xc2b2ae3d27d4eb4full;
            h2 = (h2 ^ w2) * 0x165667b19e3779f9ull;
            h3 = (h3 ^ w3) * 0xd6e8feb86659fd93ull;
            h0 ^= h0 >> 29;
            h1 ^= h1 >> 31;
            h2 ^= h2 >> 27;
            h3 ^= h3 >> 33;
        }
        for (; i < hi; i++)
            h0 = mix64(h0 ^ a[i]);
        part[c] = mix64(h0 ^ mix64(h1 ^ mix64(h2 ^ mix64(h3 ^ (uint64_t)(hi - lo)))));
    }
    uint64_t h = mix64(seed ^ (uint64_t)n);
    for (size_t c = 0; c < chunks; c++)
        h = mix64(h ^ part[c]) + c;
    free(part);
    return h;
}

static uint64_t graph_digest(const struct GraphCSR *graph, const float *w)
{
    const size_t n = graph->num_vertices, m = graph->num_edges;
    uint64_t h = digest_u32(graph->vertices->out_degree, n, 1);
    h = mix64(h ^ digest_u32(graph->inverse_vertices->edges_idx, n, 2));
    h = mix64(h ^ digest_u32(graph->inverse_vertices->out_degree, n, 3));
    h = mix64(h ^ digest_u32(graph->inverse_sorted_edges_array->edges_array_dest, m, 4));
    if (w)
        h = mix64(h ^ digest_u32(w, m, 5));
    return h;
}

static int trust_pointers(void)
{
    const char *e = getenv("ACCELGRAPH_TRUST_GRAPH_POINTERS");
    return e && *e && strcmp(e, "0") != 0;
}

void accelgraphFail(const char *what, int status)
{
    fprintf(stderr, "[accelgraph] %s failed: %s (%s)\n", what, ag_status_string(status),
            ag_last_error());
    exit(EXIT_FAILURE);
}

ag_device *accelgraphDevice(void)
{
    if (g_dev)
        return g_dev;
    int ordinal = 0;
    const char *e = getenv("ACCELGRAPH_DEVICE");
    if (e && *e)
        ordinal = atoi(e);
    int st = ag_device_open(ordinal, &g_dev);
    if (st != AG_OK)
        accelgraphFail("ag_device_open", st);
    return g_dev;
}

void accelgraphReleaseGraph(void)
{
    if (g_cached) {
        ag_graph_free(g_cached);
        g_cached = NULL;
    }
}

ag_dgraph *accelgraphUploadGraph(const struct GraphCSR *graph, int with_weights)
{
    ag_device *dev = accelgraphDevice();
    const float *w = with_weights ? graph->inverse_sorted_edges_array->edges_array_weight : NULL;
    const int trust = trust_pointers();
    const uint64_t digest = trust ? 0 : graph_digest(graph, w);
    /* ACCELGRAPH_FP32_ORDER=reference: every FP32 row sum in the reference's sequential
     * order, so ranks and the active-count stop equal the reference's bit for bit */
    const char *order = getenv("ACCELGRAPH_FP32_ORDER");
    const int ref_order = order && strcmp(order, "reference") == 0;
    if (g_cached && g_key.graph == graph && g_key.n == graph->num_vertices &&
        g_key.m == graph->num_edges && g_key.a == graph->vertices->out_degree &&
        g_key.b == graph->inverse_vertices->edges_idx &&
        g_key.c == graph->inverse_vertices->out_degree &&
        g_key.d == graph->inverse_sorted_edges_array->edges_array_dest && g_key.w == w &&
        g_key.ref_order == ref_order && (trust || g_key.digest == digest))
        return g_cached;
    accelgraphReleaseGraph();
    ag_plan_params prm;
    ag_plan_params_default(&prm);
    if (ref_order)
        prm.fp32_order = AG_ORDER_REFERENCE;
    int st = ag_graph_upload(dev, graph->num_vertices, graph->num_edges,
                             graph->vertices->out_degree, graph->inverse_vertices->edges_idx,
                             graph->inverse_vertices->out_degree,
                             graph->inverse_sorted_edges_array->edges_array_dest, w, &prm,
                             &g_cached);
    if (st != AG_OK)
        accelgraphFail("ag_graph_upload", st);
    g_key.graph = graph;
    g_key.n = graph->num_vertices;
    g_key.m = graph->num_edges;
    g_key.a = graph->vertices->out_degree;
    g_key.b = graph->inverse_vertices->edges_idx;
    g_key.c = graph->inverse_vertices->out_degree;
    g_key.d = graph->inverse_sorted_edges_array->edges_array_dest;
    g_key.w = w;
    g_key.digest = digest;
    g_key.ref_order = ref_order;
    return g_cached;
}

void accelgraphShutdown(void)
{
    accelgraphReleaseGraph();
    if (g_dev) {
        ag_device_close(g_dev);
        g_dev = NULL;
    }
}
