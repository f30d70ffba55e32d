/*
 * graph_csr.c -- CSR + inverse-CSR construction, synthetic generators and the binary
 * edge-list loader.
 *
 * The reference builds these structures in OpenGraph (absent from the reference tree:
 * 02_capi_graph/Makefile:15,227-229 name it); the pull path only consumes the result
 * (02_capi_graph/src/capi_utils/capienv.c:837-869). Construction here:
 *   1. bucket the edges by source (counting sort), sort each row by (dest, input index),
 *      drop self-loops and repeated (src,dest) pairs keeping the first occurrence;
 *   2. bucket the surviving edges by destination (counting sort) and sort each inverse
 *      row by source, giving the inverse CSR the pull gather walks.
 * Both passes are OpenMP-parallel; the output depends only on the input edge list.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#include "graphCSR.h"
#include "mt19937.h"

static void *xcalloc(size_t count, size_t size)
{
    if (count == 0)
        count = 1;
    return calloc(count, size);
}

static int cmp_u64(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return (x > y) - (x < y);
}

static void sort_u64_range(uint64_t *v, uint64_t len)
{
    if (len <= 32) {
        for (uint64_t i = 1; i < len; i++) {
            uint64_t key = v[i];
            uint64_t j = i;
            while (j > 0 && v[j - 1] > key) {
                v[j] = v[j - 1];
                j--;
            }
            v[j] = key;
        }
    } else {
        qsort(v, (size_t)len, sizeof(uint64_t), cmp_u64);
    }
}

/*
 * Bucket `m` items by key[i] < n into rows; within a row items are sorted by the 64-bit
 * value packed as (secondary << 32) | payload. Returns row offsets [n+1] and the sorted
 * packed values [m]. The packed value fully orders the items, so the result is the same
 * for every thread count.
 */
static int bucket_sort(uint32_t n, uint64_t m, const uint32_t *key, const uint32_t *secondary,
                       const uint32_t *payload, uint64_t **off_out, uint64_t **vals_out)
{
    uint64_t *off = (uint64_t *)xcalloc((size_t)n + 1, sizeof(uint64_t));
    uint64_t *cursor = (uint64_t *)xcalloc((size_t)n + 1, sizeof(uint64_t));
    uint64_t *vals = (uint64_t *)xcalloc((size_t)m, sizeof(uint64_t));
    if (!off || !cursor || !vals) {
        free(off);
        free(cursor);
        free(vals);
        return -1;
    }
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < m; i++) {
        #pragma omp atomic update
        off[key[i] + 1]++;
    }
    for (uint32_t v = 0; v < n; v++)
        off[v + 1] += off[v];
    memcpy(cursor, off, ((size_t)n + 1) * sizeof(uint64_t));
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < m; i++) {
        uint64_t pos;
        #pragma omp atomic capture
        pos = cursor[key[i]]++;
        vals[pos] = ((uint64_t)secondary[i] << 32) | (uint64_t)payload[i];
    }
    free(cursor);
    #pragma omp parallel for schedule(dynamic, 1024)
    for (uint32_t v = 0; v < n; v++)
        sort_u64_range(vals + off[v], off[v + 1] - off[v]);
    *off_out = off;
    *vals_out = vals;
    return 0;
}

struct GraphCSR *graphCSRAlloc(uint32_t n, uint32_t m, int weighted)
{
    struct GraphCSR *g = (struct GraphCSR *)calloc(1, sizeof(struct GraphCSR));
    if (!g)
        return NULL;
    g->num_vertices = n;
    g->num_edges = m;
    g->vertices = (struct Vertex *)calloc(1, sizeof(struct Vertex));
    g->inverse_vertices = (struct Vertex *)calloc(1, sizeof(struct Vertex));
    g->sorted_edges_array = (struct EdgeList *)calloc(1, sizeof(struct EdgeList));
    g->inverse_sorted_edges_array = (struct EdgeList *)calloc(1, sizeof(struct EdgeList));
    if (!g->vertices || !g->inverse_vertices || !g->sorted_edges_array ||
        !g->inverse_sorted_edges_array) {
        graphCSRFree(g);
        return NULL;
    }
    struct Vertex *vs[2] = {g->vertices, g->inverse_vertices};
    for (int k = 0; k < 2; k++) {
        vs[k]->out_degree = (uint32_t *)xcalloc(n, sizeof(uint32_t));
        vs[k]->in_degree = (uint32_t *)xcalloc(n, sizeof(uint32_t));
        vs[k]->edges_idx = (uint32_t *)xcalloc(n, sizeof(uint32_t));
        if (!vs[k]->out_degree || !vs[k]->in_degree || !vs[k]->edges_idx) {
            graphCSRFree(g);
            return NULL;
        }
    }
    struct EdgeList *es[2] = {g->sorted_edges_array, g->inverse_sorted_edges_array};
    for (int k = 0; k < 2; k++) {
        es[k]->num_edges = m;
        es[k]->num_vertices = n;
        es[k]->edges_array_src = (uint32_t *)xcalloc(m, sizeof(uint32_t));
        es[k]->edges_array_dest = (uint32_t *)xcalloc(m, sizeof(uint32_t));
        if (!es[k]->edges_array_src || !es[k]->edges_array_dest) {
            graphCSRFree(g);
            return NULL;
        }
        if (weighted) {
            es[k]->edges_array_weight = (float *)xcalloc(m, sizeof(float));
            if (!es[k]->edges_array_weight) {
                graphCSRFree(g);
                return NULL;
            }
        }
    }
    return g;
}

/* Touches every page of a freshly allocated graph's arrays from all threads (writing the
 * zeros they already hold), so that later bulk writes -- device-to-host copies into pageable
 * memory run on one thread -- do not take their page faults one page at a time. */
void graphCSRPrefault(struct GraphCSR *g)
{
    if (!g)
        return;
    const size_t n = g->num_vertices, m = g->num_edges;
    struct {
        void *p;
        size_t bytes;
    } a[12];
    int k = 0;
    struct Vertex *vs[2] = {g->vertices, g->inverse_vertices};
    for (int i = 0; i < 2; i++) {
        a[k].p = vs[i]->out_degree; a[k++].bytes = n * 4;
        a[k].p = vs[i]->in_degree;  a[k++].bytes = n * 4;
        a[k].p = vs[i]->edges_idx;  a[k++].bytes = n * 4;
    }
    struct EdgeList *es[2] = {g->sorted_edges_array, g->inverse_sorted_edges_array};
    for (int i = 0; i < 2; i++) {
        a[k].p = es[i]->edges_array_src;    a[k++].bytes = m * 4;
        a[k].p = es[i]->edges_array_dest;   a[k++].bytes = m * 4;
        a[k].p = es[i]->edges_array_weight; a[k++].bytes = es[i]->edges_array_weight ? m * 4 : 0;
    }
    for (int i = 0; i < k; i++) {
        volatile char *c = (volatile char *)a[i].p;
        const size_t bytes = a[i].bytes;
        #pragma omp parallel for schedule(static)
        for (size_t q = 0; q < bytes; q += 4096)
            c[q] = 0;
    }
}

void graphCSRFree(struct GraphCSR *g)
{
    if (!g)
        return;
    struct Vertex *vs[2] = {g->vertices, g->inverse_vertices};
    for (int k = 0; k < 2; k++) {
        if (vs[k]) {
            free(vs[k]->out_degree);
            free(vs[k]->in_degree);
            free(vs[k]->edges_idx);
            free(vs[k]);
        }
    }
    struct EdgeList *es[2] = {g->sorted_edges_array, g->inverse_sorted_edges_array};
    for (int k = 0; k < 2; k++) {
        if (es[k]) {
            free(es[k]->edges_array_src);
            free(es[k]->edges_array_dest);
            free(es[k]->edges_array_weight);
            free(es[k]);
        }
    }
    free(g);
}

struct GraphCSR *graphCSRFromEdgeList(uint32_t n, uint64_t m, const uint32_t *src,
                                      const uint32_t *dst, const float *weight, int simplify)
{
    if (m > 0 && (!src || !dst))
        return NULL;
    if (m > 0xFFFFFFFFull)
        return NULL;
    if (n == 0) {
        uint32_t mx = 0;
        int any = 0;
        for (uint64_t i = 0; i < m; i++) {
            any = 1;
            if (src[i] > mx) mx = src[i];
            if (dst[i] > mx) mx = dst[i];
        }
        n = any ? mx + 1 : 0;
    }
    for (uint64_t i = 0; i < m; i++)
        if (src[i] >= n || dst[i] >= n)
            return NULL;

    /* pass 1: rows by source, sorted by (dest, input index) */
    uint32_t *idx = (uint32_t *)xcalloc((size_t)m, sizeof(uint32_t));
    if (!idx)
        return NULL;
    for (uint64_t i = 0; i < m; i++)
        idx[i] = (uint32_t)i;
    uint64_t *foff = NULL, *fvals = NULL;
    if (bucket_sort(n, m, src, dst, idx, &foff, &fvals)) {
        free(idx);
        return NULL;
    }
    free(idx);

    /* drop self-loops and duplicates in place, per row */
    uint32_t *fdeg = (uint32_t *)xcalloc(n, sizeof(uint32_t));
    if (!fdeg) {
        free(foff);
        free(fvals);
        return NULL;
    }
    #pragma omp parallel for schedule(dynamic, 1024)
    for (uint32_t v = 0; v < n; v++) {
        uint64_t b = foff[v], e = foff[v + 1], w = b;
        for (uint64_t j = b; j < e; j++) {
            uint32_t d = (uint32_t)(fvals[j] >> 32);
            if (simplify) {
                if (d == v)
                    continue;
                if (w > b && (uint32_t)(fvals[w - 1] >> 32) == d)
                    continue;
            }
            fvals[w++] = fvals[j];
        }
        fdeg[v] = (uint32_t)(w - b);
    }
    uint64_t total = 0;
    for (uint32_t v = 0; v < n; v++)
        total += fdeg[v];

    struct GraphCSR *g = graphCSRAlloc(n, (uint32_t)total, weight != NULL);
    if (!g) {
        free(foff);
        free(fvals);
        free(fdeg);
        return NULL;
    }
    {
        uint64_t acc = 0;
        for (uint32_t v = 0; v < n; v++) {
            g->vertices->edges_idx[v] = (uint32_t)acc;
            g->vertices->out_degree[v] = fdeg[v];
            acc += fdeg[v];
        }
    }
    struct EdgeList *fe = g->sorted_edges_array;
    #pragma omp parallel for schedule(dynamic, 1024)
    for (uint32_t v = 0; v < n; v++) {
        uint64_t b = foff[v];
        uint32_t o = g->vertices->edges_idx[v];
        for (uint32_t k = 0; k < fdeg[v]; k++) {
            uint64_t pv = fvals[b + k];
            fe->edges_array_src[o + k] = v;
            fe->edges_array_dest[o + k] = (uint32_t)(pv >> 32);
            if (weight)
                fe->edges_array_weight[o + k] = weight[(uint32_t)pv];
        }
    }
    free(foff);
    free(fvals);
    free(fdeg);

    /* pass 2: inverse rows by destination, sorted by source; payload = forward index */
    uint32_t mm = g->num_edges;
    uint32_t *fidx = (uint32_t *)xcalloc(mm, sizeof(uint32_t));
    if (!fidx) {
        graphCSRFree(g);
        return NULL;
    }
    for (uint32_t j = 0; j < mm; j++)
        fidx[j] = j;
    uint64_t *ioff = NULL, *ivals = NULL;
    if (bucket_sort(n, mm, fe->edges_array_dest, fe->edges_array_src, fidx, &ioff, &ivals)) {
        free(fidx);
        graphCSRFree(g);
        return NULL;
    }
    free(fidx);
    struct EdgeList *ie = g->inverse_sorted_edges_array;
    #pragma omp parallel for schedule(dynamic, 1024)
    for (uint32_t v = 0; v < n; v++) {
        uint64_t b = ioff[v], e = ioff[v + 1];
        g->inverse_vertices->edges_idx[v] = (uint32_t)b;
        g->inverse_vertices->out_degree[v] = (uint32_t)(e - b);
        for (uint64_t j = b; j < e; j++) {
            uint64_t pv = ivals[j];
            ie->edges_array_src[j] = v;
            ie->edges_array_dest[j] = (uint32_t)(pv >> 32);
            if (weight)
                ie->edges_array_weight[j] = fe->edges_array_weight[(uint32_t)pv];
        }
    }
    free(ioff);
    free(ivals);

    /* in_degree of the forward view is the inverse row length and vice versa */
    float maxw = 0.0f;
    for (uint32_t v = 0; v < n; v++) {
        g->vertices->in_degree[v] = g->inverse_vertices->out_degree[v];
        g->inverse_vertices->in_degree[v] = g->vertices->out_degree[v];
    }
    if (weight) {
        for (uint32_t j = 0; j < mm; j++)
            if (fe->edges_array_weight[j] > maxw)
                maxw = fe->edges_array_weight[j];
    }
    g->max_weight = maxw;
    fe->max_weight = maxw;
    ie->max_weight = maxw;
    g->avg_degree = n ? (uint32_t)(mm / n) : 0;
    return g;
}

int graphCSRSetWeights(struct GraphCSR *g, const float *fw)
{
    if (!g || (!fw && g->num_edges))
        return -1;
    uint32_t m = g->num_edges, n = g->num_vertices;
    struct EdgeList *fe = g->sorted_edges_array, *ie = g->inverse_sorted_edges_array;
    if (!fe->edges_array_weight)
        fe->edges_array_weight = (float *)xcalloc(m, sizeof(float));
    if (!ie->edges_array_weight)
        ie->edges_array_weight = (float *)xcalloc(m, sizeof(float));
    if (!fe->edges_array_weight || !ie->edges_array_weight)
        return -1;
    memcpy(fe->edges_array_weight, fw, (size_t)m * sizeof(float));
    /* cursor walk: inverse rows are sorted by source and forward rows by destination, so
     * scanning sources in increasing destination order meets each forward row in order */
    uint32_t *cursor = (uint32_t *)xcalloc(n, sizeof(uint32_t));
    if (!cursor)
        return -1;
    memcpy(cursor, g->vertices->edges_idx, (size_t)n * sizeof(uint32_t));
    float maxw = 0.0f;
    for (uint32_t v = 0; v < n; v++) {
        uint32_t b = g->inverse_vertices->edges_idx[v];
        uint32_t e = b + g->inverse_vertices->out_degree[v];
        for (uint32_t j = b; j < e; j++) {
            uint32_t s = ie->edges_array_dest[j];
            uint32_t fj = cursor[s]++;
            ie->edges_array_weight[j] = fw[fj];
        }
    }
    free(cursor);
    for (uint32_t j = 0; j < m; j++)
        if (fw[j] > maxw)
            maxw = fw[j];
    g->max_weight = maxw;
    fe->max_weight = maxw;
    ie->max_weight = maxw;
    return 0;
}

/* ------------------------------------------------------------------ generators */

#define GEN_CHUNK (1u << 20)

static uint32_t chunk_seed(uint32_t seed, uint64_t chunk)
{
    return seed ^ (uint32_t)(chunk * 0x9E3779B9ull);
}

int graphGenerateRMATEdges(uint32_t scale, uint64_t m, double a, double b, double c,
                           uint32_t seed, uint32_t *src, uint32_t *dst)
{
    if (scale > 31 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return -1;
    double ab = a + b, abc = a + b + c;
    uint64_t nchunks = (m + GEN_CHUNK - 1) / GEN_CHUNK;
    #pragma omp parallel for schedule(dynamic, 1)
    for (uint64_t ch = 0; ch < nchunks; ch++) {
        mt19937state st;
        initializeMersenneState(&st, chunk_seed(seed, ch));
        uint64_t e0 = ch * GEN_CHUNK, e1 = e0 + GEN_CHUNK < m ? e0 + GEN_CHUNK : m;
        for (uint64_t e = e0; e < e1; e++) {
            uint32_t s = 0, d = 0;
            for (uint32_t lvl = 0; lvl < scale; lvl++) {
                double r = generateRandDouble(&st);
                uint32_t bit = 1u << (scale - 1 - lvl);
                if (r < a) {
                } else if (r < ab) {
                    d |= bit;
                } else if (r < abc) {
                    s |= bit;
                } else {
                    s |= bit;
                    d |= bit;
                }
            }
            src[e] = s;
            dst[e] = d;
        }
    }
    return 0;
}

struct GraphCSR *graphCSRGenerateRMAT(uint32_t scale, uint32_t edge_factor, double a, double b,
                                      double c, uint32_t seed, int simplify)
{
    if (scale > 31)
        return NULL;
    uint32_t n = 1u << scale;
    uint64_t m = (uint64_t)edge_factor << scale;
    if (m > 0xFFFFFFFFull)
        return NULL;
    uint32_t *src = (uint32_t *)xcalloc((size_t)m, sizeof(uint32_t));
    uint32_t *dst = (uint32_t *)xcalloc((size_t)m, sizeof(uint32_t));
    if (!src || !dst) {
        free(src);
        free(dst);
        return NULL;
    }
    struct GraphCSR *g = NULL;
    if (graphGenerateRMATEdges(scale, m, a, b, c, seed, src, dst) == 0)
        g = graphCSRFromEdgeList(n, m, src, dst, NULL, simplify);
    free(src);
    free(dst);
    return g;
}

struct GraphCSR *graphCSRGenerateER(uint32_t n, uint64_t m, uint32_t seed, int simplify)
{
    if (n == 0 || m > 0xFFFFFFFFull)
        return NULL;
    uint32_t *src = (uint32_t *)xcalloc((size_t)m, sizeof(uint32_t));
    uint32_t *dst = (uint32_t *)xcalloc((size_t)m, sizeof(uint32_t));
    if (!src || !dst) {
        free(src);
        free(dst);
        return NULL;
    }
    uint64_t nchunks = (m + GEN_CHUNK - 1) / GEN_CHUNK;
    #pragma omp parallel for schedule(dynamic, 1)
    for (uint64_t ch = 0; ch < nchunks; ch++) {
        mt19937state st;
        initializeMersenneState(&st, chunk_seed(seed, ch));
        uint64_t e0 = ch * GEN_CHUNK, e1 = e0 + GEN_CHUNK < m ? e0 + GEN_CHUNK : m;
        for (uint64_t e = e0; e < e1; e++) {
            src[e] = (uint32_t)(((uint64_t)generateRandInt(&st) * n) >> 32);
            dst[e] = (uint32_t)(((uint64_t)generateRandInt(&st) * n) >> 32);
        }
    }
    struct GraphCSR *g = graphCSRFromEdgeList(n, m, src, dst, NULL, simplify);
    free(src);
    free(dst);
    return g;
}

struct GraphCSR *graphCSRLoadBinaryEdgeList(const char *path, int simplify)
{
    FILE *f = fopen(path, "rb");
    if (!f)
        return NULL;
    if (fseek(f, 0, SEEK_END) != 0) {
        fclose(f);
        return NULL;
    }
    long size = ftell(f);
    rewind(f);
    if (size < 0 || size % 8 != 0) {
        fclose(f);
        return NULL;
    }
    uint64_t m = (uint64_t)size / 8;
    uint32_t *raw = (uint32_t *)xcalloc((size_t)m * 2, sizeof(uint32_t));
    uint32_t *src = (uint32_t *)xcalloc((size_t)m, sizeof(uint32_t));
    uint32_t *dst = (uint32_t *)xcalloc((size_t)m, sizeof(uint32_t));
    if (!raw || !src || !dst || fread(raw, 8, (size_t)m, f) != (size_t)m) {
        free(raw);
        free(src);
        free(dst);
        fclose(f);
        return NULL;
    }
    fclose(f);
    for (uint64_t i = 0; i < m; i++) {
        const unsigned char *p = (const unsigned char *)(raw + 2 * i);
        src[i] = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
                 ((uint32_t)p[3] << 24);
        dst[i] = (uint32_t)p[4] | ((uint32_t)p[5] << 8) | ((uint32_t)p[6] << 16) |
                 ((uint32_t)p[7] << 24);
    }
    free(raw);
    struct GraphCSR *g = graphCSRFromEdgeList(0, m, src, dst, NULL, simplify);
    free(src);
    free(dst);
    return g;
}
