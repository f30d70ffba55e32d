/*
 * graphCSR.h -- compressed-sparse-row graph with its inverse (in-edge) view.
 *
 * The reference reads these fields (02_capi_graph/src/capi_utils/capienv.c:837-869,
 * 02_capi_graph/src/algorithms/capi/pageRank.c:967-973,1029-1030):
 *   num_vertices, num_edges, vertices->{out_degree,in_degree,edges_idx},
 *   sorted_edges_array->{edges_array_src,edges_array_dest,edges_array_weight},
 *   inverse_vertices->{...}, inverse_sorted_edges_array->{...}, max_weight, avg_degree.
 * Its header lives in the absent OpenGraph submodule, so the field order below is this
 * library's own: a caller recompiled against this header is a drop-in; binary layout
 * compatibility with OpenGraph's struct is not claimed.
 *
 * Conventions (same as the reference's use of the arrays):
 *   vertices->out_degree[v]        forward out-degree of v
 *   vertices->edges_idx[v]         first forward edge of v in sorted_edges_array
 *   inverse_vertices->out_degree[v] in-degree of v (row length of the inverse CSR)
 *   inverse_vertices->edges_idx[v] first in-edge of v in inverse_sorted_edges_array
 *   inverse_sorted_edges_array->edges_array_dest[j]  source vertex of the j-th in-edge
 *   inverse_sorted_edges_array->edges_array_src[j]   the row (destination) it belongs to
 * Both views are sorted: forward rows by (src, dest), inverse rows by (dest, src).
 */
#ifndef GRAPHCSR_H
#define GRAPHCSR_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

struct Vertex
{
    uint32_t *out_degree;
    uint32_t *in_degree;
    uint32_t *edges_idx;
};

struct EdgeList
{
    uint32_t num_edges;
    uint32_t num_vertices;
    uint32_t *edges_array_src;
    uint32_t *edges_array_dest;
    float *edges_array_weight;   /* NULL when the graph is unweighted */
    float max_weight;
};

struct GraphCSR
{
    uint32_t num_edges;
    uint32_t num_vertices;
    float max_weight;
    uint32_t avg_degree;
    struct Vertex *vertices;
    struct EdgeList *sorted_edges_array;
    struct Vertex *inverse_vertices;
    struct EdgeList *inverse_sorted_edges_array;
};

/* ---- graph construction (library-owned memory; release with graphCSRFree) ---- */

/* Edge-list input: `src[i] -> dst[i]`, i < m. Vertices are 0..n-1 (n = 0 means
 * "1 + largest id"). Self-loops and duplicate edges are removed when `simplify` is set.
 * `weight` may be NULL; with weights, the first occurrence of a duplicate edge wins.
 * Returns NULL on allocation failure or an id >= n. */
struct GraphCSR *graphCSRFromEdgeList(uint32_t n, uint64_t m, const uint32_t *src,
                                      const uint32_t *dst, const float *weight,
                                      int simplify);

/* Synthetic generators (seeded, deterministic for a given seed regardless of the
 * thread count). RMAT draws `edge_factor * 2^scale` edges with quadrant
 * probabilities a,b,c,(1-a-b-c); Erdos-Renyi draws m uniform (src,dst) pairs. */
struct GraphCSR *graphCSRGenerateRMAT(uint32_t scale, uint32_t edge_factor, double a,
                                      double b, double c, uint32_t seed, int simplify);
struct GraphCSR *graphCSRGenerateER(uint32_t num_vertices, uint64_t num_edges,
                                    uint32_t seed, int simplify);

/* Raw generated edge stream (before simplification), for tests. Caller frees. */
int graphGenerateRMATEdges(uint32_t scale, uint64_t num_edges, double a, double b,
                           double c, uint32_t seed, uint32_t *src, uint32_t *dst);

/* Binary edge list: little-endian u32 (src,dst) pairs, no header, the format of the
 * reference's 04_test_graphs/TEST/graphbrew/graph.bin. */
struct GraphCSR *graphCSRLoadBinaryEdgeList(const char *path, int simplify);

/* ---- graph files (the reference driver's -z read formats, accel-graph.c:51-53, and its
 * -c/-x conversion, accel-graph.c:115-128). On failure the loaders return NULL and the
 * savers -1; graphIOLastError() then describes the failure (file, line, reason). ---- */

enum {
    GRAPH_FILE_TEXT_EDGE_LIST = 0,    /* "src dst [weight]" per line */
    GRAPH_FILE_BINARY_EDGE_LIST = 1,  /* u32 LE (src,dst) pairs, no header */
    GRAPH_FILE_CSR_SERIALIZED = 2     /* every array of struct GraphCSR, checksummed */
};

/* Text edge list: one edge per line, fields separated by spaces, tabs or commas, '#' and
 * '%' lines and blank lines skipped, further columns ignored. With `weighted` every line
 * must carry a third (float) column. Vertex count = 1 + largest id. */
struct GraphCSR *graphCSRLoadTextEdgeList(const char *path, int weighted, int simplify);

/* A graph saved by graphCSRSaveSerialized, every array as it was (no rebuild). The file is
 * checked (checksum, offsets, id ranges, the two views against each other) before use. */
struct GraphCSR *graphCSRLoadSerialized(const char *path);

/* `file_format` is one of GRAPH_FILE_*; `simplify` applies to the edge-list formats. */
struct GraphCSR *graphCSRLoadFile(const char *path, int file_format, int simplify);

/* Writers. The edge-list forms write the forward edges in CSR order (with the weights as a
 * third text column when the graph has them; the binary edge list has no weights), so
 * loading a saved simplified graph rebuilds the same arrays. Return 0 on success. */
int graphCSRSaveTextEdgeList(const struct GraphCSR *graph, const char *path);
int graphCSRSaveBinaryEdgeList(const struct GraphCSR *graph, const char *path);
int graphCSRSaveSerialized(const struct GraphCSR *graph, const char *path);
int graphCSRSaveFile(const struct GraphCSR *graph, const char *path, int file_format);

/* The last graph-file error of the calling thread ("" when none). */
const char *graphIOLastError(void);

/* Attach per-edge weights given in forward (sorted_edges_array) order; the inverse
 * weights are permuted to match. Returns 0 on success. */
int graphCSRSetWeights(struct GraphCSR *graph, const float *forward_weights);

void graphCSRFree(struct GraphCSR *graph);

/* An empty graph of n vertices and m edges with every array allocated and zeroed (weights
 * when `weighted`); what the builders fill. NULL when out of memory. */
struct GraphCSR *graphCSRAlloc(uint32_t num_vertices, uint32_t num_edges, int weighted);

#ifdef __cplusplus
}
#endif

#endif
