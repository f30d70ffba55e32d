/*
 * accelgraph.h -- C-ABI of the MI355X (gfx950) PageRank/SpMV pull engine.
 *
 * This is the device boundary that replaces the reference's CAPI runtime
 * (02_capi_graph/include/capi_utils/capienv.h:146-167: mapGraphCSRToWED, setupAFUGraphCSR,
 * startAFU, startCU, waitAFU, releaseAFU). Where the reference maps host arrays into a
 * work-element descriptor that an FPGA reads coherently, this layer uploads the arrays the
 * pull path reads (forward out-degree, inverse-CSR offsets/degrees/sources and optional
 * weights: capienv.c:837-869) into HBM once, builds a degree-binned pull plan, and runs
 * hand-written HIP kernels on one stream.
 *
 * All functions return an ag_status code (0 = AG_OK); ag_last_error() gives a message for
 * the calling thread. No torch or C++ types cross this boundary. One ag_device is driven
 * from one host thread at a time.
 */
#ifndef ACCELGRAPH_H
#define ACCELGRAPH_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AG_ABI_VERSION 1

typedef enum ag_status {
    AG_OK = 0,
    AG_ERR_INVALID = 1,      /* bad argument or size, rejected before any device work */
    AG_ERR_NO_DEVICE = 2,    /* no usable GPU */
    AG_ERR_HIP = 3,          /* a HIP runtime call failed */
    AG_ERR_OOM = 4,          /* device or host allocation failed */
    AG_ERR_TIMEOUT = 5,      /* bounded wait on the device expired */
    AG_ERR_UNSUPPORTED = 6   /* option not implemented for this entry point */
} ag_status;

typedef struct ag_device ag_device;     /* device context: stream, events, watchdog */
typedef struct ag_dgraph ag_dgraph;     /* graph resident in HBM plus its pull plan */
typedef struct ag_pr_session ag_pr_session;

/* ---------------------------------------------------------------- device */

int ag_device_count(int *count);
int ag_device_open(int ordinal, ag_device **out);
/* Fails with AG_ERR_INVALID (and closes nothing) while graphs uploaded on it are alive. */
int ag_device_close(ag_device *dev);
int ag_device_sync(ag_device *dev);           /* bounded by the watchdog deadline */
const char *ag_status_string(int status);
const char *ag_last_error(void);
int ag_abi_version(void);

/* Watchdog for waits on the device (the role of the reference's accelerator
 * watchdog, 03_capi_integration/accelerator_verification/host/accelerator_verification.c:
 * 86-135,289-320). run_timeout_ms bounds every wait; poll_interval_us is the sleep between
 * polls. Defaults come from ACCELERATOR_RUN_TIMEOUT_MS / ACCELERATOR_POLL_INTERVAL_US
 * (0 = unbounded). */
int ag_device_set_watchdog(ag_device *dev, uint64_t run_timeout_ms, uint32_t poll_interval_us);

/* Device-to-device copy bandwidth: copies `bytes` (rounded down to 16-byte units) between two
 * fresh HBM buffers `reps` times with the library's streaming copy kernel and reports the
 * mean GB/s (read + write bytes / time). The measured ceiling the pull kernels are compared
 * with, and a known-size kernel for calibrating hardware byte counters. */
int ag_device_copy_bandwidth(ag_device *dev, uint64_t bytes, uint32_t reps, double *gbps);

/* ---------------------------------------------------------------- graph + plan */

/* Row classes of the pull plan, by in-degree. */
enum {
    AG_BIN_THREAD = 0,   /* a lane per row, rows staged as LDS tiles; stream-order sums */
    AG_BIN_GROUP16 = 1,  /* 16 lanes per row */
    AG_BIN_WAVE = 2,     /* a 64-lane wavefront per row */
    AG_BIN_BLOCK = 3,    /* a 256-lane workgroup per row */
    AG_BIN_HUB = 4,      /* a row split into fixed segments, one workgroup each */
    AG_NBINS = 5
};

typedef struct ag_plan_params {
    uint32_t thread_max_deg;   /* in-degree <= this: AG_BIN_THREAD               */
    uint32_t group16_max_deg;  /* <= this: AG_BIN_GROUP16                       */
    uint32_t wave_max_deg;     /* <= this: AG_BIN_WAVE                          */
    uint32_t block_max_deg;    /* <= this: AG_BIN_BLOCK, above: AG_BIN_HUB      */
    uint32_t tile_rows;        /* max rows per thread-bin tile (<= 1024)        */
    uint32_t tile_nnz;         /* max in-edges per thread-bin tile (<= 4096)    */
    uint32_t segment_nnz;      /* in-edges per hub segment                      */
    uint32_t order;            /* row order inside a bin: 0 = vertex id, 1 = out-degree desc */
    uint32_t src_block_bits;   /* source blocking of the PageRank pull: sources are cut into
                                  blocks of 2^bits device ids and the in-edges of every row of
                                  a blocked bin are regrouped block by block, so the gathers of
                                  one pass stay inside one block (an L2-sized window of the
                                  contribution vector). AG_SRC_BLOCK_AUTO: 19 bits when the
                                  contribution vector exceeds 8 MB (n > 2^21), more if
                                  needed to stay within 256 blocks, else off;
                                  AG_SRC_BLOCK_OFF: off; else 8..30 (at most 256 blocks)   */
    uint32_t wg_per_cu;        /* resident workgroups per CU of the persistent grid (0 = auto:
                                  4, fewer if LDS limits; more in flight only queues at L2) */
    uint32_t blocked_min_bin;  /* with source blocking, rows of bins >= this are blocked.
                                  Default AG_BIN_THREAD (every row): on a blocked graph the
                                  thread bin then loses its stream-order FP32 sums (FP32
                                  results stay within rounding; integer results are exact) */
    uint32_t hot_segments;     /* with source blocking, at most this many "hot" source
                                  segments (2^min(14, bits) consecutive device ids whose
                                  in-edges average at least two per (row, segment) piece and
                                  four per source) become blocks of their own, gathered by FP32
                                  PageRank and FP32 SpMV from a copy of the segment in LDS
                                  instead of through L2 requests. Default 0 (none): on
                                  RMAT-25 128 of them made the FP32 iteration slower (the LDS
                                  pass is latency-bound at 2 workgroups per CU and doubles the
                                  partials the combine pass reads)                          */
    uint32_t fp32_order;       /* AG_ORDER_FAST (default): rows above the thread bin are
                                  summed in fixed tree / block orders. AG_ORDER_REFERENCE:
                                  every row is summed in sequence in inverse-CSR order, as
                                  the reference's CPU loop and float sum kernel do, so FP32
                                  results (and the active-count stop) equal the sequential
                                  oracle bit for bit. The plan then bins rows up to tile_nnz
                                  in-edges as thread-bin rows (one lane each, from LDS), every
                                  longer row is one hub segment added by one lane in chunks
                                  staged through LDS, and source blocking is off             */
    uint32_t tile_order;       /* order in which the workgroups of an XCD take the first-pass
                                  tiles of a blocked iteration (every row blocked).
                                  AG_TILES_AUTO (default): by ticket when the pieces number
                                  at least half the in-edges (sources spread evenly, as in an
                                  Erdos-Renyi graph), else round-robin. AG_TILES_ROUND_ROBIN:
                                  workgroup j takes tiles j, j + G, ... AG_TILES_TICKET: in
                                  order through a counter, so the tiles in flight stay
                                  consecutive (inside about one source block). The results
                                  are the same bit for bit in every order                  */
    uint32_t piece_nnz;        /* with source blocking, the most in-edges of one piece (the
                                  (row, source block) run a first-pass partial sums). 0 =
                                  auto: min(tile_nnz, 4096), or 1 when that leaves at least
                                  3/4 as many pieces as in-edges (then every in-edge is its
                                  own piece and the first pass is a plain gather); else
                                  1..tile_nnz                                                */
    uint32_t top_table;        /* with source blocking, an LDS copy of the contributions of
                                  the last 4096 device ids (the heaviest rows of the device
                                  order, which on skewed graphs are also the most gathered
                                  sources) in every first-pass workgroup of the FP32 PageRank
                                  kinds, read instead of L2 for those sources (RMAT-25: 12% of
                                  the edges, 2.9% faster; FP32 SpMV measured no gain). AG_TOP_AUTO (default): when they
                                  receive at least 2% of the blocked in-edges; AG_TOP_OFF;
                                  AG_TOP_ON                                                  */
    uint32_t tile_sort;        /* with source blocking of at most 2^20 sources per block and
                                  tiles of at most 4096 in-edges, the first pass reads each
                                  tile's source ids sorted by source, each packed with the
                                  edge's position in the tile (one 32-bit word, no extra
                                  bytes), so the lanes of a wavefront gather neighbouring
                                  sources and share more L2 lines; every term still lands at
                                  its position, so the sums are the same bit for bit. Not
                                  with hot segments. AG_SORT_AUTO (default), AG_SORT_OFF,
                                  AG_SORT_ON (the same while it is the default)              */
} ag_plan_params;

enum { AG_ORDER_FAST = 0, AG_ORDER_REFERENCE = 1 };
enum { AG_TILES_AUTO = 0, AG_TILES_ROUND_ROBIN = 1, AG_TILES_TICKET = 2 };
enum { AG_TOP_AUTO = 0, AG_TOP_OFF = 1, AG_TOP_ON = 2 };
enum { AG_SORT_AUTO = 0, AG_SORT_OFF = 1, AG_SORT_ON = 2 };

#define AG_SRC_BLOCK_AUTO 0u
#define AG_SRC_BLOCK_OFF 255u

void ag_plan_params_default(ag_plan_params *p);

typedef struct ag_graph_info {
    uint32_t num_vertices;
    uint32_t num_edges;
    uint32_t bin_rows[AG_NBINS];
    uint64_t bin_edges[AG_NBINS];
    uint32_t bin_blocks[AG_NBINS];
    uint32_t num_tiles;
    uint32_t num_segments;
    uint32_t has_weights;
    uint64_t device_bytes;
    ag_plan_params params;
    uint32_t src_block_bits;   /* blocking in effect (0: none)                          */
    uint32_t src_blocks;       /* source blocks                                         */
    uint32_t virtual_rows;     /* (row, source block) pieces of the blocked rows        */
    uint32_t blocked_row0;     /* first blocked device row                              */
    uint32_t hot_segments;     /* hot source segments in effect (blocks staged in LDS)      */
    uint32_t hot_segment_bits; /* sources per hot segment = 2^this                          */
    uint64_t hot_edges;        /* in-edges whose source lies in a hot segment               */
    uint32_t top_sources;      /* sources of the first-pass LDS table (top_table; 0: none)  */
    uint64_t top_edges;        /* blocked in-edges whose source lies in that table          */
    uint32_t sorted_tiles;     /* 1: first-pass tiles read sorted by source (tile_sort)     */
} ag_graph_info;

/* Upload one directed graph. Arrays are host arrays in the reference's layout:
 *   out_degree[n]      vertices->out_degree
 *   inv_edges_idx[n]   inverse_vertices->edges_idx
 *   inv_degree[n]      inverse_vertices->out_degree (in-degree)
 *   inv_src[m]         inverse_sorted_edges_array->edges_array_dest
 *   inv_weight[m]      inverse_sorted_edges_array->edges_array_weight, or NULL
 * params may be NULL (defaults). Validates every index on the host before upload. */
int ag_graph_upload(ag_device *dev, uint32_t n, uint32_t m, const uint32_t *out_degree,
                    const uint32_t *inv_edges_idx, const uint32_t *inv_degree,
                    const uint32_t *inv_src, const float *inv_weight,
                    const ag_plan_params *params, ag_dgraph **out);
int ag_graph_free(ag_dgraph *g);
/* Replace the fixed-point SpMV weights with caller-supplied Q16.16 codes (the reference
 * builds them on the host, 02_capi_graph/src/algorithms/capi/SPMV.c:920-926). wq[] is in the
 * caller's inverse-edge order, addressed through the same inv_edges_idx/inv_degree arrays
 * that were uploaded. */
int ag_graph_set_fixed_weights(ag_dgraph *g, const uint32_t *inv_edges_idx,
                               const uint32_t *inv_degree, const uint32_t *wq);
int ag_graph_get_info(const ag_dgraph *g, ag_graph_info *info);

/* Autotuner: builds each candidate plan for this graph, times `iterations` FP32 PageRank-pull
 * iterations (median of `repeats` runs, HIP events) and reports the fastest candidate. A plan
 * changes only the schedule -- results are identical for the integer paths and within FP32
 * rounding (tree order outside the thread bin) for the float paths. ms_out[count] may be NULL. */
int ag_autotune_plan(ag_device *dev, uint32_t n, uint32_t m, const uint32_t *out_degree,
                     const uint32_t *inv_edges_idx, const uint32_t *inv_degree,
                     const uint32_t *inv_src, const ag_plan_params *candidates, uint32_t count,
                     uint32_t iterations, uint32_t repeats, double *ms_out, uint32_t *best);

/* ---------------------------------------------------------------- PageRank pull */

enum { AG_STOP_ACTIVE = 0,   /* reference rule: stop when no vertex has |delta| >= epsilon */
       AG_STOP_L1 = 1,       /* stop at the first iteration with error_total <= l1_tolerance */
       AG_STOP_NONE = 2 };   /* run max_iterations */

enum { AG_INIT_REFERENCE = 0,  /* every scaled rank 1 - Damp (pageRank.c:57-58,71) */
       AG_INIT_UNIFORM = 1,    /* every scaled rank 1.0, i.e. 1/N after normalisation */
       AG_INIT_USER = 2 };     /* init_ranks[], scaled units, original vertex order */

enum { AG_KIND_F32 = 0, AG_KIND_FIXED64 = 64, AG_KIND_FIXED32 = 32, AG_KIND_FIXED16 = 16,
       AG_KIND_FIXED8 = 8,
       /* quantised pulls (pageRankPullQuant{32,16,8}BitGraphCSR, pageRank.c:2119-2676): per
        * iteration the f32 contributions are quantised with scale = max / (2^B - 1) and zero
        * point 0 into B-bit codes, the codes are summed exactly in u64 and the update is
        * base + damp * (scale * (float)sum). The quantiser is this library's definition
        * (the reference's quantization.h is not in its tree). */
       AG_KIND_QUANT32 = 132, AG_KIND_QUANT16 = 116, AG_KIND_QUANT8 = 108 };

enum { AG_STORE_F32 = 0, AG_STORE_BF16 = 1, AG_STORE_F16 = 2 };

typedef struct ag_pr_options {
    uint32_t kind;             /* AG_KIND_*                                                */
    uint32_t max_iterations;
    double epsilon;            /* per-vertex active threshold, scaled units               */
    uint32_t stop_rule;        /* AG_STOP_*                                               */
    double l1_tolerance;
    uint32_t init;             /* AG_INIT_*                                               */
    const float *init_ranks;   /* host [n] for AG_INIT_USER                               */
    uint32_t storage;          /* AG_KIND_F32 only: contribution storage AG_STORE_*       */
    double refine_l1;          /* 16-bit storage: switch to f32 once error_total <= this  */
    uint32_t chunk;            /* iterations enqueued between convergence polls (0 = auto:
                                  32 up to 2^24 edges, else 8)                             */
    uint32_t record_history;   /* keep per-iteration active/error/time                    */
    uint32_t direction;        /* AG_DIR_*                                                */
} ag_pr_options;

/* Direction of the sum (pageRankGraphCSR's pushpull codes, pageRank.c:874-939):
 *   PULL         gather over in-edges (codes 0, 2, 4, 9-13)
 *   PUSH         scatter every contribution along the out-edges, then update every vertex
 *                (codes 1 and 3, pageRank.c:1111-1272, 1992-2117), propagation-blocked: each
 *                out-edge writes its source's contribution to a slot of its destination's
 *                bin, and each bin's slots are added per row in LDS; AG_KIND_F32 with f32
 *                storage or a fixed-point kind. FP32 sums take the LDS atomics' order;
 *                fixed-point sums are exact and equal the pull's
 *   DATA_DRIVEN  pull over a worklist (code 6, pageRank.c:2810-2969): only listed vertices
 *                gather; a vertex whose |delta| >= epsilon takes its new rank and lists its
 *                out-neighbours for the next iteration, others keep their rank; AG_KIND_F32
 *                with f32 storage only. Unblocked plans keep explicit worklists (sums in the
 *                reference's edge order); source-blocked plans run it as a full pull that
 *                carries the lists in the contributions' sign bits (same result, pull order)
 * PUSH, and DATA_DRIVEN on an unblocked plan, build the graph's forward CSR on the device on
 * first use. */
enum { AG_DIR_PULL = 0, AG_DIR_PUSH = 1, AG_DIR_DATA_DRIVEN = 2 };

void ag_pr_options_default(ag_pr_options *o);

typedef struct ag_pr_result {
    uint32_t iterations;        /* reference meaning: index of the stopping iteration, or
                                   max_iterations when the loop ran out (pageRank.c:1021) */
    uint32_t active_vertices;   /* active count of the last executed iteration            */
    double error_total;         /* L1 delta of the last executed iteration, normalised    */
    double rank_sum;            /* FP64 sum of the normalised ranks                        */
    double loop_ms;             /* device time of the iteration loop (HIP events)          */
    double total_ms;            /* host wall time of the whole call                        */
    uint32_t executed;          /* iterations actually executed (= iterations + 1 on stop) */
    int32_t refine_iteration;   /* first iteration that wrote f32 contributions, or -1     */
} ag_pr_result;

/* One-call run (reference semantics when opts is the default): ranks_out[n] receives the
 * normalised ranks (scaled / N) in original vertex order. */
int ag_pagerank_pull(ag_dgraph *g, const ag_pr_options *opts, float *ranks_out,
                     ag_pr_result *res);

/* Stepping interface used by drivers and benchmarks. */
int ag_pr_begin(ag_dgraph *g, const ag_pr_options *opts, ag_pr_session **out);
int ag_pr_step(ag_pr_session *s, uint32_t iterations);     /* enqueue, asynchronous */
int ag_pr_poll(ag_pr_session *s, ag_pr_result *res);       /* wait and read state    */
int ag_pr_end(ag_pr_session *s, float *ranks_out, ag_pr_result *res); /* ranks may be NULL */
/* Per-iteration history of the last session run with record_history (up to cap rows). */
int ag_pr_history(ag_pr_session *s, uint32_t cap, uint32_t *active, double *error,
                  float *iter_ms, uint32_t *rows);
/* Scaled (un-normalised) ranks in original order, without ending the session. */
int ag_pr_read_scaled(ag_pr_session *s, float *scaled_out);
/* Average duration (ms) of the iteration kernel over the last ag_pr_step call, measured
 * with HIP events on the session's stream. */
int ag_pr_last_step_kernel_ms(ag_pr_session *s, double *avg_ms, uint32_t *launches);

/* ---------------------------------------------------------------- SpMV pull */

/* y[v] = sum over in-edges (u->v) of fp32(w*x[u]), accumulated in fp32; w = 0.0001f when
 * the graph has no weights. Repeats the product `repetitions` times (each overwrites y). */
int ag_spmv_pull(ag_dgraph *g, const float *x, float *y, uint32_t repetitions,
                 double *avg_kernel_ms);
/* ag_spmv_pull without any allocation (no counterpart in the reference; for callers that
 * cannot allocate device memory: serving loops, code under stream capture, harnesses that
 * forbid it). One product y = A x, inputs and outputs in host memory as for ag_spmv_pull;
 * the permuted copies of x and y live in a fixed device array of AG_NOALLOC_SCRATCH_BYTES
 * per GPU (taken at ag_device_open, held until the process ends), so the graph's vertex
 * count must satisfy 8 * num_vertices <= AG_NOALLOC_SCRATCH_BYTES (4,194,304 vertices);
 * larger graphs return AG_ERR_INVALID. Every call on one GPU shares that array: calls that
 * overlap on one GPU (several host threads) race. Nothing in the array needs clearing (x and
 * y are written whole on every call). Waits like every other call (the watchdog applies). */
#define AG_NOALLOC_SCRATCH_BYTES (32u << 20)
int ag_spmv_pull_noalloc(ag_dgraph *g, const float *x, float *y);
/* Fixed point: y_q[v] = sum of (u32)(((u64)x_q[u] * w_q) >> 16) modulo 2^64, with
 * w_q = FloatToFixed32(w) per edge, or the codes given to ag_graph_set_fixed_weights. */
int ag_spmv_pull_fixed(ag_dgraph *g, const uint32_t *x_q, uint64_t *y_q,
                       uint32_t repetitions, double *avg_kernel_ms);
/* Push forms (SPMVPushGraphCSR / SPMVPushFixedPointGraphCSR, SPMV.c:752-848, 1017-1136):
 * every source adds w*x[src] to its out-neighbours (propagation-blocked, as the PageRank push).
 * As in the reference the output is not cleared between repetitions, so y = the sum of
 * `repetitions` products (0 repetitions: y = 0). FP32: binary32 products, each repetition's
 * row sum added in the LDS atomics' order, then added to y. Fixed point: x_q and y_q in Q32.32
 * (DoubleToFixed64), weights DoubleToFixed64(w) (0.0001f unweighted), products MULFixed64V1,
 * exact u64 sums. The forward CSR and the push bins are built on the device on first use. */
int ag_spmv_push(ag_dgraph *g, const float *x, float *y, uint32_t repetitions,
                 double *avg_kernel_ms);
int ag_spmv_push_fixed(ag_dgraph *g, const uint64_t *x_q, uint64_t *y_q, uint32_t repetitions,
                       double *avg_kernel_ms);

/* ---------------------------------------------------------------- personalised PageRank */

/* Up to 64 PageRank columns pulled together: column c restarts at seeds[c] (an original vertex
 * id; AG_PPR_UNIFORM restarts uniformly, which is plain PageRank). Scaled units as the plain
 * path: the restart mass is (1 - Damp) * N, all of it on the seed; every column starts at 1.0.
 * Contributions are stored vertex-major, 64 per vertex (one 128-byte line in BF16/FP16), and
 * accumulated in FP32; with 16-bit storage the last iterations switch to FP32 contributions
 * (refine_l1, as for ag_pr_options). The L1 delta of each column is reduced in FP64. */
#define AG_PPR_MAX_COLS 64
#define AG_PPR_UNIFORM 0xFFFFFFFFu

typedef struct ag_ppr_options {
    uint32_t max_iterations;
    uint32_t stop_rule;        /* AG_STOP_L1 (every column <= l1_tolerance) or AG_STOP_NONE */
    double l1_tolerance;
    uint32_t storage;          /* AG_STORE_*                                                */
    double refine_l1;
    uint32_t chunk;            /* eager mode: iterations queued between polls (0 = auto)   */
    uint32_t use_hip_graph;    /* 1 (default): the whole solve is one captured hipGraph,
                                  replayed with one launch and one wait (captured once per
                                  graph, max_iterations and grid; the buffers stay with the
                                  graph)                                                   */
    uint32_t wg_per_cu;        /* resident workgroups per CU of the iteration kernel (0 =
                                  as many as fit, up to 8)                                  */
} ag_ppr_options;

typedef struct ag_ppr_result {
    uint32_t iterations;       /* index of the stopping iteration, or max_iterations        */
    uint32_t executed;
    double max_error;          /* largest column L1 delta of the last executed iteration    */
    double loop_ms;            /* device time of the iteration loop                         */
    int32_t refine_iteration;
} ag_ppr_result;

void ag_ppr_options_default(ag_ppr_options *o);

/* ranks_out[v * k + c]: normalised rank of vertex v (original order) in column c.
 * col_error[c] (may be NULL): last L1 delta of column c. */
int ag_ppr_pull(ag_dgraph *g, uint32_t k, const uint32_t *seeds, const ag_ppr_options *opts,
                float *ranks_out, double *col_error, ag_ppr_result *res);

/* ---------------------------------------------------------------- learned warm start */

/* A 3-layer MLP (5 inputs -> 64 -> 64 -> 1, ReLU) that predicts, for each vertex, the log
 * ratio of its converged scaled rank to its scaled rank after two iterations, from five
 * features: log(1 + in-degree), log(1 + out-degree), log(1 + mean in-degree of the
 * in-neighbours), and the log scaled ranks after one and two FP32 PageRank iterations from the
 * uniform start. Weights are kept in FP32 and fed to BF16 MFMA (FP32 accumulate); training is
 * on the GPU with the library's own forward/backward kernels and Adam. Parameter layout (FP32):
 * W1[16][64] (rows 5..15 unused), b1[64], W2[64][64], b2[64], w3[64], b3. */
#define AG_MLP_FEATURES 5
#define AG_MLP_HIDDEN 64

typedef struct ag_mlp ag_mlp;

uint32_t ag_mlp_param_count(void);
int ag_mlp_create(ag_device *dev, uint32_t seed, ag_mlp **out);      /* He-uniform init */
int ag_mlp_free(ag_mlp *m);
int ag_mlp_set_params(ag_mlp *m, const float *params);
int ag_mlp_get_params(ag_mlp *m, float *params);
/* Inputs are standardised as (x - mean) / std and the target as (y - y_mean) / y_std. */
int ag_mlp_set_normalization(ag_mlp *m, const float *in_mean, const float *in_std,
                             float y_mean, float y_std);
/* Mean-squared loss (standardised units) and its gradient on S samples, no update. */
int ag_mlp_loss_grad(ag_mlp *m, const float *X, const float *y, uint32_t S, double *loss,
                     float *grad);
/* Adam training: epochs over S samples in minibatches of `batch`, shuffled by `seed`;
 * loss_per_epoch[epochs] (may be NULL) receives each epoch's mean minibatch loss. */
int ag_mlp_train(ag_mlp *m, const float *X, const float *y, uint32_t S, uint32_t epochs,
                 uint32_t batch, float lr, uint32_t seed, float *loss_per_epoch);
/* Predictions in target units for S samples. */
int ag_mlp_predict(ag_mlp *m, const float *X, uint32_t S, float *y_out);
/* The five features of every vertex of an uploaded graph, original order, X_out[n][5]. */
int ag_mlp_features(ag_dgraph *g, float *X_out);

typedef struct ag_warm_result {
    uint32_t fallback;      /* 1 when the predicted start was rejected for the uniform one */
    double uniform_l1;      /* L1 delta of the second iteration from the uniform start      */
    double warm_l1;         /* L1 delta of one iteration from the predicted start          */
    double ms;              /* device time of features + inference + check                 */
} ag_warm_result;

/* Warm start of a fresh FP32 session begun with AG_INIT_UNIFORM: runs two iterations
 * (features), infers the start vector r2 * exp(prediction), rescales it so its total mass
 * meets the fixed-point balance (restart mass (1 - Damp) N, dangling mass dropped), runs one
 * iteration from it and keeps it only if its L1 delta beats the uniform start's second delta
 * (else continues from the uniform run). Those iterations count in the session. */
int ag_pr_warm_start(ag_pr_session *s, ag_mlp *m, ag_warm_result *res);

/* ---------------------------------------------------------------- graph construction */

/* The step before the path, on the GPU, with the same results as the host builders of
 * include/graphCSR.h array for array (the reference takes its graphs from OpenGraph's
 * builder, absent; options at 02_capi_graph/src/main/accel-graph.c:71-72,143-145):
 * the chunked MT19937 RMAT / Erdos-Renyi edge streams of graphGenerateRMATEdges /
 * graphCSRGenerateER, and the CSR + inverse CSR of graphCSRFromEdgeList (rows by source with
 * destinations ascending, self-loops and duplicates dropped when `simplify`, the first
 * occurrence keeping its weight; inverse rows by destination with sources ascending). The
 * struct GraphCSR is allocated on the host (release with graphCSRFree). */
struct GraphCSR;
int ag_generate_rmat_edges(ag_device *dev, uint32_t scale, uint64_t num_edges, double a,
                           double b, double c, uint32_t seed, uint32_t *src_out,
                           uint32_t *dst_out);
int ag_generate_er_edges(ag_device *dev, uint32_t num_vertices, uint64_t num_edges,
                         uint32_t seed, uint32_t *src_out, uint32_t *dst_out);
/* num_vertices = 0: 1 + the largest id. AG_ERR_INVALID for an id >= num_vertices. */
int ag_graph_csr_build(ag_device *dev, uint32_t num_vertices, uint64_t num_edges,
                       const uint32_t *src, const uint32_t *dst, const float *weight,
                       int simplify, struct GraphCSR **out);
int ag_graph_csr_generate_rmat(ag_device *dev, uint32_t scale, uint32_t edge_factor, double a,
                               double b, double c, uint32_t seed, int simplify,
                               struct GraphCSR **out);
int ag_graph_csr_generate_er(ag_device *dev, uint32_t num_vertices, uint64_t num_edges,
                             uint32_t seed, int simplify, struct GraphCSR **out);

/* ---------------------------------------------------------------- many small graphs */

/* A batch of independent graphs resident in HBM, solved together: every graph is cut into
 * parts (runs of rows with a bounded in-edge count), one launch per iteration runs every part
 * of every graph still iterating and a second adds each graph's part partials and applies its
 * stop rule; the whole solve is captured once into a hipGraph and replayed as one launch.
 * Inputs are the per-graph reference arrays concatenated: graph g owns vertices
 * [vert_off[g], vert_off[g+1]) and in-edges [edge_off[g], edge_off[g+1]); inv_edges_idx is
 * relative to the graph's first edge and inv_src holds graph-local vertex ids. */
typedef struct ag_batch ag_batch;

int ag_batch_upload(ag_device *dev, uint32_t num_graphs, const uint32_t *vert_off,
                    const uint32_t *edge_off, const uint32_t *out_degree,
                    const uint32_t *inv_edges_idx, const uint32_t *inv_degree,
                    const uint32_t *inv_src, ag_batch **out);
int ag_batch_free(ag_batch *b);

typedef struct ag_batch_options {
    uint32_t max_iterations;
    uint32_t stop_rule;        /* AG_STOP_ACTIVE or AG_STOP_L1 (per graph)              */
    double epsilon;
    double l1_tolerance;
    uint32_t init;             /* AG_INIT_REFERENCE or AG_INIT_UNIFORM                   */
    uint32_t use_hip_graph;    /* capture the solve once and replay it                   */
    uint32_t part_edges;       /* in-edges per part (0 = AG_BATCH_PART_EDGES); a row with
                                  more is a part by itself                                */
    uint32_t heavy_deg;        /* rows with more in-edges are summed by a wavefront in a
                                  fixed tree, the others by one lane each in the reference's
                                  stream order (0 = AG_BATCH_HEAVY_DEG)                    */
} ag_batch_options;

#define AG_BATCH_HEAVY_DEG 32
#define AG_BATCH_PART_EDGES 4096

void ag_batch_options_default(ag_batch_options *o);

/* ranks_out[vert_off[G]] normalised per graph (scaled / n_g); iterations[g] in the reference's
 * meaning; error[g] = last L1 delta; device_ms = time of the launch (HIP events). */
int ag_batch_pagerank(ag_batch *b, const ag_batch_options *opts, float *ranks_out,
                      uint32_t *iterations, double *error, double *device_ms);

/* ---------------------------------------------------------------- raw gather */

/* The bare pull gather (the FPGA sum kernel's contract): out[v] = sum over in-edges
 * (u->v) of values[u], values and out in original vertex order.
 *   AG_GATHER_F32  float -> float (fp32 adds)
 *   AG_GATHER_U64  u64 -> u64 modulo 2^64, AG_GATHER_U32/U16/U8 zero-extended -> u64 */
enum { AG_GATHER_F32 = 0, AG_GATHER_U64 = 1, AG_GATHER_U32 = 2, AG_GATHER_U16 = 3,
       AG_GATHER_U8 = 4 };
int ag_pull_gather(ag_dgraph *g, uint32_t type, const void *values, void *out);

/* Device-side fixed-point conversions (the same functions the kernels use), for tests. */
int ag_fixed_encode(ag_device *dev, uint32_t bits, const float *in, uint64_t *out, uint32_t n);
int ag_fixed_decode(ag_device *dev, uint32_t bits, const uint64_t *in, double *out, uint32_t n);

#ifdef __cplusplus
}
#endif

#endif
