/*
 * ag_oracle.c -- CPU restatement of the reference's PageRank-pull / SpMV-pull path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's CPU leg (the
 * CPU baseline timings and the FP64 reference its error figures are measured against) load
 * this library, and only as the checker or the CPU baseline; the product library
 * (accelgraph_amd/libaccelgraph.so) never links or calls it.
 *
 * Every function follows the reference step by step:
 *   or_pagerank_pull   02_capi_graph/src/algorithms/capi/pageRank.c:942-1109 (FP32) and
 *                      :1276-1990 (fixed-point 64/32/16/8): init rank = 1 - 0.85f (:50-77),
 *                      contribution rank/out_degree or 0 (:1026-1033), gather in inverse-CSR
 *                      stream order (:2890-2907; RTL golden pagerank_kernel_tb.sv:303-314),
 *                      update base + damp*next, err += |delta|/N, active += |delta| >= eps
 *                      (:1055-1069), stop on active == 0 (:1072-1074), ranks /= N and an FP64
 *                      sum (:1079-1085).
 *   or_pagerank_push   the push directions, pageRank.c:1111-1272 and :1992-2117.
 *   or_pagerank_dd_pull  the data-driven (worklist) pull, pageRank.c:2810-2969.
 *   or_pagerank_f64    the same recurrence in FP64, iterated to a tight tolerance: the
 *                      double-precision reference every FP32 result is measured against.
 *   or_spmv_f32/fixed  SPMV.c:622-751 / :849-1016 with the FPGA datapath arithmetic
 *                      (FloatPoint cu_sum_kernel_control.sv:345-367: fp32 product then stream
 *                      add; FixedPoint :155-164: (u32)((x*w) >> 16) into a u64 sum).
 *   or_spmv_push_*     the SpMV push directions, SPMV.c:752-848 / :1017-1136.
 *   or_gather_*        the bare sum kernel (cu_sum_kernel_control.sv), float and integer.
 *   or_build_csr       an independent sort-based CSR + inverse-CSR builder (the reference's
 *                      builder lives in the absent OpenGraph submodule).
 *   or_quantize        the quantised pulls' per-iteration scale and codes (pageRank.c:2119-2676;
 *                      or_pagerank_pull kinds 132/116/108 run those pulls end to end).
 *   or_fixed_*         the fixed-point formats, computed with integer arithmetic on the float's
 *                      mantissa/exponent rather than with floating-point scaling.
 *   or_ppr_residual    the FP64 residual p + 0.85 * A r - r of given (personalised) rank
 *                      columns: the defining equation of the fixed point of the update at
 *                      pageRank.c:1055-1069 (dangling mass dropped, :1029-1032). Since
 *                      ||0.85 A||_1 <= 0.85, ||r - r*||_1 <= ||residual||_1 / 0.15 bounds the
 *                      distance of r to the FP64 solution without computing it.
 *
 * Threads: `threads` = 0 runs the sequential restatement; > 0 runs the same loops under
 * OpenMP with schedule(dynamic,1024) (pageRank.c:2890), the CPU baseline. Per-vertex values
 * are identical in both modes (each row is still summed sequentially); only the order of the
 * FP64 error reduction differs.
 *
 * Parity status: or_fixed_* pins the library's own format definitions (the reference's
 * fixedPoint.h is absent; its conversions are "parity unpinned" against upstream OpenGraph).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_DAMP 0.85f

/* ------------------------------------------------------------------ fixed point */

/* floor(|x| * 2^frac + 1/2) for finite x > 0, exactly, via the float's integer mantissa.
 * Saturates at maxcode; NaN, zero and negatives give 0. */
static uint64_t or_round_half_up_scaled(double x, int frac, uint64_t maxcode)
{
    if (!(x > 0.0))
        return 0;
    if (isinf(x))
        return maxcode;
    int e2;
    double mant = frexp(x, &e2);            /* x = mant * 2^e2, mant in [0.5, 1) */
    uint64_t M = (uint64_t)ldexp(mant, 53); /* exact: 53-bit integer mantissa */
    int shift = e2 - 53 + frac;             /* x * 2^frac = M * 2^shift */
    if (shift >= 0) {
        if (shift > 11)
            return maxcode;
        uint64_t v = M << shift;
        return v > maxcode ? maxcode : v;
    }
    int rs = -shift;
    if (rs > 64)
        return 0;
    /* floor(M / 2^rs + 1/2) = floor((M + 2^(rs-1)) / 2^rs) */
    unsigned __int128 num = (unsigned __int128)M + ((unsigned __int128)1 << (rs - 1));
    unsigned __int128 q = num >> rs;
    return q > maxcode ? maxcode : (uint64_t)q;
}

uint64_t or_fixed_encode(uint32_t bits, float x)
{
    switch (bits) {
    case 64: return or_round_half_up_scaled((double)x, 32, UINT64_MAX);
    case 32: return or_round_half_up_scaled((double)x, 16, UINT32_MAX);
    case 16: return or_round_half_up_scaled((double)x, 8, UINT16_MAX);
    default: return or_round_half_up_scaled((double)x, 4, UINT8_MAX);
    }
}

/* q * 2^-frac: the u64 -> float/double conversion is round-to-nearest-even (C semantics on
 * x86-64), the power-of-two scale is exact */
double or_fixed_decode(uint32_t bits, uint64_t q)
{
    switch (bits) {
    case 64: return ldexp((double)q, -32);
    case 32: return (double)ldexpf((float)q, -16);
    case 16: return (double)ldexpf((float)q, -8);
    default: return (double)ldexpf((float)q, -4);
    }
}

void or_fixed_encode_n(uint32_t bits, const float *x, uint64_t n, uint64_t *out)
{
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; i++)
        out[i] = or_fixed_encode(bits, x[i]);
}

/* ------------------------------------------------------------------ gathers */

void or_gather_f32(uint32_t n, const uint32_t *idx, const uint32_t *deg, const uint32_t *src,
                   const float *vals, float *out)
{
    for (uint32_t v = 0; v < n; v++) {
        float acc = 0.0f;
        for (uint32_t k = 0; k < deg[v]; k++) {
            float x = vals[src[idx[v] + k]];
            acc = (k == 0) ? x : acc + x;
        }
        out[v] = acc;
    }
}

void or_gather_u64(uint32_t n, const uint32_t *idx, const uint32_t *deg, const uint32_t *src,
                   const uint64_t *vals, uint64_t *out)
{
    for (uint32_t v = 0; v < n; v++) {
        uint64_t acc = 0;
        for (uint32_t k = 0; k < deg[v]; k++)
            acc += vals[src[idx[v] + k]];
        out[v] = acc;
    }
}

/* ------------------------------------------------------------------ SpMV */

void or_spmv_f32(uint32_t n, const uint32_t *idx, const uint32_t *deg, const uint32_t *src,
                 const float *w, const float *x, float *y, int threads)
{
    #pragma omp parallel for schedule(dynamic, 1024) if (threads > 0) num_threads(threads > 0 ? threads : 1)
    for (uint32_t v = 0; v < n; v++) {
        float acc = 0.0f;
        for (uint32_t k = 0; k < deg[v]; k++) {
            uint32_t j = idx[v] + k;
            float p = x[src[j]] * (w ? w[j] : 0.0001f);
            acc = (k == 0) ? p : acc + p;
        }
        y[v] = acc;
    }
}

void or_spmv_fixed(uint32_t n, const uint32_t *idx, const uint32_t *deg, const uint32_t *src,
                   const uint32_t *wq, uint32_t wconst, const uint32_t *xq, uint64_t *yq,
                   int threads)
{
    #pragma omp parallel for schedule(dynamic, 1024) if (threads > 0) num_threads(threads > 0 ? threads : 1)
    for (uint32_t v = 0; v < n; v++) {
        uint64_t acc = 0;
        for (uint32_t k = 0; k < deg[v]; k++) {
            uint32_t j = idx[v] + k;
            uint64_t wide = (uint64_t)xq[src[j]] * (uint64_t)(wq ? wq[j] : wconst);
            acc += (uint64_t)(uint32_t)(wide >> 16);
        }
        yq[v] = acc;
    }
}

/* SPMVPushGraphCSR / SPMVPushFixedPointGraphCSR (SPMV.c:752-848, 1017-1136): every source v
 * adds w * x[v] to its out-neighbours with `omp atomic` (:815-816, :1091-1092) along the forward
 * CSR (fwd_idx/fwd_deg/fwd_dst, weights fw in the same order or NULL for 0.0001f); y is never
 * cleared between iterations, so `reps` products accumulate onto y = 0. Fixed point: x and w
 * in Q32.32 (DoubleToFixed64), products MULFixed64V1 = the 128-bit product >> 32 (the
 * library's definition: OpenGraph's is absent), exact u64 sums. Sequentially each y[u] adds
 * its terms in ascending source order, the inverse CSR's order. */
static uint64_t or_mul_fixed64(uint64_t a, uint64_t b)
{
    return (uint64_t)(((unsigned __int128)a * b) >> 32);
}

void or_spmv_push_f32(uint32_t n, const uint32_t *fwd_idx, const uint32_t *fwd_deg,
                      const uint32_t *fwd_dst, const float *fw, const float *x, uint32_t reps,
                      int threads, float *y)
{
    memset(y, 0, (size_t)n * sizeof(float));
    for (uint32_t r = 0; r < reps; r++) {
        #pragma omp parallel for schedule(dynamic, 1024) if (threads > 0) num_threads(threads > 0 ? threads : 1)
        for (uint32_t v = 0; v < n; v++) {
            for (uint32_t k = 0; k < fwd_deg[v]; k++) {
                const uint32_t j = fwd_idx[v] + k;
                const float t = (fw ? fw[j] : 0.0001f) * x[v];
                #pragma omp atomic
                y[fwd_dst[j]] += t;
            }
        }
    }
}

void or_spmv_push_fixed(uint32_t n, const uint32_t *fwd_idx, const uint32_t *fwd_deg,
                        const uint32_t *fwd_dst, const float *fw, const uint64_t *xq,
                        uint32_t reps, int threads, uint64_t *y)
{
    const uint64_t wc = or_fixed_encode(64, 0.0001f);
    memset(y, 0, (size_t)n * sizeof(uint64_t));
    for (uint32_t r = 0; r < reps; r++) {
        #pragma omp parallel for schedule(dynamic, 1024) if (threads > 0) num_threads(threads > 0 ? threads : 1)
        for (uint32_t v = 0; v < n; v++) {
            for (uint32_t k = 0; k < fwd_deg[v]; k++) {
                const uint32_t j = fwd_idx[v] + k;
                const uint64_t t = or_mul_fixed64(fw ? or_fixed_encode(64, fw[j]) : wc, xq[v]);
                #pragma omp atomic
                y[fwd_dst[j]] += t;
            }
        }
    }
}

uint64_t or_mul_fixed64_check(uint64_t a, uint64_t b) { return or_mul_fixed64(a, b); }

/* ------------------------------------------------------------------ PageRank */

enum { OR_STOP_ACTIVE = 0, OR_STOP_L1 = 1, OR_STOP_NONE = 2 };
enum { OR_INIT_REFERENCE = 0, OR_INIT_UNIFORM = 1, OR_INIT_USER = 2 };

typedef struct or_pr_out {
    uint32_t iterations;
    uint32_t executed;
    uint32_t active;
    double error_total;
    double rank_sum;
} or_pr_out;

/* ------------------------------------------------------------------ quantised contributions
 *
 * pageRankPullQuant{32,16,8}BitGraphCSR (pageRank.c:2119-2303, 2305-2489, 2491-2676), per
 * iteration: f32 contributions rank/out_degree (:2203-2210), their min and max and a scale
 * (getMinMax_32 / GetScale_32, zero point 0: :2213-2216), codes quantize_32(c, scale, 0)
 * (:2221-2226), an integer gather of the codes into u64 sums on the FPGA (:2234-2239), and
 * the update base + damp * (scale * sum) in float (:2250-2253). quantization.h lives in the
 * absent OpenGraph submodule, so the three quantiser functions are this library's own:
 *   scale(B)  = mx > 0 ? (float)((double)mx / (2^B - 1)) : 1, mx = largest finite c > 0
 *               (with the zero point fixed at 0 the code range starts at 0, so the range
 *               that must fit is [0, mx]; the minimum takes no part);
 *   code(c)   = 0 unless c > 0, else min(2^B - 1, floor((double)c * inv + 0.5)) with
 *               inv = 1.0 / (double)scale (one double product and one double sum, rounded
 *               to nearest each).
 * Parity of these three against upstream OpenGraph is unpinned (its source is absent). */
#define OR_QUANT_BASE 100u

static float or_quant_scale(float mx, uint32_t bits)
{
    const double maxcode = (double)((1ull << bits) - 1ull);
    return mx > 0.0f ? (float)((double)mx / maxcode) : 1.0f;
}

static uint64_t or_quant_code(float c, double inv, uint32_t bits)
{
    const double maxcode = (double)((1ull << bits) - 1ull);
    if (!(c > 0.0f))
        return 0;
    const double t = floor((double)c * inv + 0.5);
    return t >= maxcode ? (uint64_t)maxcode : (uint64_t)t;
}

/* scale and codes of one iteration's f32 contributions; returns the scale */
float or_quantize(uint32_t n, const float *c, uint32_t bits, uint64_t *codes)
{
    float mx = 0.0f;
    for (uint32_t v = 0; v < n; v++)
        if (c[v] > mx && !isinf(c[v]))
            mx = c[v];
    const float scale = or_quant_scale(mx, bits);
    const double inv = 1.0 / (double)scale;
    for (uint32_t v = 0; v < n; v++)
        codes[v] = or_quant_code(c[v], inv, bits);
    return scale;
}

/* contribution of one vertex in the representation of `kind` (0 = f32, else fixed width) */
static inline void contrib_store(uint32_t kind, void *c, uint32_t v, float rank, uint32_t od)
{
    float q = od ? rank / (float)od : 0.0f;
    switch (kind) {
    case 0: ((float *)c)[v] = q; break;
    case 64: ((uint64_t *)c)[v] = od ? or_fixed_encode(64, q) : 0; break;
    case 32: ((uint32_t *)c)[v] = od ? (uint32_t)or_fixed_encode(32, q) : 0; break;
    case 16: ((uint16_t *)c)[v] = od ? (uint16_t)or_fixed_encode(16, q) : 0; break;
    default: ((uint8_t *)c)[v] = od ? (uint8_t)or_fixed_encode(8, q) : 0; break;
    }
}

static inline uint64_t contrib_load_q(uint32_t kind, const void *c, uint32_t u)
{
    switch (kind) {
    case 64: return ((const uint64_t *)c)[u];
    case 32: return ((const uint32_t *)c)[u];
    case 16: return ((const uint16_t *)c)[u];
    default: return ((const uint8_t *)c)[u];
    }
}

/* next rank from the gathered sum, in the reference's order of operations */
static inline float update_rank(uint32_t kind, float base, float damp, float fsum, uint64_t qsum)
{
    switch (kind) {
    case 0: return base + damp * fsum;
    case 64: return (float)((double)base + (double)damp * or_fixed_decode(64, qsum));
    default: return base + damp * (float)or_fixed_decode(kind, qsum);
    }
}

int or_pagerank_pull(uint32_t n, const uint32_t *outdeg, const uint32_t *idx,
                     const uint32_t *deg, const uint32_t *src, uint32_t kind,
                     uint32_t max_iter, double eps, uint32_t stop_rule, double l1_tol,
                     uint32_t init, const float *init_ranks, int threads, float *ranks_out,
                     float *scaled_out, uint32_t *hist_active, double *hist_err,
                     or_pr_out *out)
{
    const int quant = kind == OR_QUANT_BASE + 32 || kind == OR_QUANT_BASE + 16 ||
                      kind == OR_QUANT_BASE + 8;
    if (kind != 0 && kind != 64 && kind != 32 && kind != 16 && kind != 8 && !quant)
        return -1;
    const uint32_t qbits = quant ? kind - OR_QUANT_BASE : 0;
    const float damp = OR_DAMP;
    const float base = 1.0f - damp;
    /* quantised kinds: f32 contributions, then u64 codes (any width fits) */
    size_t csize = (kind == 0 || quant) ? 4 : kind / 8;
    float *rank = (float *)malloc((size_t)(n ? n : 1) * sizeof(float));
    void *contrib = malloc((size_t)(n ? n : 1) * csize);
    uint64_t *codes = quant ? (uint64_t *)malloc((size_t)(n ? n : 1) * sizeof(uint64_t)) : NULL;
    if (!rank || !contrib || (quant && !codes)) {
        free(rank);
        free(contrib);
        free(codes);
        return -2;
    }
    for (uint32_t v = 0; v < n; v++)
        rank[v] = init == OR_INIT_REFERENCE ? base : init == OR_INIT_UNIFORM ? 1.0f : init_ranks[v];
    memset(out, 0, sizeof(*out));
    uint32_t it;
    double error_total = 0.0;
    uint32_t active = 0;
    int stopped = 0;
    const int nt = threads > 0 ? threads : 1;
    for (it = 0; it < max_iter; it++) {
        error_total = 0.0;
        active = 0;
        #pragma omp parallel for schedule(static) if (threads > 0) num_threads(nt)
        for (uint32_t v = 0; v < n; v++)
            contrib_store(quant ? 0 : kind, contrib, v, rank[v], outdeg[v]);
        const float qscale = quant ? or_quantize(n, (const float *)contrib, qbits, codes) : 0.0f;
        #pragma omp parallel for schedule(dynamic, 1024) if (threads > 0) num_threads(nt) reduction(+ : error_total, active)
        for (uint32_t v = 0; v < n; v++) {
            float fsum = 0.0f;
            uint64_t qsum = 0;
            const uint32_t b = idx[v], d = deg[v];
            if (kind == 0) {
                const float *c = (const float *)contrib;
                for (uint32_t k = 0; k < d; k++) {
                    float x = c[src[b + k]];
                    fsum = (k == 0) ? x : fsum + x;
                }
            } else if (quant) {
                for (uint32_t k = 0; k < d; k++)
                    qsum += codes[src[b + k]];
            } else {
                for (uint32_t k = 0; k < d; k++)
                    qsum += contrib_load_q(kind, contrib, src[b + k]);
            }
            float prev = rank[v];
            float next = quant ? base + damp * (qscale * (float)qsum)
                               : update_rank(kind, base, damp, fsum, qsum);
            rank[v] = next;
            double error = fabs((double)(next - prev));
            error_total += error / n;
            if (error >= eps)
                active++;
        }
        if (hist_active)
            hist_active[it] = active;
        if (hist_err)
            hist_err[it] = error_total;
        int stop = (stop_rule == OR_STOP_ACTIVE && active == 0) ||
                   (stop_rule == OR_STOP_L1 && error_total <= l1_tol);
        if (stop) {
            stopped = 1;
            break;
        }
    }
    out->iterations = it;
    out->executed = stopped ? it + 1 : it;
    out->active = active;
    out->error_total = error_total;
    double sum = 0.0;
    for (uint32_t v = 0; v < n; v++) {
        if (scaled_out)
            scaled_out[v] = rank[v];
        float r = rank[v] / n;
        if (ranks_out)
            ranks_out[v] = r;
        sum += r;
    }
    out->rank_sum = sum;
    free(rank);
    free(contrib);
    free(codes);
    return 0;
}

/* pageRankPushGraphCSR / pageRankPushFixedPointGraphCSR (pageRank.c:1111-1272, 1992-2117):
 * the pull's loop with next[u] formed by scattering every source's contribution along its
 * out-edges (forward CSR fwd_idx/fwd_deg/fwd_dst), `#pragma omp atomic` adds as at :2062-2063,
 * then the update loop of the pull. kind 0 (FP32) or a fixed-point width. Sequentially the
 * scatter adds each vertex's in-edges in ascending source order, which is the inverse CSR's
 * order, so the result equals or_pagerank_pull's; with threads the FP32 additions take the
 * threads' order and fixed-point sums stay exact. */
int or_pagerank_push(uint32_t n, const uint32_t *outdeg, const uint32_t *fwd_idx,
                     const uint32_t *fwd_deg, const uint32_t *fwd_dst, uint32_t kind,
                     uint32_t max_iter, double eps, uint32_t stop_rule, double l1_tol,
                     uint32_t init, const float *init_ranks, int threads, float *ranks_out,
                     float *scaled_out, uint32_t *hist_active, double *hist_err,
                     or_pr_out *out)
{
    if (kind != 0 && kind != 64 && kind != 32 && kind != 16 && kind != 8)
        return -1;
    const float damp = OR_DAMP;
    const float base = 1.0f - damp;
    const size_t nn = n ? n : 1;
    const size_t csize = kind == 0 ? 4 : kind / 8;
    float *rank = (float *)malloc(nn * sizeof(float));
    void *contrib = malloc(nn * csize);
    float *next_f = (float *)calloc(nn, sizeof(float));
    uint64_t *next_q = (uint64_t *)calloc(nn, sizeof(uint64_t));
    if (!rank || !contrib || !next_f || !next_q) {
        free(rank);
        free(contrib);
        free(next_f);
        free(next_q);
        return -2;
    }
    for (uint32_t v = 0; v < n; v++)
        rank[v] = init == OR_INIT_REFERENCE ? base : init == OR_INIT_UNIFORM ? 1.0f : init_ranks[v];
    memset(out, 0, sizeof(*out));
    uint32_t it, active = 0;
    double error_total = 0.0;
    int stopped = 0;
    const int nt = threads > 0 ? threads : 1;
    for (it = 0; it < max_iter; it++) {
        error_total = 0.0;
        active = 0;
        #pragma omp parallel for schedule(static) if (threads > 0) num_threads(nt)
        for (uint32_t v = 0; v < n; v++)
            contrib_store(kind, contrib, v, rank[v], outdeg[v]);
        #pragma omp parallel for schedule(dynamic, 1024) if (threads > 0) num_threads(nt)
        for (uint32_t v = 0; v < n; v++) {
            const uint32_t b = fwd_idx[v], d = fwd_deg[v];
            if (kind == 0) {
                const float c = ((const float *)contrib)[v];
                for (uint32_t k = 0; k < d; k++) {
                    #pragma omp atomic
                    next_f[fwd_dst[b + k]] += c;
                }
            } else {
                const uint64_t q = contrib_load_q(kind, contrib, v);
                for (uint32_t k = 0; k < d; k++) {
                    #pragma omp atomic
                    next_q[fwd_dst[b + k]] += q;
                }
            }
        }
        #pragma omp parallel for schedule(static) if (threads > 0) num_threads(nt) reduction(+ : error_total, active)
        for (uint32_t v = 0; v < n; v++) {
            const float prev = rank[v];
            const float next = update_rank(kind, base, damp, next_f[v], next_q[v]);
            next_f[v] = 0.0f;
            next_q[v] = 0;
            rank[v] = next;
            const double error = fabs((double)(next - prev));
            error_total += error / n;
            if (error >= eps)
                active++;
        }
        if (hist_active)
            hist_active[it] = active;
        if (hist_err)
            hist_err[it] = error_total;
        if ((stop_rule == OR_STOP_ACTIVE && active == 0) ||
            (stop_rule == OR_STOP_L1 && error_total <= l1_tol)) {
            stopped = 1;
            break;
        }
    }
    out->iterations = it;
    out->executed = stopped ? it + 1 : it;
    out->active = active;
    out->error_total = error_total;
    double sum = 0.0;
    for (uint32_t v = 0; v < n; v++) {
        if (scaled_out)
            scaled_out[v] = rank[v];
        const float r = rank[v] / n;
        if (ranks_out)
            ranks_out[v] = r;
        sum += r;
    }
    out->rank_sum = sum;
    free(rank);
    free(contrib);
    free(next_f);
    free(next_q);
    return 0;
}

/* pageRankDataDrivenPullGraphCSR (pageRank.c:2810-2969). Every vertex starts
 * on the worklist (:2849-2856). Per iteration the contributions rank/out_degree are taken from
 * the ranks at the iteration's start (:2870-2876); each listed vertex sums its in-edges in
 * inverse-CSR order (:2888-2894), adds |new - old| / N to the error (:2896-2899) and, only when
 * |new - old| >= eps, takes the new rank, counts as active and lists its out-neighbours for
 * the next iteration (:2901-2912). The loop stops when no vertex was active (:2922-2923).
 * The contributions are fixed for the whole iteration and each vertex writes only its own
 * rank, so the per-vertex values do not depend on the order vertices are visited in. */
int or_pagerank_dd_pull(uint32_t n, const uint32_t *outdeg, const uint32_t *idx,
                        const uint32_t *deg, const uint32_t *src, const uint32_t *fwd_idx,
                        const uint32_t *fwd_deg, const uint32_t *fwd_dst, uint32_t max_iter,
                        double eps, uint32_t init, const float *init_ranks, int threads,
                        float *ranks_out, float *scaled_out, uint32_t *hist_active,
                        double *hist_err, or_pr_out *out)
{
    const float damp = OR_DAMP;
    const float base = 1.0f - damp;
    const size_t nn = n ? n : 1;
    float *rank = (float *)malloc(nn * sizeof(float));
    float *contrib = (float *)malloc(nn * sizeof(float));
    uint8_t *work = (uint8_t *)malloc(nn);
    uint8_t *next_work = (uint8_t *)malloc(nn);
    if (!rank || !contrib || !work || !next_work) {
        free(rank);
        free(contrib);
        free(work);
        free(next_work);
        return -2;
    }
    for (uint32_t v = 0; v < n; v++) {
        rank[v] = init == OR_INIT_REFERENCE ? base : init == OR_INIT_UNIFORM ? 1.0f : init_ranks[v];
        work[v] = 1;
        next_work[v] = 0;
    }
    memset(out, 0, sizeof(*out));
    uint32_t it, active = 0;
    double error_total = 0.0;
    int stopped = 0;
    const int nt = threads > 0 ? threads : 1;
    for (it = 0; it < max_iter; it++) {
        error_total = 0.0;
        active = 0;
        #pragma omp parallel for schedule(static) if (threads > 0) num_threads(nt)
        for (uint32_t v = 0; v < n; v++)
            contrib[v] = outdeg[v] ? rank[v] / (float)outdeg[v] : 0.0f;
        /* threads > 0: the reference's OpenMP loop (schedule(dynamic,1024), :2890); the marks
         * are idempotent stores, so the per-vertex values equal the sequential run's */
        #pragma omp parallel for schedule(dynamic, 1024) if (threads > 0) num_threads(nt) reduction(+ : error_total, active)
        for (uint32_t v = 0; v < n; v++) {
            if (!work[v])
                continue;
            float sum = 0.0f;
            for (uint32_t k = 0; k < deg[v]; k++)
                sum += contrib[src[idx[v] + k]];
            const float next = base + damp * sum;
            const double error = fabs((double)(next - rank[v]));
            error_total += error / n;
            if (error >= eps) {
                rank[v] = next;
                for (uint32_t k = 0; k < fwd_deg[v]; k++) {
                    #pragma omp atomic write
                    next_work[fwd_dst[fwd_idx[v] + k]] = 1;
                }
                active++;
            }
        }
        uint8_t *t = work;
        work = next_work;
        next_work = t;
        memset(next_work, 0, nn);
        if (hist_active)
            hist_active[it] = active;
        if (hist_err)
            hist_err[it] = error_total;
        if (active == 0) {
            stopped = 1;
            break;
        }
    }
    out->iterations = it;
    out->executed = stopped ? it + 1 : it;
    out->active = active;
    out->error_total = error_total;
    double sum = 0.0;
    for (uint32_t v = 0; v < n; v++) {
        if (scaled_out)
            scaled_out[v] = rank[v];
        const float r = rank[v] / n;
        if (ranks_out)
            ranks_out[v] = r;
        sum += r;
    }
    out->rank_sum = sum;
    free(rank);
    free(contrib);
    free(work);
    free(next_work);
    return 0;
}

/* FP64 power iteration of the same recurrence, from `init` (scaled units, NULL = 1.0),
 * until the normalised L1 delta <= tol or max_iter; returns iterations run. */
int or_pagerank_f64(uint32_t n, const uint32_t *outdeg, const uint32_t *idx, const uint32_t *deg,
                    const uint32_t *src, double tol, uint32_t max_iter, int threads,
                    double *ranks_out)
{
    const double damp = 0.85, base = 1.0 - 0.85;
    double *rank = (double *)malloc((size_t)(n ? n : 1) * sizeof(double));
    double *contrib = (double *)malloc((size_t)(n ? n : 1) * sizeof(double));
    if (!rank || !contrib) {
        free(rank);
        free(contrib);
        return -1;
    }
    const int nt = threads > 0 ? threads : 1;
    for (uint32_t v = 0; v < n; v++)
        rank[v] = 1.0;
    uint32_t it = 0;
    for (it = 0; it < max_iter; it++) {
        #pragma omp parallel for schedule(static) if (threads > 0) num_threads(nt)
        for (uint32_t v = 0; v < n; v++)
            contrib[v] = outdeg[v] ? rank[v] / outdeg[v] : 0.0;
        double l1 = 0.0;
        #pragma omp parallel for schedule(dynamic, 1024) if (threads > 0) num_threads(nt) reduction(+ : l1)
        for (uint32_t v = 0; v < n; v++) {
            double s = 0.0;
            for (uint32_t k = 0; k < deg[v]; k++)
                s += contrib[src[idx[v] + k]];
            double next = base + damp * s;
            l1 += fabs(next - rank[v]);
            rank[v] = next;
        }
        if (n == 0 || l1 / n <= tol) {
            it++;
            break;
        }
    }
    for (uint32_t v = 0; v < n; v++)
        ranks_out[v] = rank[v] / n;
    free(rank);
    free(contrib);
    return (int)it;
}

/* Personalised variant: restart mass (1 - 0.85) * n on `seed` (UINT32_MAX = uniform restart,
 * i.e. plain PageRank), start 1.0 everywhere, FP64, until the normalised L1 delta <= tol. */
int or_ppr_f64(uint32_t n, const uint32_t *outdeg, const uint32_t *idx, const uint32_t *deg,
               const uint32_t *src, uint32_t seed, double tol, uint32_t max_iter, int threads,
               double *ranks_out)
{
    const double damp = 0.85, base = 1.0 - 0.85;
    double *rank = (double *)malloc((size_t)(n ? n : 1) * sizeof(double));
    double *contrib = (double *)malloc((size_t)(n ? n : 1) * sizeof(double));
    if (!rank || !contrib) {
        free(rank);
        free(contrib);
        return -1;
    }
    const int nt = threads > 0 ? threads : 1;
    for (uint32_t v = 0; v < n; v++)
        rank[v] = 1.0;
    uint32_t it = 0;
    for (it = 0; it < max_iter; it++) {
        #pragma omp parallel for schedule(static) if (threads > 0) num_threads(nt)
        for (uint32_t v = 0; v < n; v++)
            contrib[v] = outdeg[v] ? rank[v] / outdeg[v] : 0.0;
        double l1 = 0.0;
        #pragma omp parallel for schedule(dynamic, 1024) if (threads > 0) num_threads(nt) reduction(+ : l1)
        for (uint32_t v = 0; v < n; v++) {
            double s = 0.0;
            for (uint32_t k = 0; k < deg[v]; k++)
                s += contrib[src[idx[v] + k]];
            double p = seed == 0xFFFFFFFFu ? base : (v == seed ? base * n : 0.0);
            double next = p + damp * s;
            l1 += fabs(next - rank[v]);
            rank[v] = next;
        }
        if (n == 0 || l1 / n <= tol) {
            it++;
            break;
        }
    }
    for (uint32_t v = 0; v < n; v++)
        ranks_out[v] = rank[v] / n;
    free(rank);
    free(contrib);
    return (int)it;
}

/* FP64 residual of k rank columns, vertex-major ranks[v * k + c], normalised (divided by n).
 * Column c restarts on seeds[c] (UINT32_MAX: uniform restart, plain PageRank): the exact
 * solution satisfies r_v = p_v + 0.85 * sum_{u -> v} r_u / outdeg_u with p_v = 0.15 / n
 * (uniform) or 0.15 * [v == seed]. Writes, per column, the L1 norm of the residual and of
 * the column. */
int or_ppr_residual(uint32_t n, const uint32_t *outdeg, const uint32_t *idx,
                    const uint32_t *deg, const uint32_t *src, uint32_t k, const uint32_t *seeds,
                    const float *ranks, int threads, double *res_l1, double *rank_l1)
{
    if (k == 0 || k > 256)
        return -1;
    const double damp = 0.85, base = 1.0 - 0.85;
    const int nt = threads > 0 ? threads : 1;
    for (uint32_t c = 0; c < k; c++)
        res_l1[c] = rank_l1[c] = 0.0;
    #pragma omp parallel num_threads(nt)
    {
        double *acc = (double *)calloc(3 * (size_t)k, sizeof(double));
        double *res = acc + k, *nrm = acc + 2 * (size_t)k;
        #pragma omp for schedule(dynamic, 1024)
        for (uint32_t v = 0; v < n; v++) {
            for (uint32_t c = 0; c < k; c++)
                acc[c] = 0.0;
            for (uint32_t e = 0; e < deg[v]; e++) {
                const uint32_t u = src[idx[v] + e];
                const double inv = 1.0 / (double)outdeg[u];
                const float *ru = ranks + (size_t)u * k;
                for (uint32_t c = 0; c < k; c++)
                    acc[c] += (double)ru[c] * inv;
            }
            const float *rv = ranks + (size_t)v * k;
            for (uint32_t c = 0; c < k; c++) {
                const double p = seeds[c] == 0xFFFFFFFFu ? base / n : (v == seeds[c] ? base : 0.0);
                res[c] += fabs(p + damp * acc[c] - (double)rv[c]);
                nrm[c] += fabs((double)rv[c]);
            }
        }
        #pragma omp critical
        for (uint32_t c = 0; c < k; c++) {
            res_l1[c] += res[c];
            rank_l1[c] += nrm[c];
        }
        free(acc);
    }
    return 0;
}

/* ------------------------------------------------------------------ CSR builder */

static int cmp_u64(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return (x > y) - (x < y);
}

/* Sort-based construction: forward edges sorted by (src, dst) with self-loops and duplicate
 * pairs dropped; the inverse view sorted by (dst, src). Outputs must be sized n (degrees,
 * offsets) and m (edge arrays); *m_out receives the simplified edge count. */
int or_build_csr(uint32_t n, uint64_t m, const uint32_t *src, const uint32_t *dst,
                 uint32_t *out_degree, uint32_t *edges_idx, uint32_t *fwd_dest,
                 uint32_t *in_degree, uint32_t *inv_edges_idx, uint32_t *inv_src,
                 uint64_t *m_out)
{
    uint64_t *key = (uint64_t *)malloc((size_t)(m ? m : 1) * sizeof(uint64_t));
    if (!key)
        return -1;
    uint64_t k = 0;
    for (uint64_t i = 0; i < m; i++) {
        if (src[i] >= n || dst[i] >= n) {
            free(key);
            return -2;
        }
        if (src[i] != dst[i])
            key[k++] = ((uint64_t)src[i] << 32) | dst[i];
    }
    qsort(key, (size_t)k, sizeof(uint64_t), cmp_u64);
    uint64_t u = 0;
    for (uint64_t i = 0; i < k; i++)
        if (u == 0 || key[i] != key[u - 1])
            key[u++] = key[i];
    memset(out_degree, 0, (size_t)n * sizeof(uint32_t));
    memset(in_degree, 0, (size_t)n * sizeof(uint32_t));
    for (uint64_t i = 0; i < u; i++) {
        out_degree[key[i] >> 32]++;
        in_degree[(uint32_t)key[i]]++;
        fwd_dest[i] = (uint32_t)key[i];
    }
    uint64_t acc = 0;
    for (uint32_t v = 0; v < n; v++) {
        edges_idx[v] = (uint32_t)acc;
        acc += out_degree[v];
    }
    for (uint64_t i = 0; i < u; i++)
        key[i] = ((uint64_t)(uint32_t)key[i] << 32) | (key[i] >> 32);
    qsort(key, (size_t)u, sizeof(uint64_t), cmp_u64);
    acc = 0;
    for (uint32_t v = 0; v < n; v++) {
        inv_edges_idx[v] = (uint32_t)acc;
        acc += in_degree[v];
    }
    for (uint64_t i = 0; i < u; i++)
        inv_src[i] = (uint32_t)key[i];
    *m_out = u;
    free(key);
    return 0;
}

int or_max_threads(void) { return omp_get_max_threads(); }
