// ag_ppr.hip -- 64-column personalised PageRank pull.
//
// The plain iteration of pageRank.c:942-1109 applied to 64 rank vectors at once, each with
// its own restart vector. Contributions are stored vertex-major, 64 per vertex, so an in-edge
// gathers one contiguous 128-byte line (BF16/FP16) or 256-byte line (FP32): a half-wavefront
// (32 lanes, two columns per lane) reads it with one coalesced load. Work items reuse the pull
// plan's bins: rows of the THREAD/GROUP16/WAVE bins take one half-wavefront each (eight rows
// per workgroup), BLOCK rows a whole workgroup, HUB rows fixed segments whose 64-wide partials
// a one-workgroup finish kernel adds in segment order. The per-column L1 delta is reduced in
// FP64 in a fixed order; the run stops when every column is below the tolerance.
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "ag_internal.h"
#include "ag_pull_kernel.h"

#define PPR_W 64          // columns per vertex (fixed stride)
#define PPR_HW 32         // lanes per half-wavefront
#define PPR_ROWS_PER_WG (AG_WG / PPR_HW)

namespace {

struct PprArgs {
    const void *c16[2];       // 16-bit contributions [N][64] (bf16 or fp16 bits)
    const float *c32[2];      // f32 contributions [N][64]
    float *rank;              // [N][64]
    const uint32_t *outdeg;
    const uint32_t *seed;     // [64] device row of each column's seed, AG_PPR_UNIFORM = uniform
    float *hub_part;          // [n_segs][64]
    double *part;             // [grid][64] per-workgroup column error partials
    double *colerr;           // [64] last column errors
    double *hist;             // [max_iterations] max column error per iteration
    PrState *st;
    uint32_t nparts;
    uint32_t k;
    uint32_t storage;         // AG_STORE_*
    uint32_t stop_rule;
    double l1_tol, refine_l1;
    float base, damp, basen;
    double dn;
};

struct Mode {
    bool r16, w16;
    uint32_t rd, wr;
};

__device__ __forceinline__ Mode ppr_mode(const PprArgs &A, uint32_t iter)
{
    Mode m;
    m.rd = iter & 1u;
    m.wr = m.rd ^ 1u;
    if (A.storage == AG_STORE_F32) {
        m.r16 = m.w16 = false;
        return m;
    }
    const int32_t ri = *(volatile const int32_t *)&A.st->refine_iter;
    m.w16 = !(ri >= 0 && (int32_t)iter >= ri);
    m.r16 = !(ri >= 0 && (int32_t)iter >= ri + 1);
    return m;
}

__device__ __forceinline__ float2 unpack16(uint32_t v, bool f16)
{
    if (f16) {
        _Float16 lo = __builtin_bit_cast(_Float16, (uint16_t)(v & 0xFFFFu));
        _Float16 hi = __builtin_bit_cast(_Float16, (uint16_t)(v >> 16));
        return make_float2((float)lo, (float)hi);
    }
    return make_float2(__uint_as_float(v << 16), __uint_as_float(v & 0xFFFF0000u));
}

__device__ __forceinline__ uint32_t to_bf16(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}

__device__ __forceinline__ uint32_t pack16(float x, float y, bool f16)
{
    if (f16) {
        const uint16_t a = __builtin_bit_cast(uint16_t, (_Float16)x);
        const uint16_t b = __builtin_bit_cast(uint16_t, (_Float16)y);
        return (uint32_t)a | ((uint32_t)b << 16);
    }
    return to_bf16(x) | (to_bf16(y) << 16);
}

// Contribution lines in flight per half-wavefront: the 64-column rows are coalesced 128/256-byte
// reads from HBM, so unlike the scalar gathers this kernel gains from more requests in flight.
#define PPR_UNROLL 4

// Sum of this lane's two columns over in-edges lo + start, lo + start + step, ... < hi.
// PPR_UNROLL independent accumulators, combined in a fixed tree.
__device__ __forceinline__ float2 ppr_edges(const PullGeom &G, const PprArgs &A, const Mode &m,
                                            uint32_t lo, uint32_t hi, uint32_t start,
                                            uint32_t step, uint32_t h)
{
    float2 acc[PPR_UNROLL];
#pragma unroll
    for (int u = 0; u < PPR_UNROLL; u++)
        acc[u] = make_float2(0.f, 0.f);
    const bool f16 = A.storage == AG_STORE_F16;
    const uint32_t *c16 = (const uint32_t *)A.c16[m.rd];
    const float2 *c32 = (const float2 *)A.c32[m.rd];
    for (uint32_t j = lo + start; j < hi; j += PPR_UNROLL * step) {
        uint32_t s[PPR_UNROLL];
#pragma unroll
        for (int u = 0; u < PPR_UNROLL; u++) {
            const uint32_t jj = j + u * step;
            s[u] = jj < hi ? __builtin_nontemporal_load(&G.src[jj]) : 0u;
        }
        float2 v[PPR_UNROLL];
#pragma unroll
        for (int u = 0; u < PPR_UNROLL; u++) {
            const uint32_t jj = j + u * step;
            v[u] = make_float2(0.f, 0.f);
            if (jj < hi)
                v[u] = m.r16 ? unpack16(c16[(size_t)s[u] * PPR_HW + h], f16)
                             : c32[(size_t)s[u] * PPR_HW + h];
        }
#pragma unroll
        for (int u = 0; u < PPR_UNROLL; u++) {
            acc[u].x += v[u].x;
            acc[u].y += v[u].y;
        }
    }
#pragma unroll
    for (int w = PPR_UNROLL / 2; w >= 1; w >>= 1)
#pragma unroll
        for (int u = 0; u < w; u++) {
            acc[u].x += acc[u + w].x;
            acc[u].y += acc[u + w].y;
        }
    return acc[0];
}

// Damping update of columns 2h, 2h+1 of row v; error accumulated per column.
__device__ __forceinline__ void ppr_finish(const PprArgs &A, const Mode &m, uint32_t v,
                                           uint32_t h, float2 sum, double &ex, double &ey)
{
    const uint32_t c0 = 2 * h;
    float nw[2];
    const float sums[2] = {sum.x, sum.y};
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const uint32_t sd = A.seed[c0 + q];
        const float p = sd == AG_PPR_UNIFORM ? A.base : (sd == v ? A.basen : 0.0f);
        nw[q] = __fadd_rn(p, __fmul_rn(A.damp, sums[q]));
    }
    float2 *rk = (float2 *)(A.rank + (size_t)v * PPR_W) + h;
    const float2 old = *rk;
    *rk = make_float2(nw[0], nw[1]);
    ex += __ddiv_rn((double)fabsf(__fsub_rn(nw[0], old.x)), A.dn);
    ey += __ddiv_rn((double)fabsf(__fsub_rn(nw[1], old.y)), A.dn);
    const uint32_t od = A.outdeg[v];
    const float inv = od ? __uint2float_rn(od) : 1.0f;
    const float q0 = od ? __fdiv_rn(nw[0], inv) : 0.0f;
    const float q1 = od ? __fdiv_rn(nw[1], inv) : 0.0f;
    if (m.w16)
        ((uint32_t *)A.c16[m.wr])[(size_t)v * PPR_HW + h] = pack16(q0, q1, A.storage == AG_STORE_F16);
    else
        ((float2 *)A.c32[m.wr])[(size_t)v * PPR_HW + h] = make_float2(q0, q1);
}

// Column errors of one workgroup: s_err[hw][64] summed over half-wavefronts in order.
template <uint32_t NT>
__device__ __forceinline__ void ppr_column_errors(double ex, double ey, double *s_err, double *out)
{
    const uint32_t tid = threadIdx.x, h = tid % PPR_HW, hw = tid / PPR_HW;
    s_err[hw * PPR_W + 2 * h] = ex;
    s_err[hw * PPR_W + 2 * h + 1] = ey;
    __syncthreads();
    if (tid < PPR_W) {
        double t = 0.0;
        for (uint32_t w = 0; w < NT / PPR_HW; w++)
            t += s_err[w * PPR_W + tid];
        out[tid] = t;
    }
    __syncthreads();
}

__global__ __launch_bounds__(AG_WG) void ppr_kernel(PullGeom G, PprArgs A)
{
    __shared__ float s_red[(AG_WG / PPR_HW) * PPR_W];
    __shared__ double s_err[(AG_WG / PPR_HW) * PPR_W];
    __shared__ double s_tot[PPR_W];
    if (*(volatile const uint32_t *)&A.st->done)
        return;
    const uint32_t iter = *(volatile const uint32_t *)&A.st->iter;
    const Mode m = ppr_mode(A, iter);
    const uint32_t tid = threadIdx.x, h = tid % PPR_HW, hw = tid / PPR_HW;
    double ex = 0.0, ey = 0.0;
    const uint32_t lite_rows = G.bin_row[AG_BIN_BLOCK];   // THREAD + GROUP16 + WAVE rows
    const uint32_t lite_items = (lite_rows + PPR_ROWS_PER_WG - 1) / PPR_ROWS_PER_WG;
    const uint32_t nblock = G.bin_row[AG_BIN_BLOCK + 1] - G.bin_row[AG_BIN_BLOCK];
    const uint32_t nseg = G.slot_blk0[1] - G.slot_blk0[0];
    const uint32_t total = nseg + nblock + lite_items;
    for (uint32_t it = blockIdx.x; it < total; it += gridDim.x) {
        if (it < nseg + nblock) {
            // a whole workgroup on one hub segment or one BLOCK row
            uint32_t lo, hi, row = 0;
            const bool hub = it < nseg;
            if (hub) {
                lo = G.seg_lo[it];
                hi = G.seg_hi[it];
            } else {
                row = G.bin_row[AG_BIN_BLOCK] + (it - nseg);
                lo = G.off[row];
                hi = G.off[row + 1];
            }
            const float2 a = ppr_edges(G, A, m, lo, hi, hw, AG_WG / PPR_HW, h);
            s_red[hw * PPR_W + 2 * h] = a.x;
            s_red[hw * PPR_W + 2 * h + 1] = a.y;
            __syncthreads();
            if (hw == 0) {
                float2 t = make_float2(0.f, 0.f);
                for (uint32_t w = 0; w < AG_WG / PPR_HW; w++) {
                    t.x += s_red[w * PPR_W + 2 * h];
                    t.y += s_red[w * PPR_W + 2 * h + 1];
                }
                if (hub)
                    ((float2 *)(A.hub_part + (size_t)it * PPR_W))[h] = t;
                else
                    ppr_finish(A, m, row, h, t, ex, ey);
            }
            __syncthreads();
        } else {
            const uint32_t row = (it - nseg - nblock) * PPR_ROWS_PER_WG + hw;
            if (row < lite_rows) {
                const float2 a = ppr_edges(G, A, m, G.off[row], G.off[row + 1], 0, 1, h);
                ppr_finish(A, m, row, h, a, ex, ey);
            }
        }
    }
    ppr_column_errors<AG_WG>(ex, ey, s_err, s_tot);
    if (tid < PPR_W)
        A.part[(size_t)blockIdx.x * PPR_W + tid] = s_tot[tid];
}

__global__ __launch_bounds__(AG_FIN_WG) void ppr_finish_kernel(PullGeom G, PprArgs A)
{
    __shared__ double s_err[(AG_FIN_WG / PPR_HW) * PPR_W];
    __shared__ double s_tot[PPR_W];
    if (*(volatile const uint32_t *)&A.st->done)
        return;
    const uint32_t iter = *(volatile const uint32_t *)&A.st->iter;
    const Mode m = ppr_mode(A, iter);
    const uint32_t tid = threadIdx.x, h = tid % PPR_HW, hw = tid / PPR_HW;
    double ex = 0.0, ey = 0.0;
    for (uint32_t hub = hw; hub < G.n_hubs; hub += AG_FIN_WG / PPR_HW) {
        const uint32_t s0 = G.hub_seg0[hub], s1 = G.hub_seg0[hub + 1];
        float2 t = ((const float2 *)(A.hub_part + (size_t)s0 * PPR_W))[h];
        for (uint32_t q = s0 + 1; q < s1; q++) {
            const float2 u = ((const float2 *)(A.hub_part + (size_t)q * PPR_W))[h];
            t.x += u.x;
            t.y += u.y;
        }
        ppr_finish(A, m, G.bin_row[AG_BIN_HUB] + hub, h, t, ex, ey);
    }
    ppr_column_errors<AG_FIN_WG>(ex, ey, s_err, s_tot);
    if (tid < PPR_W) {
        double t = s_tot[tid];
        for (uint32_t b = 0; b < A.nparts; b++)
            t += A.part[(size_t)b * PPR_W + tid];
        s_tot[tid] = t;
    }
    __syncthreads();
    if (tid != 0)
        return;
    double E = 0.0;
    for (uint32_t c = 0; c < A.k; c++) {
        A.colerr[c] = s_tot[c];
        E = s_tot[c] > E ? s_tot[c] : E;
    }
    PrState *st = A.st;
    A.hist[iter] = E;
    bool can_stop = true;
    if (A.storage != AG_STORE_F32) {
        const int32_t ri = st->refine_iter;
        if (ri < 0) {
            refine_decide(st, iter, E, A.refine_l1);
            can_stop = false;
        } else {
            can_stop = (int32_t)iter >= ri + 1;
        }
    }
    if (can_stop && A.stop_rule == AG_STOP_L1 && E <= A.l1_tol) {
        st->done = 1;
        st->stop_iter = iter;
    }
    st->last_err = E;
    st->iter = iter + 1;
}

__global__ void ppr_init_kernel(uint32_t n, uint32_t storage, void *c16, float *c32, float *rank,
                                const uint32_t *outdeg)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;  // (vertex, lane)
    if (i >= (uint64_t)n * PPR_HW)
        return;
    const uint32_t v = (uint32_t)(i / PPR_HW);
    ((float2 *)rank)[i] = make_float2(1.0f, 1.0f);
    const uint32_t od = outdeg[v];
    const float q = od ? __fdiv_rn(1.0f, __uint2float_rn(od)) : 0.0f;
    if (storage == AG_STORE_F32)
        ((float2 *)c32)[i] = make_float2(q, q);
    else
        ((uint32_t *)c16)[i] = pack16(q, q, storage == AG_STORE_F16);
}

__global__ void ppr_out_kernel(uint32_t n, uint32_t k, const float *rank, const uint32_t *d2o,
                               float nf, float *out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)n * k)
        return;
    const uint32_t v = (uint32_t)(i / k), c = (uint32_t)(i % k);
    out[(size_t)d2o[v] * k + c] = __fdiv_rn(rank[(size_t)v * PPR_W + c], nf);
}

} // namespace

extern "C" void ag_ppr_options_default(ag_ppr_options *o)
{
    memset(o, 0, sizeof(*o));
    o->max_iterations = 100;
    o->stop_rule = AG_STOP_L1;
    o->l1_tolerance = 1e-6;
    o->storage = AG_STORE_BF16;
    o->refine_l1 = 1e-3;
    o->chunk = 0;
}

extern "C" int ag_ppr_pull(ag_dgraph *g, uint32_t k, const uint32_t *seeds,
                           const ag_ppr_options *opts, float *ranks_out, double *col_error,
                           ag_ppr_result *res)
{
    if (!g || k == 0 || k > AG_PPR_MAX_COLS || !seeds || (g->n && !ranks_out))
        return ag_set_error(AG_ERR_INVALID, "bad argument (k must be in [1, 64])");
    ag_ppr_options o;
    if (opts)
        o = *opts;
    else
        ag_ppr_options_default(&o);
    if (o.stop_rule != AG_STOP_L1 && o.stop_rule != AG_STOP_NONE)
        return ag_set_error(AG_ERR_INVALID, "PPR supports AG_STOP_L1 and AG_STOP_NONE");
    if (o.storage > AG_STORE_F16)
        return ag_set_error(AG_ERR_INVALID, "unknown storage %u", o.storage);
    const uint32_t n = g->n;
    // seed vertex -> device row
    uint32_t dseed[AG_PPR_MAX_COLS];
    for (uint32_t c = 0; c < AG_PPR_MAX_COLS; c++)
        dseed[c] = AG_PPR_UNIFORM;
    for (uint32_t c = 0; c < k; c++) {
        if (seeds[c] != AG_PPR_UNIFORM && seeds[c] >= n)
            return ag_set_error(AG_ERR_INVALID, "seed %u of column %u out of range", seeds[c], c);
    }
    {
        uint32_t *o2d = (uint32_t *)malloc((size_t)(n ? n : 1) * sizeof(uint32_t));
        if (!o2d)
            return ag_set_error(AG_ERR_OOM, "host allocation failed");
        for (uint32_t r = 0; r < n; r++)
            o2d[g->h_d2o[r]] = r;
        for (uint32_t c = 0; c < k; c++)
            if (seeds[c] != AG_PPR_UNIFORM)
                dseed[c] = o2d[seeds[c]];
        free(o2d);
    }
    memset(res, 0, sizeof(*res));
    res->refine_iteration = -1;
    if (n == 0)
        return AG_OK;
    AG_HIP(hipSetDevice(g->dev->ordinal));
    hipStream_t s = g->dev->stream;
    const uint32_t maxit = o.max_iterations ? o.max_iterations : 1;
    // 64-column rows are wide coalesced reads: this kernel keeps 8 workgroups per CU
    const uint32_t grid = ag_pull_grid(g, (AG_WG / PPR_HW) * PPR_W * 12u, 8u) ?
                          ag_pull_grid(g, (AG_WG / PPR_HW) * PPR_W * 12u, 8u) : 1;
    const size_t nv = (size_t)n * PPR_W;
    void *buf[12] = {0};
    int nb = 0;
    int st = AG_OK;
    auto alloc = [&](void **p, size_t bytes) {
        if (st != AG_OK)
            return;
        hipError_t e = hipMalloc(p, bytes ? bytes : 8);
        if (e != hipSuccess)
            st = ag_hip_fail(e, "PPR allocation", __FILE__, __LINE__);
        else
            buf[nb++] = *p;
    };
    PprArgs A;
    memset(&A, 0, sizeof(A));
    void *c16a = NULL, *c16b = NULL, *c32a = NULL, *c32b = NULL, *rank = NULL, *seedd = NULL,
         *hubp = NULL, *part = NULL, *colerr = NULL, *hist = NULL, *stv = NULL, *outd = NULL;
    alloc(&rank, nv * sizeof(float));
    alloc(&c32a, nv * sizeof(float));
    alloc(&c32b, nv * sizeof(float));
    if (o.storage != AG_STORE_F32) {
        alloc(&c16a, nv * 2);
        alloc(&c16b, nv * 2);
    }
    alloc(&seedd, sizeof(dseed));
    alloc(&hubp, (size_t)(g->n_segs ? g->n_segs : 1) * PPR_W * sizeof(float));
    alloc(&part, (size_t)grid * PPR_W * sizeof(double));
    alloc(&colerr, PPR_W * sizeof(double));
    alloc(&hist, (size_t)maxit * sizeof(double));
    alloc(&stv, sizeof(PrState));
    PrState *sth = NULL;
    if (st == AG_OK) {
        hipError_t e = hipHostMalloc((void **)&sth, sizeof(PrState), hipHostMallocDefault);
        if (e != hipSuccess)
            st = ag_hip_fail(e, "hipHostMalloc", __FILE__, __LINE__);
    }
    hipEvent_t e0 = NULL, e1 = NULL;
    if (st == AG_OK) {
        memset(sth, 0, sizeof(PrState));
        sth->refine_iter = -1;
        hipError_t e = hipMemcpyAsync(stv, sth, sizeof(PrState), hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(seedd, dseed, sizeof(dseed), hipMemcpyHostToDevice, s);
        if (e == hipSuccess) {
            const uint64_t tot = (uint64_t)n * PPR_HW;
            ppr_init_kernel<<<(uint32_t)((tot + 255) / 256), 256, 0, s>>>(
                n, o.storage, c16a, (float *)c32a, (float *)rank, g->outdeg);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipEventCreate(&e0);
        if (e == hipSuccess)
            e = hipEventCreate(&e1);
        if (e == hipSuccess)
            e = hipEventRecord(e0, s);
        if (e != hipSuccess)
            st = ag_hip_fail(e, "PPR setup", __FILE__, __LINE__);
    }
    if (st == AG_OK) {
        A.c16[0] = c16a;
        A.c16[1] = c16b;
        A.c32[0] = (const float *)c32a;
        A.c32[1] = (const float *)c32b;
        A.rank = (float *)rank;
        A.outdeg = g->outdeg;
        A.seed = (const uint32_t *)seedd;
        A.hub_part = (float *)hubp;
        A.part = (double *)part;
        A.colerr = (double *)colerr;
        A.hist = (double *)hist;
        A.st = (PrState *)stv;
        A.nparts = grid;
        A.k = k;
        A.storage = o.storage;
        A.stop_rule = o.stop_rule;
        A.l1_tol = o.l1_tolerance;
        A.refine_l1 = o.refine_l1;
        A.damp = 0.85f;
        volatile float d = 0.85f;
        A.base = 1.0f - d;
        A.basen = A.base * (float)n;
        A.dn = (double)n;
        PullGeom G = ag_make_geom(g);
        const uint32_t chunk = o.chunk ? o.chunk : 8;
        uint32_t enq = 0;
        while (enq < o.max_iterations && st == AG_OK) {
            const uint32_t kk = o.max_iterations - enq < chunk ? o.max_iterations - enq : chunk;
            for (uint32_t i = 0; i < kk; i++) {
                ppr_kernel<<<grid, AG_WG, 0, s>>>(G, A);
                ppr_finish_kernel<<<1, AG_FIN_WG, 0, s>>>(G, A);
            }
            enq += kk;
            hipError_t e = hipGetLastError();
            if (e == hipSuccess)
                e = hipEventRecord(e1, s);
            if (e == hipSuccess)
                e = hipMemcpyAsync(sth, stv, sizeof(PrState), hipMemcpyDeviceToHost, s);
            st = (e == hipSuccess) ? ag_wait_stream(g->dev) : ag_hip_fail(e, "PPR launch", __FILE__, __LINE__);
            if (st == AG_OK && sth->done)
                break;
        }
    }
    if (st == AG_OK) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            ms = 0.0f;   // timing is informational; the ranks are already valid
        res->loop_ms = ms;
        res->iterations = sth->done ? sth->stop_iter : sth->iter;
        res->executed = sth->done ? sth->stop_iter + 1 : sth->iter;
        res->max_error = sth->last_err;
        res->refine_iteration = sth->refine_iter;
        // normalised ranks, original order; reuse c32b as staging
        outd = c32b;
        const uint64_t tot = (uint64_t)n * k;
        ppr_out_kernel<<<(uint32_t)((tot + 255) / 256), 256, 0, s>>>(n, k, (const float *)rank,
                                                                     g->d2o, (float)n, (float *)outd);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(ranks_out, outd, tot * sizeof(float), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess && col_error)
            e = hipMemcpyAsync(col_error, colerr, (size_t)k * sizeof(double), hipMemcpyDeviceToHost, s);
        st = (e == hipSuccess) ? ag_wait_stream(g->dev) : ag_hip_fail(e, "PPR output", __FILE__, __LINE__);
    }
    if (e0)
        (void)hipEventDestroy(e0);
    if (e1)
        (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(s);
    for (int i = 0; i < nb; i++)
        (void)hipFree(buf[i]);
    if (sth)
        (void)hipHostFree(sth);
    return st;
}
