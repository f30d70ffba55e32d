// ag_ppr.hip -- 64-column personalised PageRank pull.
//
// The plain iteration of pageRank.c:942-1109 applied to 64 rank vectors at once, each with
// its own restart vector. Contributions are stored vertex-major, 64 per vertex, so an in-edge
// gathers one contiguous 128-byte line (BF16/FP16) or 256-byte line (FP32): a group of 16
// lanes (four columns per lane, 8-byte or 16-byte loads) reads it with one coalesced load, and
// a group keeps the lines of 16 in-edges in flight at once, their source ids fetched by one
// coalesced load and passed lane to lane. Work items reuse the pull plan's bins: THREAD and
// GROUP16 rows take one group each (sixteen rows per workgroup), WAVE rows a wavefront (four
// groups), BLOCK rows a whole workgroup, HUB rows fixed segments whose 64-wide partials a
// one-workgroup finish kernel adds in segment order. The per-column L1 delta is reduced in
// FP64 in a fixed order; the run stops when every column is below the tolerance. The solve is
// captured once into a hipGraph and replayed.
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "ag_internal.h"
#include "ag_pull_kernel.h"

#define PPR_W 64          // columns per vertex (fixed stride)
#define PPR_GL 16         // lanes per group: one row's 64 columns, four per lane
#define PPR_GROUPS (AG_WG / PPR_GL)
#define PPR_FIN_GROUPS (AG_FIN_WG / PPR_GL)

namespace {

// Run options read by the kernels from device memory, so one captured solve serves any
// seeds, column count, storage and stop rule.
struct PprConf {
    uint32_t seed[PPR_W];     // device row of each column's seed, AG_PPR_UNIFORM = uniform
    // FP16 storage holds contribution * 2^-f16_bias (the sum is multiplied back by 2^f16_bias,
    // exactly): every scaled rank, so every contribution, is at most N, and N * 2^-f16_bias
    // stays below FP16's largest finite value
    int32_t f16_bias;
    uint32_t k;
    uint32_t storage;         // AG_STORE_*
    uint32_t stop_rule;
    uint32_t pad;
    double l1_tol, refine_l1;
};

struct PprArgs {
    const void *c16[2];       // 16-bit contributions [N][64] (bf16 or fp16 bits)
    const float *c32[2];      // f32 contributions [N][64]
    float *rank;              // [N][64]
    const uint32_t *outdeg;
    const PprConf *conf;
    float *hub_part;          // [n_segs][64]
    double *part;             // [grid][64] per-workgroup column error partials
    double *colerr;           // [64] last column errors
    double *hist;             // [max_iterations] max column error per iteration
    PrState *st;
    uint32_t nparts;
    uint32_t n;
    float base, damp, basen;
    double dn;
};

struct Mode {
    bool r16, w16, f16;
    uint32_t rd, wr;
};

__device__ __forceinline__ Mode ppr_mode(const PprArgs &A, uint32_t storage, uint32_t iter)
{
    Mode m;
    m.rd = iter & 1u;
    m.wr = m.rd ^ 1u;
    m.f16 = storage == AG_STORE_F16;
    if (storage == AG_STORE_F32) {
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

__device__ __forceinline__ void add4(float4 &a, const float4 &b)
{
    a.x += b.x;
    a.y += b.y;
    a.z += b.z;
    a.w += b.w;
}

// This lane's source id in the first 16-edge chunk of group gi over in-edges [lo, hi)
// (0 past hi).
__device__ __forceinline__ uint32_t ppr_first_ids(const PullGeom &G, uint32_t lo, uint32_t hi,
                                                  uint32_t gi, uint32_t q)
{
    const uint32_t j = lo + gi * PPR_GL + q;
    return j < hi ? __builtin_nontemporal_load(&G.src[j]) : 0u;
}

// Sum of columns 4q..4q+3 over the in-edges [lo, hi) owned by group gi of the ng groups that
// share the row: the edges are cut into 16-edge chunks, chunk c belongs to group c % ng. A
// chunk's source ids arrive in one coalesced load (one per lane; `sl` holds the first chunk's,
// loaded by the caller) and are passed lane to lane, so every contribution line of the chunk
// is in flight at once: 16 loads of 8 bytes per lane (16-bit storage) or two rounds of 8 loads
// of 16 bytes (FP32); the next chunk's ids are loaded before them. Per lane the terms are added
// in edge order into two accumulators (even and odd edges), combined at the end.
__device__ __forceinline__ float4 ppr_group_sum(const PullGeom &G, const PprArgs &A,
                                                const Mode &m, uint32_t lo, uint32_t hi,
                                                uint32_t gi, uint32_t ng, uint32_t q,
                                                uint32_t sl)
{
    float4 a0 = make_float4(0.f, 0.f, 0.f, 0.f), a1 = a0;
    const uint2 *c16 = (const uint2 *)A.c16[m.rd];
    const float4 *c32 = (const float4 *)A.c32[m.rd];
    const uint32_t stride = ng * PPR_GL;
    for (uint32_t base = lo + gi * PPR_GL; base < hi; base += stride) {
        const uint32_t cnt = hi - base < PPR_GL ? hi - base : PPR_GL;
        const uint32_t nb = base + stride;
        const uint32_t sl_next =
            (nb < hi && nb + q < hi) ? __builtin_nontemporal_load(&G.src[nb + q]) : 0u;
        if (m.r16) {
            uint2 v[PPR_GL];
#pragma unroll
            for (uint32_t t = 0; t < PPR_GL; t++) {
                const uint32_t sv = (uint32_t)__shfl((int)sl, (int)t, PPR_GL);
                v[t] = t < cnt ? c16[(size_t)sv * PPR_GL + q] : make_uint2(0u, 0u);
            }
#pragma unroll
            for (uint32_t t = 0; t < PPR_GL; t++) {
                const float2 lo2 = unpack16(v[t].x, m.f16), hi2 = unpack16(v[t].y, m.f16);
                add4((t & 1u) ? a1 : a0, make_float4(lo2.x, lo2.y, hi2.x, hi2.y));
            }
        } else {
#pragma unroll
            for (uint32_t h = 0; h < PPR_GL; h += PPR_GL / 2) {
                float4 v[PPR_GL / 2];
#pragma unroll
                for (uint32_t t = 0; t < PPR_GL / 2; t++) {
                    const uint32_t sv = (uint32_t)__shfl((int)sl, (int)(h + t), PPR_GL);
                    v[t] = h + t < cnt ? c32[(size_t)sv * PPR_GL + q]
                                       : make_float4(0.f, 0.f, 0.f, 0.f);
                }
#pragma unroll
                for (uint32_t t = 0; t < PPR_GL / 2; t++)
                    add4((t & 1u) ? a1 : a0, v[t]);
            }
        }
        sl = sl_next;
    }
    add4(a0, a1);
    return a0;
}

// Damping update of columns 4q..4q+3 of row v; each column's |delta| / N is accumulated in
// err (FP64), the next contribution rank / out_degree stored in 16 bits or FP32.
__device__ __forceinline__ void ppr_finish(const PprArgs &A, const Mode &m, uint32_t v,
                                           uint32_t q, float4 sum, const uint32_t *seed,
                                           double *err)
{
    float sums[4] = {sum.x, sum.y, sum.z, sum.w};
    if (m.r16 && m.f16) {   // FP16 terms were stored scaled by 2^-bias: exact power-of-two undo
#pragma unroll
        for (int c = 0; c < 4; c++)
            sums[c] = ldexpf(sums[c], A.conf->f16_bias);
    }
    float nw[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const float p = seed[c] == AG_PPR_UNIFORM ? A.base : (seed[c] == v ? A.basen : 0.0f);
        nw[c] = __fadd_rn(p, __fmul_rn(A.damp, sums[c]));
    }
    float4 *rk = (float4 *)(A.rank + (size_t)v * PPR_W) + q;
    const float4 old = *rk;
    *rk = make_float4(nw[0], nw[1], nw[2], nw[3]);
    const float olds[4] = {old.x, old.y, old.z, old.w};
#pragma unroll
    for (int c = 0; c < 4; c++)
        err[c] += __ddiv_rn((double)fabsf(__fsub_rn(nw[c], olds[c])), A.dn);
    const uint32_t od = A.outdeg[v];
    const float inv = od ? __uint2float_rn(od) : 1.0f;
    float qv[4];
#pragma unroll
    for (int c = 0; c < 4; c++)
        qv[c] = od ? __fdiv_rn(nw[c], inv) : 0.0f;
    if (m.w16) {
        if (m.f16) {
#pragma unroll
            for (int c = 0; c < 4; c++)
                qv[c] = ldexpf(qv[c], -A.conf->f16_bias);
        }
        ((uint2 *)A.c16[m.wr])[(size_t)v * PPR_GL + q] =
            make_uint2(pack16(qv[0], qv[1], m.f16), pack16(qv[2], qv[3], m.f16));
    }
    else
        ((float4 *)A.c32[m.wr])[(size_t)v * PPR_GL + q] = make_float4(qv[0], qv[1], qv[2], qv[3]);
}

// Column errors of one workgroup of NG groups: s_err[group][64] summed over groups in order.
template <uint32_t NG>
__device__ __forceinline__ void ppr_column_errors(const double *err, double *s_err, double *out)
{
    const uint32_t tid = threadIdx.x, q = tid % PPR_GL, gi = tid / PPR_GL;
#pragma unroll
    for (int c = 0; c < 4; c++)
        s_err[gi * PPR_W + 4 * q + c] = err[c];
    __syncthreads();
    if (tid < PPR_W) {
        double t = 0.0;
        for (uint32_t w = 0; w < NG; w++)
            t += s_err[w * PPR_W + tid];
        out[tid] = t;
    }
    __syncthreads();
}

// One iteration's gather and update. Work items, heaviest first: hub segments and BLOCK rows
// (16 groups on one row; partials combined in LDS in group order), WAVE rows (the four groups
// of a wavefront on one row; combined by two lane exchanges), THREAD and GROUP16 rows (one
// group per row, sixteen rows per workgroup).
__global__ __launch_bounds__(AG_WG) void ppr_kernel(PullGeom G, PprArgs A)
{
    __shared__ float4 s_red[PPR_GROUPS * PPR_GL];
    __shared__ double s_err[PPR_GROUPS * PPR_W];
    __shared__ double s_tot[PPR_W];
    if (*(volatile const uint32_t *)&A.st->done)
        return;
    const uint32_t iter = *(volatile const uint32_t *)&A.st->iter;
    const Mode m = ppr_mode(A, A.conf->storage, iter);
    const uint32_t tid = threadIdx.x, q = tid % PPR_GL, gi = tid / PPR_GL;
    const uint32_t wl = (tid % AG_WAVE) / PPR_GL;   // group inside its wavefront
    uint32_t seed[4];
#pragma unroll
    for (int c = 0; c < 4; c++)
        seed[c] = A.conf->seed[4 * q + c];
    double err[4] = {0.0, 0.0, 0.0, 0.0};
    const uint32_t lite_rows = G.bin_row[AG_BIN_WAVE];    // THREAD + GROUP16 rows
    const uint32_t wave0 = G.bin_row[AG_BIN_WAVE], wave1 = G.bin_row[AG_BIN_WAVE + 1];
    const uint32_t nwave = (wave1 - wave0 + AG_WG / AG_WAVE - 1) / (AG_WG / AG_WAVE);
    const uint32_t nlite = (lite_rows + PPR_GROUPS - 1) / PPR_GROUPS;
    const uint32_t nblock = G.bin_row[AG_BIN_BLOCK + 1] - G.bin_row[AG_BIN_BLOCK];
    const uint32_t nseg = G.slot_blk0[1] - G.slot_blk0[0];
    const uint32_t heavy = nseg + nblock;
    const uint32_t total = heavy + nwave + nlite;
    const uint32_t lite0 = heavy + nwave;
    uint32_t it = blockIdx.x;
    for (; it < lite0; it += gridDim.x) {
        if (it < heavy) {
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
            s_red[gi * PPR_GL + q] = ppr_group_sum(G, A, m, lo, hi, gi, PPR_GROUPS, q,
                                                   ppr_first_ids(G, lo, hi, gi, q));
            __syncthreads();
            if (gi == 0) {
                float4 t = s_red[q];
                for (uint32_t w = 1; w < PPR_GROUPS; w++)
                    add4(t, s_red[w * PPR_GL + q]);
                if (hub)
                    ((float4 *)(A.hub_part + (size_t)it * PPR_W))[q] = t;
                else
                    ppr_finish(A, m, row, q, t, seed, err);
            }
            __syncthreads();
        } else if (it < heavy + nwave) {
            const uint32_t row = wave0 + (it - heavy) * (AG_WG / AG_WAVE) + tid / AG_WAVE;
            const bool valid = row < wave1;
            float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
            if (valid) {
                const uint32_t lo = G.off[row], hi = G.off[row + 1];
                a = ppr_group_sum(G, A, m, lo, hi, wl, AG_WAVE / PPR_GL, q,
                                  ppr_first_ids(G, lo, hi, wl, q));
            }
            // (g0 + g1) + (g2 + g3), the same value in every group (addition commutes)
#pragma unroll
            for (int msk = PPR_GL; msk < AG_WAVE; msk <<= 1) {
                a.x += __shfl_xor(a.x, msk, AG_WAVE);
                a.y += __shfl_xor(a.y, msk, AG_WAVE);
                a.z += __shfl_xor(a.z, msk, AG_WAVE);
                a.w += __shfl_xor(a.w, msk, AG_WAVE);
            }
            if (valid && wl == 0)
                ppr_finish(A, m, row, q, a, seed, err);
        }
    }
    // THREAD and GROUP16 rows, a group each, two rows ahead: while row k is gathered, the
    // first source ids of row k + 1 and the offsets of row k + 2 are in flight
    {
        const uint32_t step = gridDim.x * PPR_GROUPS;
        uint32_t row = it < total ? (it - lite0) * PPR_GROUPS + gi : lite_rows;
        uint32_t lo = 0, hi = 0, lo1 = 0, hi1 = 0, sl = 0;
        if (row < lite_rows) {
            lo = G.off[row];
            hi = G.off[row + 1];
            sl = ppr_first_ids(G, lo, hi, 0, q);
        }
        if (row < lite_rows && row + step < lite_rows) {
            lo1 = G.off[row + step];
            hi1 = G.off[row + step + 1];
        }
        for (; row < lite_rows; row += step) {
            uint32_t lo2 = 0, hi2 = 0;
            if (row + 2 * step < lite_rows) {
                lo2 = G.off[row + 2 * step];
                hi2 = G.off[row + 2 * step + 1];
            }
            const uint32_t sl1 = ppr_first_ids(G, lo1, hi1, 0, q);
            const float4 a = ppr_group_sum(G, A, m, lo, hi, 0, 1, q, sl);
            ppr_finish(A, m, row, q, a, seed, err);
            lo = lo1;
            hi = hi1;
            sl = sl1;
            lo1 = lo2;
            hi1 = hi2;
        }
    }
    ppr_column_errors<PPR_GROUPS>(err, s_err, s_tot);
    if (tid < PPR_W)
        A.part[(size_t)blockIdx.x * PPR_W + tid] = s_tot[tid];
}

// Hub rows (their segment partials added in segment order), the column errors of the
// iteration (fixed order) and the stop / refinement decisions. One workgroup.
__global__ __launch_bounds__(AG_FIN_WG) void ppr_finish_kernel(PullGeom G, PprArgs A)
{
    __shared__ double s_err[PPR_FIN_GROUPS * PPR_W];
    __shared__ double s_tot[PPR_W];
    if (*(volatile const uint32_t *)&A.st->done)
        return;
    const PprConf &C = *A.conf;
    const uint32_t iter = *(volatile const uint32_t *)&A.st->iter;
    const Mode m = ppr_mode(A, C.storage, iter);
    const uint32_t tid = threadIdx.x, q = tid % PPR_GL, gi = tid / PPR_GL;
    uint32_t seed[4];
#pragma unroll
    for (int c = 0; c < 4; c++)
        seed[c] = C.seed[4 * q + c];
    double err[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t hub = gi; hub < G.n_hubs; hub += PPR_FIN_GROUPS) {
        const uint32_t s0 = G.hub_seg0[hub], s1 = G.hub_seg0[hub + 1];
        float4 t = ((const float4 *)(A.hub_part + (size_t)s0 * PPR_W))[q];
        for (uint32_t sg = s0 + 1; sg < s1; sg++)
            add4(t, ((const float4 *)(A.hub_part + (size_t)sg * PPR_W))[q]);
        ppr_finish(A, m, G.bin_row[AG_BIN_HUB] + hub, q, t, seed, err);
    }
    ppr_column_errors<PPR_FIN_GROUPS>(err, s_err, s_tot);
    // the iteration kernel's per-workgroup column errors: lane (chunk j, column c) adds the
    // workgroups of chunk j in order (8 loads in flight), then each column adds its hub term
    // and the chunks in chunk order -- a fixed order for any grid
    {
        constexpr uint32_t NCH = AG_FIN_WG / PPR_W, U = 8;
        const uint32_t c = tid % PPR_W, j = tid / PPR_W;
        const uint32_t len = (A.nparts + NCH - 1) / NCH;
        const uint32_t b0 = j * len, b1 = b0 + len < A.nparts ? b0 + len : A.nparts;
        double t = 0.0;
        for (uint32_t b = b0; b < b1; b += U) {
            double v[U];
#pragma unroll
            for (uint32_t u = 0; u < U; u++)
                v[u] = b + u < b1 ? A.part[(size_t)(b + u) * PPR_W + c] : 0.0;
#pragma unroll
            for (uint32_t u = 0; u < U; u++)
                t += v[u];
        }
        s_err[j * PPR_W + c] = t;   // s_err is free again (ppr_column_errors ends synced)
        __syncthreads();
        if (tid < PPR_W) {
            double tot = s_tot[tid];
            for (uint32_t q2 = 0; q2 < NCH; q2++)
                tot += s_err[q2 * PPR_W + tid];
            s_tot[tid] = tot;
        }
        __syncthreads();
    }
    if (tid != 0)
        return;
    double E = 0.0;
    bool finite = true;
    for (uint32_t c = 0; c < C.k; c++) {
        A.colerr[c] = s_tot[c];
        E = s_tot[c] > E ? s_tot[c] : E;
        finite = finite && isfinite(s_tot[c]);
    }
    if (!finite)   // a column that overflowed never satisfies the stop rule
        E = INFINITY;
    PrState *st = A.st;
    A.hist[iter] = E;
    bool can_stop = true;
    if (C.storage != AG_STORE_F32) {
        const int32_t ri = st->refine_iter;
        if (ri < 0) {
            refine_decide(st, iter, E, C.refine_l1);
            can_stop = false;
        } else {
            can_stop = (int32_t)iter >= ri + 1;
        }
    }
    if (can_stop && C.stop_rule == AG_STOP_L1 && E <= C.l1_tol) {
        st->done = 1;
        st->stop_iter = iter;
    }
    st->last_err = E;
    st->iter = iter + 1;
}

// Start of a solve: every rank 1.0, contributions 1 / out_degree (0 for sinks) in the storage
// of the first iteration, and the iteration state reset (thread 0; no other thread reads it
// in this launch).
__global__ void ppr_init_kernel(PprArgs A)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;  // (vertex, lane)
    if (i == 0) {
        PrState *st = A.st;
        st->iter = 0;
        st->done = 0;
        st->stop_iter = 0;
        st->refine_iter = -1;
        st->last_active = 0;
        st->last_err = 0.0;
        st->best_err = 0.0;
        st->stall = 0;
    }
    if (i >= (uint64_t)A.n * PPR_GL)
        return;
    const uint32_t v = (uint32_t)(i / PPR_GL);
    ((float4 *)A.rank)[i] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    const uint32_t od = A.outdeg[v];
    const float c = od ? __fdiv_rn(1.0f, __uint2float_rn(od)) : 0.0f;
    const uint32_t storage = A.conf->storage;
    if (storage == AG_STORE_F32) {
        ((float4 *)A.c32[0])[i] = make_float4(c, c, c, c);
    } else {
        const bool f16 = storage == AG_STORE_F16;
        const float cs = f16 ? ldexpf(c, -A.conf->f16_bias) : c;
        const uint32_t p = pack16(cs, cs, f16);
        ((uint2 *)A.c16[0])[i] = make_uint2(p, p);
    }
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

// Per-graph PPR workspace, kept between calls (the captured solve holds its pointers).
struct ag_ppr_ws {
    uint32_t n, grid_max, hist_cap;
    uint32_t grid;             // workgroups of this call's iteration kernel (<= grid_max)
    void *c16[2];
    float *c32[2];
    float *rank;
    float *hub_part;
    double *part, *colerr, *hist;
    PrState *st;
    PprConf *conf;
    PrState *st_host;          // pinned
    PprConf *conf_host;        // pinned
    hipGraph_t graph;
    hipGraphExec_t exec;
    uint32_t exec_maxit;       // iterations in the captured solve (0: none captured)
    uint32_t exec_grid;        // its iteration-kernel grid
};

void ag_ppr_free_ws(ag_dgraph *g)
{
    ag_ppr_ws *w = g->ppr;
    if (!w)
        return;
    (void)hipStreamSynchronize(g->dev->stream);
    if (w->exec)
        (void)hipGraphExecDestroy(w->exec);
    if (w->graph)
        (void)hipGraphDestroy(w->graph);
    void *ptrs[] = {w->c16[0], w->c16[1], w->c32[0], w->c32[1], w->rank, w->hub_part, w->part,
                    w->colerr, w->hist, w->st, w->conf};
    for (void *p : ptrs)
        if (p)
            (void)hipFree(p);
    if (w->st_host)
        (void)hipHostFree(w->st_host);
    if (w->conf_host)
        (void)hipHostFree(w->conf_host);
    free(w);
    g->ppr = NULL;
}

// Allocates the workspace on first use; grows the history (dropping the captured solve) when a
// call asks for more iterations than it holds.
static int ppr_workspace(ag_dgraph *g, uint32_t maxit, ag_ppr_ws **out)
{
    ag_ppr_ws *w = g->ppr;
    if (!w) {
        w = (ag_ppr_ws *)calloc(1, sizeof(ag_ppr_ws));
        if (!w)
            return ag_set_error(AG_ERR_OOM, "host allocation failed");
        g->ppr = w;
        w->n = g->n;
        // 16-lane groups with up to 16 contribution lines in flight per lane: as many
        // workgroups per CU as fit, up to 8
        w->grid_max = ag_pull_grid(g, PPR_GROUPS * PPR_GL * 16u + PPR_GROUPS * PPR_W * 8u, 8u);
        if (w->grid_max == 0)
            w->grid_max = 1;
        w->grid = w->grid_max;
        const size_t nv = (size_t)g->n * PPR_W;
        AG_HIP(hipMalloc(&w->rank, nv * sizeof(float)));
        AG_HIP(hipMalloc((void **)&w->c32[0], nv * sizeof(float)));
        AG_HIP(hipMalloc((void **)&w->c32[1], nv * sizeof(float)));
        AG_HIP(hipMalloc(&w->c16[0], nv * 2));
        AG_HIP(hipMalloc(&w->c16[1], nv * 2));
        AG_HIP(hipMalloc((void **)&w->hub_part,
                         (size_t)(g->n_segs ? g->n_segs : 1) * PPR_W * sizeof(float)));
        AG_HIP(hipMalloc((void **)&w->part, (size_t)w->grid_max * PPR_W * sizeof(double)));
        AG_HIP(hipMalloc((void **)&w->colerr, PPR_W * sizeof(double)));
        AG_HIP(hipMalloc((void **)&w->st, sizeof(PrState)));
        AG_HIP(hipMalloc((void **)&w->conf, sizeof(PprConf)));
        AG_HIP(hipHostMalloc((void **)&w->st_host, sizeof(PrState), hipHostMallocDefault));
        AG_HIP(hipHostMalloc((void **)&w->conf_host, sizeof(PprConf), hipHostMallocDefault));
    }
    if (w->hist_cap < maxit) {
        if (w->exec) {
            (void)hipGraphExecDestroy(w->exec);
            (void)hipGraphDestroy(w->graph);
            w->exec = NULL;
            w->graph = NULL;
            w->exec_maxit = 0;
        }
        if (w->hist)
            AG_HIP(hipFree(w->hist));
        w->hist = NULL;
        w->hist_cap = 0;
        const uint32_t cap = maxit > 512u ? maxit : 512u;
        AG_HIP(hipMalloc((void **)&w->hist, (size_t)cap * sizeof(double)));
        w->hist_cap = cap;
    }
    *out = w;
    return AG_OK;
}

static PprArgs ppr_args(const ag_dgraph *g, const ag_ppr_ws *w)
{
    PprArgs A;
    memset(&A, 0, sizeof(A));
    A.c16[0] = w->c16[0];
    A.c16[1] = w->c16[1];
    A.c32[0] = w->c32[0];
    A.c32[1] = w->c32[1];
    A.rank = w->rank;
    A.outdeg = g->outdeg;
    A.conf = w->conf;
    A.hub_part = w->hub_part;
    A.part = w->part;
    A.colerr = w->colerr;
    A.hist = w->hist;
    A.st = w->st;
    A.nparts = w->grid;
    A.n = g->n;
    A.damp = 0.85f;
    volatile float d = 0.85f;
    A.base = 1.0f - d;
    A.basen = A.base * (float)g->n;
    A.dn = (double)g->n;
    return A;
}

static hipError_t ppr_enqueue_init(const ag_dgraph *g, const PprArgs &A, hipStream_t s)
{
    const uint64_t tot = (uint64_t)g->n * PPR_GL;
    ppr_init_kernel<<<(uint32_t)((tot + 255) / 256), 256, 0, s>>>(A);
    return hipGetLastError();
}

static hipError_t ppr_enqueue_iterations(const ag_dgraph *g, const ag_ppr_ws *w,
                                         const PprArgs &A, uint32_t count, hipStream_t s)
{
    const PullGeom G = ag_make_geom(g);
    for (uint32_t i = 0; i < count; i++) {
        ppr_kernel<<<w->grid, AG_WG, 0, s>>>(G, A);
        ppr_finish_kernel<<<1, AG_FIN_WG, 0, s>>>(G, A);
    }
    return hipGetLastError();
}

extern "C" void ag_ppr_options_default(ag_ppr_options *o)
{
    memset(o, 0, sizeof(*o));
    o->max_iterations = 100;
    o->stop_rule = AG_STOP_L1;
    o->l1_tolerance = 1e-6;
    o->storage = AG_STORE_BF16;
    o->refine_l1 = 1e-3;
    o->chunk = 0;
    o->use_hip_graph = 1;
}

// A solve is: the init launch, then max_iterations x (gather/update, finish); the finish
// launch of the stopping iteration sets a device flag that turns every later launch into a
// no-op. With use_hip_graph the whole solve is captured once per graph and iteration budget
// and replayed with one hipGraphLaunch and one wait; otherwise it is queued eagerly, `chunk`
// iterations between polls of the flag.
extern "C" int ag_ppr_pull(ag_dgraph *g, uint32_t k, const uint32_t *seeds,
                           const ag_ppr_options *opts, float *ranks_out, double *col_error,
                           ag_ppr_result *res)
{
    if (!g || k == 0 || k > AG_PPR_MAX_COLS || !seeds || (g->n && !ranks_out) || !res)
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
    for (uint32_t c = 0; c < k; c++) {
        if (seeds[c] != AG_PPR_UNIFORM && seeds[c] >= n)
            return ag_set_error(AG_ERR_INVALID, "seed %u of column %u out of range", seeds[c], c);
    }
    memset(res, 0, sizeof(*res));
    res->refine_iteration = -1;
    if (n == 0)
        return AG_OK;
    AG_HIP(hipSetDevice(g->dev->ordinal));
    hipStream_t s = g->dev->stream;
    const uint32_t maxit = o.max_iterations ? o.max_iterations : 1;
    ag_ppr_ws *w = NULL;
    int st = ppr_workspace(g, maxit, &w);
    if (st != AG_OK)
        return st;
    st = ag_wait_stream(g->dev);   // the pinned option / state buffers may still feed a copy
    if (st != AG_OK)
        return st;
    // seed vertex -> device row; unused columns run as uniform restarts and are not reported
    PprConf *conf = w->conf_host;
    memset(conf, 0, sizeof(*conf));
    for (uint32_t c = 0; c < PPR_W; c++)
        conf->seed[c] = AG_PPR_UNIFORM;
    {
        uint32_t *o2d = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
        if (!o2d)
            return ag_set_error(AG_ERR_OOM, "host allocation failed");
        for (uint32_t r = 0; r < n; r++)
            o2d[g->h_d2o[r]] = r;
        for (uint32_t c = 0; c < k; c++)
            if (seeds[c] != AG_PPR_UNIFORM)
                conf->seed[c] = o2d[seeds[c]];
        free(o2d);
    }
    conf->k = k;
    {
        // every scaled rank is at most N (the mass of a column never exceeds N), so a
        // contribution times 2^-bias stays at most 2^15, inside FP16's range
        int32_t bits = 0;
        while (bits < 32 && (1ull << bits) < (uint64_t)n)
            bits++;
        conf->f16_bias = bits > 15 ? bits - 15 : 0;
    }
    conf->storage = o.storage;
    conf->stop_rule = o.stop_rule;
    conf->l1_tol = o.l1_tolerance;
    conf->refine_l1 = o.refine_l1;
    AG_HIP(hipMemcpyAsync(w->conf, conf, sizeof(PprConf), hipMemcpyHostToDevice, s));
    {
        uint32_t grid = w->grid_max;
        if (o.wg_per_cu) {
            const uint64_t want = (uint64_t)g->dev->num_cus * o.wg_per_cu;
            grid = want < grid ? (uint32_t)want : grid;
        }
        w->grid = grid ? grid : 1u;
    }
    const PprArgs A = ppr_args(g, w);
    hipEvent_t e0 = NULL, e1 = NULL;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess)
        e = hipEventCreate(&e1);
    if (e == hipSuccess)
        e = hipEventRecord(e0, s);
    if (e == hipSuccess && o.use_hip_graph) {
        if (w->exec && (w->exec_maxit != maxit || w->exec_grid != w->grid)) {
            (void)hipGraphExecDestroy(w->exec);
            (void)hipGraphDestroy(w->graph);
            w->exec = NULL;
            w->graph = NULL;
            w->exec_maxit = 0;
        }
        if (!w->exec) {
            e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                hipError_t qe = ppr_enqueue_init(g, A, s);
                if (qe == hipSuccess)
                    qe = ppr_enqueue_iterations(g, w, A, maxit, s);
                hipError_t ce = hipStreamEndCapture(s, &w->graph);
                e = qe != hipSuccess ? qe : ce;
                if (qe != hipSuccess && ce == hipSuccess) {
                    (void)hipGraphDestroy(w->graph);
                    w->graph = NULL;
                }
            }
            if (e == hipSuccess)
                e = hipGraphInstantiate(&w->exec, w->graph, NULL, NULL, 0);
            if (e == hipSuccess) {
                w->exec_maxit = maxit;
                w->exec_grid = w->grid;
            }
        }
        if (e == hipSuccess)
            e = hipGraphLaunch(w->exec, s);
        if (e == hipSuccess)
            e = hipEventRecord(e1, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(w->st_host, w->st, sizeof(PrState), hipMemcpyDeviceToHost, s);
        st = (e == hipSuccess) ? ag_wait_stream(g->dev) : ag_hip_fail(e, "PPR graph launch", __FILE__, __LINE__);
    } else if (e == hipSuccess) {
        e = ppr_enqueue_init(g, A, s);
        const uint32_t chunk = o.chunk ? o.chunk : 8;
        uint32_t enq = 0;
        while (e == hipSuccess && st == AG_OK && enq < maxit) {
            const uint32_t kk = maxit - enq < chunk ? maxit - enq : chunk;
            e = ppr_enqueue_iterations(g, w, A, kk, s);
            enq += kk;
            if (e == hipSuccess)
                e = hipEventRecord(e1, s);
            if (e == hipSuccess)
                e = hipMemcpyAsync(w->st_host, w->st, sizeof(PrState), hipMemcpyDeviceToHost, s);
            if (e == hipSuccess)
                st = ag_wait_stream(g->dev);
            if (st == AG_OK && e == hipSuccess && w->st_host->done)
                break;
        }
        if (e != hipSuccess)
            st = ag_hip_fail(e, "PPR launch", __FILE__, __LINE__);
    } else {
        st = ag_hip_fail(e, "PPR events", __FILE__, __LINE__);
    }
    if (st == AG_OK) {
        const PrState *sth = w->st_host;
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            ms = 0.0f;   // timing is informational; the ranks are already valid
        res->loop_ms = ms;
        res->iterations = sth->done ? sth->stop_iter : sth->iter;
        res->executed = sth->done ? sth->stop_iter + 1 : sth->iter;
        res->max_error = sth->last_err;
        res->refine_iteration = sth->refine_iter;
        // normalised ranks, original order; c32[1] is free as staging once the solve ended
        float *outd = w->c32[1];
        const uint64_t tot = (uint64_t)n * k;
        ppr_out_kernel<<<(uint32_t)((tot + 255) / 256), 256, 0, s>>>(n, k, w->rank, g->d2o,
                                                                     (float)n, outd);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(ranks_out, outd, tot * sizeof(float), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess && col_error)
            e = hipMemcpyAsync(col_error, w->colerr, (size_t)k * sizeof(double), hipMemcpyDeviceToHost, s);
        st = (e == hipSuccess) ? ag_wait_stream(g->dev) : ag_hip_fail(e, "PPR output", __FILE__, __LINE__);
    }
    if (e0)
        (void)hipEventDestroy(e0);
    if (e1)
        (void)hipEventDestroy(e1);
    return st;
}
