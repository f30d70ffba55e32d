// ag_spmv.hip -- SpMV pull and the raw pull gather on the GPU.
//
// SpMV pull (02_capi_graph/src/algorithms/capi/SPMV.c:622-751): y[v] = sum over in-edges
// (u->v) of w(u,v) * x[u]. The FPGA FloatPoint datapath multiplies in binary32 and then
// accumulates in stream order (cu_SPMV/CSR/PULL/FloatPoint/cu/cu_sum_kernel_control.sv:
// 345-367); the FixedPoint datapath adds (u32)(((u64)x * w) >> SCALEF) into a u64 sum with
// SCALEF = 16 (FixedPoint/cu/cu_sum_kernel_control.sv:155-164, FixedPoint/pkg/
// globals_cu_pkg.sv:37). Each call computes y = A x once per repetition (the RTL overwrites
// y every run). Unweighted graphs use the constant weight 0.0001f (SPMV.c:1206). The push
// forms (SPMV.c:752-848, 1017-1136) scatter along the forward CSR and accumulate across
// repetitions, as the reference's output is never cleared there.
//
// The raw gather is the sum kernel itself (cu_PageRank/CSR/PULL/*/cu/cu_sum_kernel_control.sv)
// on caller-supplied per-vertex values; it exists so the gather arithmetic can be checked
// against the reference's kernel-level known-answer fixtures.
#include <string.h>

#include "fixedPoint.h"
#include "ag_internal.h"
#include "ag_fixed_dev.h"
#include "ag_pull_kernel.h"
#include "ag_push.h"

namespace {

template <bool WEIGHTED>
struct SpmvF32Pol {
    using Acc = float;
    const float *x;     // device order
    const float *w;     // [m] or null
    float *y;           // device order
    __device__ __forceinline__ float edge(uint32_t j, uint32_t s) const
    {
        return edge_val(j, x[s]);
    }
    // the same term from x[s] already in hand (hot-segment LDS copy)
    __device__ __forceinline__ float edge_val(uint32_t j, float xs) const
    {
        return __fmul_rn(xs, WEIGHTED ? __builtin_nontemporal_load(&w[j]) : 0.0001f);
    }
    __device__ __forceinline__ void finish(uint32_t r, float sum, Red &) const { y[r] = sum; }
};

// Hot pass of the FP32 SpMV (see pull_hot_items): x gathered from LDS copies of its hot
// segments, per-edge weights in virtual-row order.
template <bool WEIGHTED>
__global__ __launch_bounds__(AG_WG) void spmv_hot_kernel(BlockGeom B, SpmvF32Pol<WEIGHTED> pol)
{
    __shared__ uint4 s_tab[AG_HOT_TAB_BYTES / 16];
    __shared__ float s_vals[AG_HOT_NNZ];
    __shared__ uint32_t s_long[AG_LONG_MAX];
    __shared__ uint32_t s_nlong;
    pull_hot_items(B, pol, pol.x, s_tab, s_vals, s_long, &s_nlong);
}

template <bool WEIGHTED>
struct SpmvFixPol {
    using Acc = uint64_t;
    const uint32_t *xq;
    const uint32_t *wq;
    uint32_t wconst;
    uint64_t *y;
    __device__ __forceinline__ uint64_t edge(uint32_t j, uint32_t s) const
    {
        const uint64_t wide =
            (uint64_t)xq[s] * (uint64_t)(WEIGHTED ? __builtin_nontemporal_load(&wq[j]) : wconst);
        return (uint64_t)(uint32_t)(wide >> 16);
    }
    __device__ __forceinline__ void finish(uint32_t r, uint64_t sum, Red &) const { y[r] = sum; }
};

template <typename T>
__global__ void permute_in_kernel(uint32_t n, const T *orig, const uint32_t *d2o, T *dev)
{
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n)
        dev[r] = orig[d2o[r]];
}

template <typename T>
__global__ void permute_out_kernel(uint32_t n, const T *dev, const uint32_t *d2o, T *orig)
{
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n)
        orig[d2o[r]] = dev[r];
}

// SpMV push (SPMV.c:752-848, 1017-1136): every source adds w * x[src] into its out-edges'
// destinations. The reference never clears the output between iterations, so repetitions
// accumulate (y = reps * A x). Propagation-blocked (ag_push.h): phase 1 writes every
// out-edge's product into its bin slot, phase 2 adds each bin's slots per row in LDS and adds
// the row's sum to y. FP32: binary32 products, summed in the LDS atomics' order (a hub row's
// bin: a fixed tree). Fixed point: Q32.32 x and weights, MULFixed64V1 products, exact u64 sums.
template <bool WEIGHTED>
__global__ __launch_bounds__(AG_WG) void spmv_push_scatter_kernel(uint32_t m, const uint32_t *fsrc,
                                                                  const uint32_t *pos,
                                                                  const float *fw, const float *x,
                                                                  float *vals)
{
    pb_scatter(m, pos, vals, [&](uint32_t j) {
        const float w = WEIGHTED ? __builtin_nontemporal_load(&fw[j]) : 0.0001f;
        return __fmul_rn(w, x[__builtin_nontemporal_load(&fsrc[j])]);
    });
}

template <bool WEIGHTED>
__global__ __launch_bounds__(AG_WG) void spmv_push_fixed_scatter_kernel(
    uint32_t m, const uint32_t *fsrc, const uint32_t *pos, const float *fw, const uint64_t *xq,
    uint64_t wc, uint64_t *vals)
{
    pb_scatter(m, pos, vals, [&](uint32_t j) {
        const uint64_t w =
            WEIGHTED ? ag_dev_double_to_fixed64((double)__builtin_nontemporal_load(&fw[j])) : wc;
        return ag_dev_mul_fixed64(w, xq[__builtin_nontemporal_load(&fsrc[j])]);
    });
}

__global__ __launch_bounds__(AG_WG) void spmv_push_bins_kernel(PushBins P, const float *vals,
                                                               float *y)
{
    __shared__ float s_acc[AG_PB_ROWS];
    __shared__ float s_wave[AG_WG / AG_WAVE];
    pb_bins(P, vals, s_acc, s_wave, [&](uint32_t r, float sum) { y[r] = y[r] + sum; });
}

__global__ __launch_bounds__(AG_WG) void spmv_push_fixed_bins_kernel(PushBins P,
                                                                     const uint64_t *vals,
                                                                     uint64_t *y)
{
    __shared__ unsigned long long s_acc[AG_PB_ROWS];
    __shared__ unsigned long long s_wave[AG_WG / AG_WAVE];
    pb_bins(P, vals, s_acc, s_wave,
            [&](uint32_t r, unsigned long long sum) { y[r] += (uint64_t)sum; });
}

__global__ __launch_bounds__(AG_WG) void zero_u32_kernel(uint32_t *p, uint32_t count)
{
    for (uint32_t i = blockIdx.x * AG_WG + threadIdx.x; i < count; i += gridDim.x * AG_WG)
        p[i] = 0u;
}

__global__ void quantize_weights_kernel(uint32_t m, const float *w, uint32_t *wq)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < m)
        wq[j] = ag_dev_float_to_fixed32(w[j]);
}

template <typename VT, typename A>
struct Gather {
    using Acc = A;
    const VT *v;
    A *out;
    __device__ __forceinline__ A edge(uint32_t, uint32_t s) const { return (A)v[s]; }
    __device__ __forceinline__ void finish(uint32_t r, A sum, Red &) const { out[r] = sum; }
};

} // namespace

// Timed launches of one pull policy `reps` times; *avg_ms is the mean launch duration from
// HIP events on the device stream.
template <class P>
static int run_pull(ag_dgraph *g, const P &pol, uint32_t reps, double *avg_ms)
{
    hipStream_t stream = g->dev->stream;
    PullGeom G = ag_make_geom(g);
    hipEvent_t e0 = NULL, e1 = NULL;
    if (avg_ms) {
        AG_HIP(hipEventCreate(&e0));
        AG_HIP(hipEventCreate(&e1));
        AG_HIP(hipEventRecord(e0, stream));
    }
    const uint32_t grid = ag_pull_grid(g, AG_TILE_NNZ_MAX * (uint32_t)sizeof(typename P::Acc));
    for (uint32_t i = 0; i < reps && g->total_blocks; i++) {
        pull_kernel<P><<<grid, AG_WG, 0, stream>>>(G, pol);
        if (g->n_hubs)
            pull_finish_kernel<P><<<1, AG_FIN_WG, 0, stream>>>(G, pol);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && avg_ms)
        e = hipEventRecord(e1, stream);
    int st = (e == hipSuccess) ? ag_wait_stream(g->dev) : ag_hip_fail(e, "pull launch", __FILE__, __LINE__);
    if (st == AG_OK && avg_ms) {
        float ms = 0.0f;
        e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess)
            st = ag_hip_fail(e, "hipEventElapsedTime", __FILE__, __LINE__);
        *avg_ms = reps ? (double)ms / reps : 0.0;
    }
    if (e0)
        (void)hipEventDestroy(e0);
    if (e1)
        (void)hipEventDestroy(e1);
    return st;
}

// The hot pass of a blocked SpMV: FP32 only (the fixed-point form gathers everything in
// pass 1). Returns whether it was launched.
template <class P>
static bool spmv_hot_launch(const ag_dgraph *, const BlockGeom &, const P &, hipStream_t)
{
    return false;
}
template <bool WEIGHTED>
static bool spmv_hot_launch(const ag_dgraph *g, const BlockGeom &B,
                            const SpmvF32Pol<WEIGHTED> &pv, hipStream_t stream)
{
    if (!g->blk.n_hgrp)
        return false;
    static const uint32_t res = ag_resident_per_cu(spmv_hot_kernel<WEIGHTED>);
    spmv_hot_kernel<WEIGHTED><<<ag_blocked_grid(g, res, 2u), AG_WG, 0, stream>>>(B, pv);
    return true;
}

// Source-blocked form of run_pull: `pv` reads per-edge data in virtual-row order, `pu` in
// regular order (unblocked rows); the hot pass (FP32), pass 1 then the combine pass per
// repetition.
template <class P>
static int run_pull_blocked(ag_dgraph *g, const P &pv, const P &pu, uint32_t reps,
                            double *avg_ms)
{
    hipStream_t stream = g->dev->stream;
    const PullGeom U = ag_make_unblocked_geom(g);
    const BlockGeom B = ag_make_block_geom(g);
    BlockGeom B1 = B;   // pass 1: past the hot tiles when the hot pass gathers them
    static const uint32_t res1 = ag_resident_per_cu(pull_blocked_kernel<P, P>);
    static const uint32_t res2 = ag_resident_per_cu(pull_combine_kernel<P>);
    const uint32_t g1 = ag_blocked_grid(g, res1, 4u), g2 = ag_blocked_grid(g, res2, 8u);
    hipEvent_t e0 = NULL, e1 = NULL;
    if (avg_ms) {
        AG_HIP(hipEventCreate(&e0));
        AG_HIP(hipEventCreate(&e1));
        AG_HIP(hipEventRecord(e0, stream));
    }
    for (uint32_t i = 0; i < reps; i++) {
        if (spmv_hot_launch(g, B, pv, stream))
            ag_block_geom_skip_hot(g, B1);
        pull_blocked_kernel<P, P><<<g1, AG_WG, 0, stream>>>(U, B1, pv, pu);
        pull_combine_kernel<P><<<g2, AG_WG, 0, stream>>>(B, pu);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && avg_ms)
        e = hipEventRecord(e1, stream);
    int st = (e == hipSuccess) ? ag_wait_stream(g->dev)
                               : ag_hip_fail(e, "blocked pull launch", __FILE__, __LINE__);
    if (st == AG_OK && avg_ms) {
        float ms = 0.0f;
        e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess)
            st = ag_hip_fail(e, "hipEventElapsedTime", __FILE__, __LINE__);
        *avg_ms = reps ? (double)ms / reps : 0.0;
    }
    if (e0)
        (void)hipEventDestroy(e0);
    if (e1)
        (void)hipEventDestroy(e1);
    return st;
}

// One FP32 SpMV on device buffers: x (host, original order) -> xo -> permuted into xd (device
// order) -> y in device order into yd -> permuted back into xo -> y (host). yd may be xo (x
// in original order is dead once permuted), never xd.
static int spmv_f32_device(ag_dgraph *g, const float *x, float *y, float *xo, float *xd,
                           float *yd, uint32_t reps, double *avg_ms)
{
    const uint32_t n = g->n;
    hipStream_t stream = g->dev->stream;
    const uint32_t blk = (n + AG_WG - 1) / AG_WG;
    AG_HIP(hipMemcpyAsync(xo, x, (size_t)n * sizeof(float), hipMemcpyHostToDevice, stream));
    permute_in_kernel<float><<<blk, AG_WG, 0, stream>>>(n, xo, g->d2o, xd);
    AG_HIP(hipGetLastError());
    int st;
    if (g->blk.bits && g->w)
        st = run_pull_blocked(g, SpmvF32Pol<true>{xd, g->blk.vw, yd},
                              SpmvF32Pol<true>{xd, g->w, yd}, reps, avg_ms);
    else if (g->blk.bits)
        st = run_pull_blocked(g, SpmvF32Pol<false>{xd, NULL, yd}, SpmvF32Pol<false>{xd, NULL, yd},
                              reps, avg_ms);
    else if (g->w)
        st = run_pull(g, SpmvF32Pol<true>{xd, g->w, yd}, reps, avg_ms);
    else
        st = run_pull(g, SpmvF32Pol<false>{xd, NULL, yd}, reps, avg_ms);
    if (st != AG_OK)
        return st;
    float *yo = (yd == xo) ? xd : xo;   // y in original order
    permute_out_kernel<float><<<blk, AG_WG, 0, stream>>>(n, yd, g->d2o, yo);
    AG_HIP(hipGetLastError());
    AG_HIP(hipMemcpyAsync(y, yo, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, stream));
    return ag_wait_stream(g->dev);
}

extern "C" int ag_spmv_pull(ag_dgraph *g, const float *x, float *y, uint32_t reps, double *avg_ms)
{
    if (!g || (g->n && (!x || !y)))
        return ag_set_error(AG_ERR_INVALID, "null argument");
    if (avg_ms)
        *avg_ms = 0.0;
    const uint32_t n = g->n;
    if (n == 0)
        return AG_OK;
    if (reps == 0)
        reps = 1;
    AG_HIP(hipSetDevice(g->dev->ordinal));
    int st = ag_ensure_vec(g, (size_t)n * sizeof(uint64_t));
    if (st != AG_OK)
        return st;
    float *xo = (float *)g->vec_a;                    // x, original order (then y original)
    float *xd = (float *)g->vec_b;                    // x, device order
    float *yd = (float *)((char *)g->vec_b + (size_t)n * sizeof(float));  // y, device order
    return spmv_f32_device(g, x, y, xo, xd, yd, reps, avg_ms);
}

extern "C" int ag_spmv_pull_noalloc(ag_dgraph *g, const float *x, float *y)
{
    if (!g || (g->n && (!x || !y)))
        return ag_set_error(AG_ERR_INVALID, "null argument");
    const uint32_t n = g->n;
    if (n == 0)
        return AG_OK;
    if ((uint64_t)n * 2u * sizeof(float) > (uint64_t)AG_NOALLOC_SCRATCH_BYTES)
        return ag_set_error(AG_ERR_INVALID, "ag_spmv_pull_noalloc: %u vertices need %llu bytes "
                            "of scratch, more than its fixed %u (at most %u vertices)", n,
                            (unsigned long long)n * 2u * sizeof(float),
                            (unsigned)AG_NOALLOC_SCRATCH_BYTES,
                            (unsigned)(AG_NOALLOC_SCRATCH_BYTES / (2u * sizeof(float))));
    if (!g->dev->scratch)
        return ag_set_error(AG_ERR_UNSUPPORTED, "ag_spmv_pull_noalloc: no scratch on device %d",
                            g->dev->ordinal);
    AG_HIP(hipSetDevice(g->dev->ordinal));
    // two halves of the scratch: x in original order, then (once permuted) y in device order;
    // x in device order, then y in original order
    float *a = (float *)g->dev->scratch;
    float *b = a + n;
    return spmv_f32_device(g, x, y, a, b, a, 1, NULL);
}

static uint32_t host_fixed32(float x)
{
    double s = (double)x * 65536.0 + 0.5;
    if (!(s >= 1.0))
        return 0;
    if (s >= 4294967296.0)
        return 0xFFFFFFFFu;
    return (uint32_t)s;
}

extern "C" int ag_spmv_pull_fixed(ag_dgraph *g, const uint32_t *xq, uint64_t *yq, uint32_t reps,
                                  double *avg_ms)
{
    if (!g || (g->n && (!xq || !yq)))
        return ag_set_error(AG_ERR_INVALID, "null argument");
    if (avg_ms)
        *avg_ms = 0.0;
    const uint32_t n = g->n;
    if (n == 0)
        return AG_OK;
    if (reps == 0)
        reps = 1;
    AG_HIP(hipSetDevice(g->dev->ordinal));
    int st = ag_ensure_vec(g, (size_t)n * sizeof(uint64_t) * 2);
    if (st != AG_OK)
        return st;
    hipStream_t stream = g->dev->stream;
    if (g->w && !g->wq && g->m) {
        AG_HIP(hipMalloc((void **)&g->wq, (size_t)g->m * sizeof(uint32_t)));
        g->device_bytes += (size_t)g->m * sizeof(uint32_t);
        quantize_weights_kernel<<<(g->m + AG_WG - 1) / AG_WG, AG_WG, 0, stream>>>(g->m, g->w, g->wq);
        AG_HIP(hipGetLastError());
    }
    // Blocked graphs read weight codes in virtual-row order: quantised from vw. Codes set by
    // the caller exist only in regular order, so such a graph keeps the regular schedule.
    const bool blocked = g->blk.bits && !g->wq_user;
    if (blocked && g->w && !g->vwq && g->blk.mb) {
        AG_HIP(hipMalloc((void **)&g->vwq, (size_t)g->blk.mb * sizeof(uint32_t)));
        g->device_bytes += (size_t)g->blk.mb * sizeof(uint32_t);
        quantize_weights_kernel<<<(g->blk.mb + AG_WG - 1) / AG_WG, AG_WG, 0, stream>>>(
            g->blk.mb, g->blk.vw, g->vwq);
        AG_HIP(hipGetLastError());
    }
    uint32_t *xo = (uint32_t *)g->vec_a;
    uint64_t *yo = (uint64_t *)g->vec_a;              // reused after x is permuted
    uint32_t *xd = (uint32_t *)g->vec_b;
    uint64_t *yd = (uint64_t *)((char *)g->vec_b + (size_t)n * sizeof(uint64_t));
    const uint32_t blk = (n + AG_WG - 1) / AG_WG;
    AG_HIP(hipMemcpyAsync(xo, xq, (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
    permute_in_kernel<uint32_t><<<blk, AG_WG, 0, stream>>>(n, xo, g->d2o, xd);
    AG_HIP(hipGetLastError());
    const uint32_t wc = host_fixed32(0.0001f);
    if (blocked && g->wq)
        st = run_pull_blocked(g, SpmvFixPol<true>{xd, g->vwq, 0u, yd},
                              SpmvFixPol<true>{xd, g->wq, 0u, yd}, reps, avg_ms);
    else if (blocked)
        st = run_pull_blocked(g, SpmvFixPol<false>{xd, NULL, wc, yd},
                              SpmvFixPol<false>{xd, NULL, wc, yd}, reps, avg_ms);
    else if (g->wq)
        st = run_pull(g, SpmvFixPol<true>{xd, g->wq, 0u, yd}, reps, avg_ms);
    else
        st = run_pull(g, SpmvFixPol<false>{xd, NULL, wc, yd}, reps, avg_ms);
    if (st != AG_OK)
        return st;
    permute_out_kernel<uint64_t><<<blk, AG_WG, 0, stream>>>(n, yd, g->d2o, yo);
    AG_HIP(hipGetLastError());
    AG_HIP(hipMemcpyAsync(yq, yo, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
    return ag_wait_stream(g->dev);
}

// reps propagation-blocked push passes into yd (ywords u32 words, zeroed first) through a
// bin buffer of m values of vbytes each, timed with HIP events. launch(p1, p2, vals) enqueues
// one pass with phase-1 and phase-2 grids p1 and p2.
template <class L>
static int run_push(ag_dgraph *g, uint32_t reps, double *avg_ms, void *yd, uint32_t ywords,
                    size_t vbytes, const L &launch)
{
    hipStream_t stream = g->dev->stream;
    void *vals = NULL;
    AG_HIP(hipMalloc(&vals, (g->m ? g->m : 1u) * vbytes));
    const uint32_t zb = (ywords + AG_WG - 1) / AG_WG;
    zero_u32_kernel<<<zb < 4096u ? (zb ? zb : 1u) : 4096u, AG_WG, 0, stream>>>((uint32_t *)yd,
                                                                               ywords);
    hipEvent_t e0 = NULL, e1 = NULL;
    hipError_t e = hipSuccess;
    if (avg_ms) {
        e = hipEventCreate(&e0);
        if (e == hipSuccess)
            e = hipEventCreate(&e1);
        if (e == hipSuccess)
            e = hipEventRecord(e0, stream);
    }
    const uint32_t p1 = (g->dev->num_cus * 16u + AG_NXCD - 1u) / AG_NXCD * AG_NXCD;
    const uint32_t p2 = g->pb_nbins < g->dev->num_cus * 8u ? g->pb_nbins : g->dev->num_cus * 8u;
    for (uint32_t i = 0; e == hipSuccess && i < reps && g->m; i++)
        launch(p1, p2, vals);
    if (e == hipSuccess)
        e = hipGetLastError();
    if (e == hipSuccess && avg_ms)
        e = hipEventRecord(e1, stream);
    int st = (e == hipSuccess) ? ag_wait_stream(g->dev) : ag_hip_fail(e, "push launch", __FILE__, __LINE__);
    if (st == AG_OK && avg_ms) {
        float ms = 0.0f;
        e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess)
            st = ag_hip_fail(e, "hipEventElapsedTime", __FILE__, __LINE__);
        *avg_ms = reps ? (double)ms / reps : 0.0;
    }
    if (e0)
        (void)hipEventDestroy(e0);
    if (e1)
        (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(stream);
    (void)hipFree(vals);
    return st;
}

extern "C" int ag_spmv_push(ag_dgraph *g, const float *x, float *y, uint32_t reps, double *avg_ms)
{
    if (!g || (g->n && (!x || !y)))
        return ag_set_error(AG_ERR_INVALID, "null argument");
    if (avg_ms)
        *avg_ms = 0.0;
    const uint32_t n = g->n;
    if (n == 0)
        return AG_OK;
    AG_HIP(hipSetDevice(g->dev->ordinal));
    int st = ag_graph_push_bins(g);
    if (st == AG_OK)
        st = ag_ensure_vec(g, (size_t)n * sizeof(uint64_t));
    if (st != AG_OK)
        return st;
    hipStream_t stream = g->dev->stream;
    float *xo = (float *)g->vec_a;
    float *xd = (float *)g->vec_b;
    float *yd = (float *)((char *)g->vec_b + (size_t)n * sizeof(float));
    const uint32_t blk = (n + AG_WG - 1) / AG_WG;
    AG_HIP(hipMemcpyAsync(xo, x, (size_t)n * sizeof(float), hipMemcpyHostToDevice, stream));
    permute_in_kernel<float><<<blk, AG_WG, 0, stream>>>(n, xo, g->d2o, xd);
    AG_HIP(hipGetLastError());
    const uint32_t m = g->m;
    const float *fw = g->fwd_w;
    const PushBins P{g->pb_nbins, g->pb_row, g->pb_off, g->pb_pos, g->pb_dst};
    st = run_push(g, reps, avg_ms, yd, n, sizeof(float), [&](uint32_t p1, uint32_t p2, void *v) {
        if (fw)
            spmv_push_scatter_kernel<true><<<p1, AG_WG, 0, stream>>>(m, g->fwd_src, g->pb_pos, fw,
                                                                     xd, (float *)v);
        else
            spmv_push_scatter_kernel<false><<<p1, AG_WG, 0, stream>>>(m, g->fwd_src, g->pb_pos,
                                                                      NULL, xd, (float *)v);
        spmv_push_bins_kernel<<<p2, AG_WG, 0, stream>>>(P, (const float *)v, yd);
    });
    if (st != AG_OK)
        return st;
    permute_out_kernel<float><<<blk, AG_WG, 0, stream>>>(n, yd, g->d2o, xo);
    AG_HIP(hipGetLastError());
    AG_HIP(hipMemcpyAsync(y, xo, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, stream));
    return ag_wait_stream(g->dev);
}

extern "C" int ag_spmv_push_fixed(ag_dgraph *g, const uint64_t *xq, uint64_t *yq, uint32_t reps,
                                  double *avg_ms)
{
    if (!g || (g->n && (!xq || !yq)))
        return ag_set_error(AG_ERR_INVALID, "null argument");
    if (avg_ms)
        *avg_ms = 0.0;
    const uint32_t n = g->n;
    if (n == 0)
        return AG_OK;
    AG_HIP(hipSetDevice(g->dev->ordinal));
    int st = ag_graph_push_bins(g);
    if (st == AG_OK)
        st = ag_ensure_vec(g, (size_t)n * sizeof(uint64_t) * 2);
    if (st != AG_OK)
        return st;
    hipStream_t stream = g->dev->stream;
    uint64_t *xo = (uint64_t *)g->vec_a;
    uint64_t *xd = (uint64_t *)g->vec_b;
    uint64_t *yd = (uint64_t *)((char *)g->vec_b + (size_t)n * sizeof(uint64_t));
    const uint32_t blk = (n + AG_WG - 1) / AG_WG;
    AG_HIP(hipMemcpyAsync(xo, xq, (size_t)n * sizeof(uint64_t), hipMemcpyHostToDevice, stream));
    permute_in_kernel<uint64_t><<<blk, AG_WG, 0, stream>>>(n, xo, g->d2o, xd);
    AG_HIP(hipGetLastError());
    const uint32_t m = g->m;
    const float *fw = g->fwd_w;
    const uint64_t wc = DoubleToFixed64((double)0.0001f);   // SPMV.c:1080
    const PushBins P{g->pb_nbins, g->pb_row, g->pb_off, g->pb_pos, g->pb_dst};
    st = run_push(g, reps, avg_ms, yd, 2u * n, sizeof(uint64_t),
                  [&](uint32_t p1, uint32_t p2, void *v) {
        if (fw)
            spmv_push_fixed_scatter_kernel<true><<<p1, AG_WG, 0, stream>>>(
                m, g->fwd_src, g->pb_pos, fw, xd, 0ull, (uint64_t *)v);
        else
            spmv_push_fixed_scatter_kernel<false><<<p1, AG_WG, 0, stream>>>(
                m, g->fwd_src, g->pb_pos, NULL, xd, wc, (uint64_t *)v);
        spmv_push_fixed_bins_kernel<<<p2, AG_WG, 0, stream>>>(P, (const uint64_t *)v, yd);
    });
    if (st != AG_OK)
        return st;
    permute_out_kernel<uint64_t><<<blk, AG_WG, 0, stream>>>(n, yd, g->d2o, xo);
    AG_HIP(hipGetLastError());
    AG_HIP(hipMemcpyAsync(yq, xo, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
    return ag_wait_stream(g->dev);
}

template <typename VT, typename A>
static int gather_typed(ag_dgraph *g, const void *values, void *out)
{
    const uint32_t n = g->n;
    hipStream_t stream = g->dev->stream;
    int st = ag_ensure_vec(g, (size_t)n * sizeof(uint64_t) * 2);
    if (st != AG_OK)
        return st;
    VT *vo = (VT *)g->vec_a;
    A *ao = (A *)g->vec_a;
    VT *vd = (VT *)g->vec_b;
    A *ad = (A *)((char *)g->vec_b + (size_t)n * sizeof(uint64_t));
    const uint32_t blk = (n + AG_WG - 1) / AG_WG;
    AG_HIP(hipMemcpyAsync(vo, values, (size_t)n * sizeof(VT), hipMemcpyHostToDevice, stream));
    permute_in_kernel<VT><<<blk, AG_WG, 0, stream>>>(n, vo, g->d2o, vd);
    AG_HIP(hipGetLastError());
    st = run_pull(g, Gather<VT, A>{vd, ad}, 1, NULL);
    if (st != AG_OK)
        return st;
    permute_out_kernel<A><<<blk, AG_WG, 0, stream>>>(n, ad, g->d2o, ao);
    AG_HIP(hipGetLastError());
    AG_HIP(hipMemcpyAsync(out, ao, (size_t)n * sizeof(A), hipMemcpyDeviceToHost, stream));
    return ag_wait_stream(g->dev);
}

extern "C" int ag_pull_gather(ag_dgraph *g, uint32_t type, const void *values, void *out)
{
    if (!g || (g->n && (!values || !out)))
        return ag_set_error(AG_ERR_INVALID, "null argument");
    if (g->n == 0)
        return AG_OK;
    AG_HIP(hipSetDevice(g->dev->ordinal));
    switch (type) {
    case AG_GATHER_F32: return gather_typed<float, float>(g, values, out);
    case AG_GATHER_U64: return gather_typed<uint64_t, uint64_t>(g, values, out);
    case AG_GATHER_U32: return gather_typed<uint32_t, uint64_t>(g, values, out);
    case AG_GATHER_U16: return gather_typed<uint16_t, uint64_t>(g, values, out);
    case AG_GATHER_U8: return gather_typed<uint8_t, uint64_t>(g, values, out);
    default: return ag_set_error(AG_ERR_INVALID, "unknown gather type %u", type);
    }
}
