// ag_pagerank.hip -- PageRank pull on the GPU.
//
// Reference host loop (02_capi_graph/src/algorithms/capi/pageRank.c:942-1109), per iteration:
//   contrib[v] = out_degree[v] ? rank[v] / out_degree[v] : 0          (:1026-1033)
//   next[v]    = sum over in-edges (u->v) of contrib[u]                (FPGA gather, :1041-1046)
//   rank[v]    = base_pr + damp * next[v]; err += |delta| / N;
//   active    += |delta| >= epsilon                                    (:1055-1069)
//   stop when active == 0                                              (:1072-1074)
// and finally rank[v] /= N with an FP64 sum (:1079-1085).
//
// One iteration is either
//   regular:        pr_*_kernel (gather + update for every row) then pr_*_finish, or
//   source-blocked: pr_*_blk_kernel (per-block partial sums, XCD-sliced) then pr_*_cmb_kernel
//                   (partials added per row in block order + update) then pr_*_finish.
// The update is fused into the launch that completes a row's sum: the damping update, the
// per-workgroup L1 / active partials and the next iteration's contribution; `next` never
// exists in memory. Contributions ping-pong between two buffers by iteration parity. The
// one-workgroup finish kernel adds the per-workgroup partials in a fixed order, records the
// iteration's error and active count, decides the 16-bit -> FP32 refinement and sets a done
// flag that turns every later launch into a no-op, so many iterations can be queued without
// a host round trip per iteration.
//
// Fixed-point variants (pageRank.c:1276-1990) gather u64/u32/u16/u8 contributions into an
// exact u64 sum (wraps modulo 2^64) and convert with the formats of fixedPoint.h. Quantised
// variants (pageRank.c:2119-2676) gather u32/u16/u8 codes of a per-iteration scale into an
// exact u64 sum; a small pr_quant_kernel after each finish writes the next codes.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <type_traits>

#include "ag_internal.h"
#include "ag_fixed_dev.h"
#include "ag_pull_kernel.h"
#include "ag_push.h"

// Accumulator of the data-driven pull on a source-blocked plan: the FP32 sum of the in-edges
// and whether any in-neighbour took a new rank in the previous iteration (0 or 1, reduced
// with max). Sums are added exactly as the FP32 pull adds them.
struct DdAcc {
    float s;
    float f;
    DdAcc() = default;
    __device__ __forceinline__ DdAcc(int) : s(0.0f), f(0.0f) {}
    __device__ __forceinline__ DdAcc(float s_, float f_) : s(s_), f(f_) {}
};
__device__ __forceinline__ DdAcc operator+(DdAcc a, DdAcc b)
{
    return DdAcc(__fadd_rn(a.s, b.s), fmaxf(a.f, b.f));
}
template <>
__device__ __forceinline__ DdAcc ag_shfl_xor<DdAcc>(DdAcc v, int mask, int width)
{
    return DdAcc(__shfl_xor(v.s, mask, width), __shfl_xor(v.f, mask, width));
}
template <>
__device__ __forceinline__ uint64_t acc_bits<DdAcc>(DdAcc v)
{
    return (uint64_t)__float_as_uint(v.s) | ((uint64_t)__float_as_uint(v.f) << 32);
}
template <>
__device__ __forceinline__ DdAcc acc_from_bits<DdAcc>(uint64_t b)
{
    return DdAcc(__uint_as_float((uint32_t)b), __uint_as_float((uint32_t)(b >> 32)));
}

namespace {

// ---------------------------------------------------------------- 16-bit storage

struct bf16s {
    uint16_t bits;
};

__device__ __forceinline__ float ld_f32(float v) { return v; }
__device__ __forceinline__ float ld_f32(bf16s v) { return __uint_as_float((uint32_t)v.bits << 16); }
__device__ __forceinline__ float ld_f32(_Float16 v) { return (float)v; }

template <typename T>
__device__ __forceinline__ T st_from_f32(float x);
template <>
__device__ __forceinline__ float st_from_f32<float>(float x) { return x; }
template <>
__device__ __forceinline__ bf16s st_from_f32<bf16s>(float x)
{
    // round to nearest even; inputs are finite
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return bf16s{(uint16_t)(u >> 16)};
}
template <>
__device__ __forceinline__ _Float16 st_from_f32<_Float16>(float x) { return (_Float16)x; }

// ---------------------------------------------------------------- iteration control

struct PrCtl {
    PrState *st;
    double *part_err;      // [nparts] per-workgroup partials of the main launch
    uint32_t *part_act;    // [nparts]
    uint32_t *part_cmax;   // [nparts] quantised kinds only (else null): largest contribution
    double *hist_err;      // [max_iterations]
    uint32_t *hist_act;    // [max_iterations]
    uint32_t nparts;       // partials of all passes of one iteration
    uint32_t part_base;    // first partial of this launch's workgroups
    uint32_t stop_rule;
    uint32_t mixed;        // 16-bit storage with f32 refinement
    double l1_tol;
    double refine_l1;
};

// Record iteration `iter` and decide whether the run stops (one thread of the finish kernel).
__device__ __forceinline__ void pr_control(const PrCtl &C, uint32_t iter, double E, uint32_t A)
{
    PrState *st = C.st;
    C.hist_err[iter] = E;
    C.hist_act[iter] = A;
    bool can_stop = true;
    if (C.mixed) {
        const int32_t ri = st->refine_iter;
        if (ri < 0) {
            refine_decide(st, iter, E, C.refine_l1);
            can_stop = false;
        } else {
            can_stop = (int32_t)iter >= ri + 1;
        }
    }
    bool stop = false;
    if (C.stop_rule == AG_STOP_ACTIVE)
        stop = (A == 0);
    else if (C.stop_rule == AG_STOP_L1)
        stop = (E <= C.l1_tol);
    if (can_stop && stop) {
        st->done = 1;
        st->stop_iter = iter;
    }
    st->last_err = E;
    st->last_active = A;
    st->iter = iter + 1;
}

// Per-workgroup partials of the main launch (plain stores; read after the launch boundary).
__device__ __forceinline__ void pr_store_partial(const PrCtl &C, Red red)
{
    __shared__ double s_e[AG_WG / AG_WAVE];
    __shared__ uint32_t s_a[AG_WG / AG_WAVE];
    Red t = block_reduce_red<AG_WG>(red, s_e, s_a);
    if (C.part_cmax) {
        const uint32_t m = block_max_u32<AG_WG>(red.cmax, s_a);
        if (threadIdx.x == 0)
            C.part_cmax[C.part_base + blockIdx.x] = m;
    }
    if (threadIdx.x == 0) {
        C.part_err[C.part_base + blockIdx.x] = t.err;
        C.part_act[C.part_base + blockIdx.x] = t.active;
    }
}

// Finish kernel tail: hub rows already folded into `red`; add the main launch's partials in
// index order and close the iteration.
__device__ __forceinline__ void pr_close(const PrCtl &C, uint32_t iter, Red red)
{
    __shared__ double s_e[AG_FIN_WG / AG_WAVE];
    __shared__ uint32_t s_a[AG_FIN_WG / AG_WAVE];
    uint32_t mx = red.cmax;
    for (uint32_t i = threadIdx.x; i < C.nparts; i += AG_FIN_WG) {
        red.err += C.part_err[i];
        red.active += C.part_act[i];
        if (C.part_cmax)
            mx = max(mx, C.part_cmax[i]);
    }
    Red t = block_reduce_red<AG_FIN_WG>(red, s_e, s_a);
    if (C.part_cmax) {
        // the quantiser of the next iteration reads this (after the launch boundary)
        mx = block_max_u32<AG_FIN_WG>(mx, s_a);
        if (threadIdx.x == 0)
            C.st->qmax_bits = mx;
    }
    if (threadIdx.x == 0)
        pr_control(C, iter, t.err, t.active);
}

// ---------------------------------------------------------------- FP32 policy

template <typename RT, typename WT>
struct PrF32Pol {
    using Acc = float;
    const RT *cin;
    WT *cout;
    float *rank;
    const uint32_t *outdeg;
    float base, damp;
    double dn, eps;

    __device__ __forceinline__ float edge(uint32_t, uint32_t s) const { return ld_f32(cin[s]); }
    // the same term from a stored contribution already in hand (hot-segment LDS copy)
    __device__ __forceinline__ float edge_val(uint32_t, RT v) const { return ld_f32(v); }
    __device__ __forceinline__ void finish(uint32_t r, float sum, Red &red) const
    {
        const float nw = __fadd_rn(base, __fmul_rn(damp, sum));
        const float old = __builtin_nontemporal_load(&rank[r]);
        __builtin_nontemporal_store(nw, &rank[r]);
        const double e = (double)fabsf(__fsub_rn(nw, old));
        red.err += __ddiv_rn(e, dn);
        red.active += (e >= eps) ? 1u : 0u;
        const uint32_t od = __builtin_nontemporal_load(&outdeg[r]);
        cout[r] = st_from_f32<WT>(od ? __fdiv_rn(nw, __uint2float_rn(od)) : 0.0f);
    }
};

struct PrF32Args {
    void *c32[2];      // f32 contribution buffers (ping-pong)
    void *c16[2];      // 16-bit contribution buffers (mixed storage only)
    float *rank;
    const uint32_t *outdeg;
    float base, damp;
    double dn, eps;
};

// Runs `body(policy)` with the policy of this iteration's storage mode.
template <typename HT, class F>
__device__ __forceinline__ void pr_f32_dispatch(const PrState *st, uint32_t iter,
                                                const PrF32Args &A, bool mixed, F &&body)
{
    const uint32_t rd = iter & 1u, wr = rd ^ 1u;
    if (!mixed) {
        body(PrF32Pol<float, float>{(const float *)A.c32[rd], (float *)A.c32[wr], A.rank,
                                    A.outdeg, A.base, A.damp, A.dn, A.eps});
        return;
    }
    const int32_t ri = *(volatile const int32_t *)&st->refine_iter;
    const bool wf32 = ri >= 0 && (int32_t)iter >= ri;
    const bool rf32 = ri >= 0 && (int32_t)iter >= ri + 1;
    if (rf32)
        body(PrF32Pol<float, float>{(const float *)A.c32[rd], (float *)A.c32[wr], A.rank,
                                    A.outdeg, A.base, A.damp, A.dn, A.eps});
    else if (wf32)
        body(PrF32Pol<HT, float>{(const HT *)A.c16[rd], (float *)A.c32[wr], A.rank, A.outdeg,
                                 A.base, A.damp, A.dn, A.eps});
    else
        body(PrF32Pol<HT, HT>{(const HT *)A.c16[rd], (HT *)A.c16[wr], A.rank, A.outdeg, A.base,
                              A.damp, A.dn, A.eps});
}

template <typename HT>
__device__ __forceinline__ void pr_f32_main(const PullGeom &G, const PrCtl &C,
                                            const PrF32Args &A, bool mixed)
{
    __shared__ float s_vals[AG_TILE_NNZ_MAX];
    __shared__ float s_wave[AG_WG / AG_WAVE];
    const PrState *st = C.st;
    if (*(volatile const uint32_t *)&st->done)
        return;
    const uint32_t iter = *(volatile const uint32_t *)&st->iter;
    Red red{};
    pr_f32_dispatch<HT>(st, iter, A, mixed,
                        [&](const auto &p) { pull_items(G, p, red, s_vals, s_wave); });
    pr_store_partial(C, red);
}

template <typename HT>
__device__ __forceinline__ void pr_f32_fin(const PullGeom &G, const PrCtl &C,
                                           const PrF32Args &A, bool mixed)
{
    const PrState *st = C.st;
    if (*(volatile const uint32_t *)&st->done)
        return;
    const uint32_t iter = *(volatile const uint32_t *)&st->iter;
    Red red{};
    pr_f32_dispatch<HT>(st, iter, A, mixed, [&](const auto &p) { pull_hub_finish(G, p, red); });
    pr_close(C, iter, red);
}

// Source-blocked iteration: pass 1 (partials of the blocked rows, full update of the others),
// then pass 2 (each blocked row's partials in block order, then the update).
// When the plan keeps an LDS table of the last sources (B.top0 < B.n_src), every workgroup
// first copies their contributions, in the storage this iteration gathers, and pass 1 reads
// those sources from it.
template <typename HT>
__device__ __forceinline__ void pr_f32_blk(const PullGeom &U, const BlockGeom &B,
                                           const PrCtl &C, const PrF32Args &A, bool mixed)
{
    __shared__ float s_vals[AG_TILE_NNZ_MAX];
    __shared__ float s_wave[AG_WG / AG_WAVE];
    __shared__ BlkLds L;
    __shared__ float s_top[AG_TOP_IDS];
    const PrState *st = C.st;
    if (*(volatile const uint32_t *)&st->done)
        return;
    const uint32_t iter = *(volatile const uint32_t *)&st->iter;
    Red red{};
    pr_f32_dispatch<HT>(st, iter, A, mixed, [&](const auto &p) {
        using RT = std::remove_cv_t<std::remove_pointer_t<decltype(p.cin)>>;
        static_assert(sizeof(RT) <= sizeof(float), "the table holds 4-byte contributions");
        if (B.top0 < B.n_src) {
            RT *tab = (RT *)s_top;
            const uint32_t cnt = B.n_src - B.top0;   // <= AG_TOP_IDS (the plan's table)
            for (uint32_t i = threadIdx.x; i < cnt; i += AG_WG)
                tab[i] = p.cin[B.top0 + i];
            __syncthreads();
            pull_blocked_items(U, B, p, p, red, s_vals, s_wave, L, TopTab<RT>{tab, B.top0});
        } else {
            pull_blocked_items(U, B, p, p, red, s_vals, s_wave, L);
        }
    });
    pr_store_partial(C, red);
}

template <typename HT>
__device__ __forceinline__ void pr_f32_cmb(const BlockGeom &B, const PrCtl &C,
                                           const PrF32Args &A, bool mixed)
{
    __shared__ float s_vals[AG_TILE_NNZ_MAX];
    __shared__ float s_wave[AG_WG / AG_WAVE];
    __shared__ BlkLds L;
    const PrState *st = C.st;
    if (*(volatile const uint32_t *)&st->done)
        return;
    const uint32_t iter = *(volatile const uint32_t *)&st->iter;
    Red red{};
    pr_f32_dispatch<HT>(st, iter, A, mixed, [&](const auto &p) {
        pull_combine_items(B, p, red, s_vals, s_wave, L);
    });
    pr_store_partial(C, red);
}

// Hot pass of a source-blocked iteration: the hot-segment tiles, gathered from an LDS copy
// of each tile group's segment of this iteration's contributions (partials only).
template <typename HT>
__device__ __forceinline__ void pr_f32_hot(const BlockGeom &B, const PrCtl &C,
                                           const PrF32Args &A, bool mixed)
{
    __shared__ uint4 s_tab[AG_HOT_TAB_BYTES / 16];
    __shared__ float s_vals[AG_HOT_NNZ];
    __shared__ uint32_t s_long[AG_LONG_MAX];
    __shared__ uint32_t s_nlong;
    const PrState *st = C.st;
    if (*(volatile const uint32_t *)&st->done)
        return;
    const uint32_t iter = *(volatile const uint32_t *)&st->iter;
    pr_f32_dispatch<HT>(st, iter, A, mixed, [&](const auto &p) {
        pull_hot_items(B, p, p.cin, s_tab, s_vals, s_long, &s_nlong);
    });
}

__global__ __launch_bounds__(AG_WG) void pr_f32_hot_kernel(BlockGeom B, PrCtl C, PrF32Args A)
{
    pr_f32_hot<float>(B, C, A, false);
}
__global__ __launch_bounds__(AG_WG) void pr_bf16_hot_kernel(BlockGeom B, PrCtl C, PrF32Args A)
{
    pr_f32_hot<bf16s>(B, C, A, true);
}
__global__ __launch_bounds__(AG_WG) void pr_f16_hot_kernel(BlockGeom B, PrCtl C, PrF32Args A)
{
    pr_f32_hot<_Float16>(B, C, A, true);
}

__global__ __launch_bounds__(AG_WG) void pr_f32_blk_kernel(PullGeom U, BlockGeom B, PrCtl C,
                                                           PrF32Args A)
{
    pr_f32_blk<float>(U, B, C, A, false);
}
__global__ __launch_bounds__(AG_WG) void pr_bf16_blk_kernel(PullGeom U, BlockGeom B, PrCtl C,
                                                            PrF32Args A)
{
    pr_f32_blk<bf16s>(U, B, C, A, true);
}
__global__ __launch_bounds__(AG_WG) void pr_f16_blk_kernel(PullGeom U, BlockGeom B, PrCtl C,
                                                           PrF32Args A)
{
    pr_f32_blk<_Float16>(U, B, C, A, true);
}
__global__ __launch_bounds__(AG_WG) void pr_f32_cmb_kernel(BlockGeom B, PrCtl C, PrF32Args A)
{
    pr_f32_cmb<float>(B, C, A, false);
}
__global__ __launch_bounds__(AG_WG) void pr_bf16_cmb_kernel(BlockGeom B, PrCtl C, PrF32Args A)
{
    pr_f32_cmb<bf16s>(B, C, A, true);
}
__global__ __launch_bounds__(AG_WG) void pr_f16_cmb_kernel(BlockGeom B, PrCtl C, PrF32Args A)
{
    pr_f32_cmb<_Float16>(B, C, A, true);
}

__global__ __launch_bounds__(AG_WG) void pr_f32_kernel(PullGeom G, PrCtl C, PrF32Args A)
{
    pr_f32_main<float>(G, C, A, false);
}
__global__ __launch_bounds__(AG_WG) void pr_bf16_kernel(PullGeom G, PrCtl C, PrF32Args A)
{
    pr_f32_main<bf16s>(G, C, A, true);
}
__global__ __launch_bounds__(AG_WG) void pr_f16_kernel(PullGeom G, PrCtl C, PrF32Args A)
{
    pr_f32_main<_Float16>(G, C, A, true);
}
__global__ __launch_bounds__(AG_FIN_WG) void pr_f32_finish(PullGeom G, PrCtl C, PrF32Args A)
{
    pr_f32_fin<float>(G, C, A, false);
}
__global__ __launch_bounds__(AG_FIN_WG) void pr_bf16_finish(PullGeom G, PrCtl C, PrF32Args A)
{
    pr_f32_fin<bf16s>(G, C, A, true);
}
__global__ __launch_bounds__(AG_FIN_WG) void pr_f16_finish(PullGeom G, PrCtl C, PrF32Args A)
{
    pr_f32_fin<_Float16>(G, C, A, true);
}

// ---------------------------------------------------------------- fixed-point policy

template <int BITS>
struct FixT;
template <> struct FixT<64> { using T = uint64_t; };
template <> struct FixT<32> { using T = uint32_t; };
template <> struct FixT<16> { using T = uint16_t; };
template <> struct FixT<8> { using T = uint8_t; };

template <int BITS>
__device__ __forceinline__ typename FixT<BITS>::T fix_encode(float quotient)
{
    if constexpr (BITS == 64)
        return ag_dev_double_to_fixed64((double)quotient);
    else if constexpr (BITS == 32)
        return ag_dev_float_to_fixed32(quotient);
    else if constexpr (BITS == 16)
        return ag_dev_float_to_fixed16(quotient);
    else
        return ag_dev_float_to_fixed8(quotient);
}

template <int BITS>
__device__ __forceinline__ float fix_update(float base, float damp, uint64_t q)
{
    if constexpr (BITS == 64) {
        // float + float * double is evaluated in double and stored to float (pageRank.c:1400)
        return __double2float_rn(__dadd_rn((double)base, __dmul_rn((double)damp,
                                                                   ag_dev_fixed64_to_double(q))));
    } else {
        float f;
        if constexpr (BITS == 32)
            f = ag_dev_fixed32_to_float(q);
        else if constexpr (BITS == 16)
            f = ag_dev_fixed16_to_float(q);
        else
            f = ag_dev_fixed8_to_float(q);
        return __fadd_rn(base, __fmul_rn(damp, f));
    }
}

template <int BITS>
struct PrFixPol {
    using Acc = uint64_t;
    using CT = typename FixT<BITS>::T;
    const CT *cin;
    CT *cout;
    float *rank;
    const uint32_t *outdeg;
    float base, damp;
    double dn, eps;

    __device__ __forceinline__ uint64_t edge(uint32_t, uint32_t s) const { return (uint64_t)cin[s]; }
    __device__ __forceinline__ void finish(uint32_t r, uint64_t q, Red &red) const
    {
        const float nw = fix_update<BITS>(base, damp, q);
        const float old = __builtin_nontemporal_load(&rank[r]);
        __builtin_nontemporal_store(nw, &rank[r]);
        const double e = (double)fabsf(__fsub_rn(nw, old));
        red.err += __ddiv_rn(e, dn);
        red.active += (e >= eps) ? 1u : 0u;
        const uint32_t od = __builtin_nontemporal_load(&outdeg[r]);
        cout[r] = od ? fix_encode<BITS>(__fdiv_rn(nw, __uint2float_rn(od))) : (CT)0;
    }
};

template <int BITS>
__device__ __forceinline__ PrFixPol<BITS> fix_policy(uint32_t iter, const PrF32Args &A)
{
    using CT = typename FixT<BITS>::T;
    const uint32_t rd = iter & 1u, wr = rd ^ 1u;
    return PrFixPol<BITS>{(const CT *)A.c32[rd], (CT *)A.c32[wr], A.rank, A.outdeg, A.base,
                          A.damp, A.dn, A.eps};
}

// ---------------------------------------------------------------- quantised policy
//
// pageRankPullQuant{32,16,8}BitGraphCSR (pageRank.c:2119-2676): per iteration the f32
// contributions get a scale from their maximum (zero point 0), are quantised to B-bit codes,
// the codes are summed exactly in u64 and the update is base + damp * (scale * (float)sum).
// Here the update writes the f32 contributions and a per-workgroup maximum, the finish kernel
// reduces the maximum, and pr_quant_kernel writes the codes the next gather reads; the u64
// sums are exact, so any summation order gives the same ranks. The quantiser (the reference's
// quantization.h is not in its tree) is defined as
//   scale = mx > 0 ? (float)((double)mx / (2^B - 1)) : 1     mx = largest finite c >= 0
//   code  = c > 0 ? min(2^B - 1, floor((double)c * (1 / (double)scale) + 0.5)) : 0

// bits of a contribution if it takes part in the maximum (finite and not negative), else 0
__device__ __forceinline__ uint32_t quant_max_bits(float c)
{
    const uint32_t b = __float_as_uint(c);
    return b < 0x7F800000u ? b : 0u;
}

template <int BITS>
__device__ __forceinline__ float quant_scale(uint32_t mx_bits)
{
    const float mx = __uint_as_float(mx_bits);
    const double maxcode = (double)((1ull << BITS) - 1ull);
    return mx > 0.0f ? __double2float_rn(__ddiv_rn((double)mx, maxcode)) : 1.0f;
}

template <int BITS>
__device__ __forceinline__ typename FixT<BITS>::T quant_code(float c, double inv)
{
    using CT = typename FixT<BITS>::T;
    const double maxcode = (double)((1ull << BITS) - 1ull);
    if (!(c > 0.0f))
        return (CT)0;
    const double t = floor(__dadd_rn(__dmul_rn((double)c, inv), 0.5));
    return t >= maxcode ? (CT)((1ull << BITS) - 1ull) : (CT)(uint32_t)t;
}

template <int BITS>
struct PrQuantPol {
    using Acc = uint64_t;
    using CT = typename FixT<BITS>::T;
    const CT *codes;
    float *cf;            // f32 contributions of the next iteration
    float *rank;
    const uint32_t *outdeg;
    float base, damp, scale;
    double dn, eps;

    __device__ __forceinline__ uint64_t edge(uint32_t, uint32_t s) const { return (uint64_t)codes[s]; }
    __device__ __forceinline__ void finish(uint32_t r, uint64_t q, Red &red) const
    {
        const float nw = __fadd_rn(base, __fmul_rn(damp, __fmul_rn(scale, __ull2float_rn(q))));
        const float old = __builtin_nontemporal_load(&rank[r]);
        __builtin_nontemporal_store(nw, &rank[r]);
        const double e = (double)fabsf(__fsub_rn(nw, old));
        red.err += __ddiv_rn(e, dn);
        red.active += (e >= eps) ? 1u : 0u;
        const uint32_t od = __builtin_nontemporal_load(&outdeg[r]);
        const float c = od ? __fdiv_rn(nw, __uint2float_rn(od)) : 0.0f;
        cf[r] = c;
        red.cmax = max(red.cmax, quant_max_bits(c));
    }
};

// Integer-sum policies: KC = 64/32/16/8 fixed point, 100 + B quantised B-bit codes.
template <int KC>
__device__ __forceinline__ auto int_policy(const PrState *st, uint32_t iter, const PrF32Args &A)
{
    if constexpr (KC > 100) {
        using CT = typename FixT<KC - 100>::T;
        const float scale = *(volatile const float *)&st->qscale;
        return PrQuantPol<KC - 100>{(const CT *)A.c32[1], (float *)A.c32[0], A.rank, A.outdeg,
                                    A.base, A.damp, scale, A.dn, A.eps};
    } else {
        return fix_policy<KC>(iter, A);
    }
}

// Codes of the next iteration from its f32 contributions; records the scale they use.
template <int BITS>
__global__ __launch_bounds__(AG_WG) void pr_quant_kernel(uint32_t n, PrState *st, const float *cf,
                                                         typename FixT<BITS>::T *codes)
{
    if (*(volatile const uint32_t *)&st->done)
        return;
    const float scale = quant_scale<BITS>(*(volatile const uint32_t *)&st->qmax_bits);
    const double inv = __ddiv_rn(1.0, (double)scale);
    for (uint32_t v = blockIdx.x * AG_WG + threadIdx.x; v < n; v += gridDim.x * AG_WG)
        codes[v] = quant_code<BITS>(cf[v], inv);
    if (blockIdx.x == 0 && threadIdx.x == 0)
        st->qscale = scale;
}

// Largest initial contribution (st->qmax_bits starts at 0).
__global__ __launch_bounds__(AG_WG) void pr_quant_init_max(uint32_t n, const float *cf, PrState *st)
{
    __shared__ uint32_t s_w[AG_WG / AG_WAVE];
    uint32_t m = 0u;
    for (uint32_t v = blockIdx.x * AG_WG + threadIdx.x; v < n; v += gridDim.x * AG_WG)
        m = max(m, quant_max_bits(cf[v]));
    m = block_max_u32<AG_WG>(m, s_w);
    if (threadIdx.x == 0)
        atomicMax(&st->qmax_bits, m);
}

// Integer-sum iteration kernels (fixed point and quantised), KC as in int_policy.
template <int KC>
__global__ __launch_bounds__(AG_WG) void pr_int_kernel(PullGeom G, PrCtl C, PrF32Args A)
{
    __shared__ uint64_t s_vals[AG_TILE_NNZ_MAX];
    __shared__ uint64_t s_wave[AG_WG / AG_WAVE];
    const PrState *st = C.st;
    if (*(volatile const uint32_t *)&st->done)
        return;
    const uint32_t iter = *(volatile const uint32_t *)&st->iter;
    Red red{};
    pull_items(G, int_policy<KC>(st, iter, A), red, s_vals, s_wave);
    pr_store_partial(C, red);
}

template <int KC>
__global__ __launch_bounds__(AG_WG) void pr_int_blk_kernel(PullGeom U, BlockGeom B, PrCtl C,
                                                             PrF32Args A)
{
    __shared__ uint64_t s_vals[AG_TILE_NNZ_MAX];
    __shared__ uint64_t s_wave[AG_WG / AG_WAVE];
    __shared__ BlkLds L;
    const PrState *st = C.st;
    if (*(volatile const uint32_t *)&st->done)
        return;
    const uint32_t iter = *(volatile const uint32_t *)&st->iter;
    Red red{};
    const auto p = int_policy<KC>(st, iter, A);
    pull_blocked_items(U, B, p, p, red, s_vals, s_wave, L);
    pr_store_partial(C, red);
}

template <int KC>
__global__ __launch_bounds__(AG_WG) void pr_int_cmb_kernel(BlockGeom B, PrCtl C, PrF32Args A)
{
    __shared__ uint64_t s_vals[AG_TILE_NNZ_MAX];
    __shared__ uint64_t s_wave[AG_WG / AG_WAVE];
    __shared__ BlkLds L;
    const PrState *st = C.st;
    if (*(volatile const uint32_t *)&st->done)
        return;
    const uint32_t iter = *(volatile const uint32_t *)&st->iter;
    Red red{};
    pull_combine_items(B, int_policy<KC>(st, iter, A), red, s_vals, s_wave, L);
    pr_store_partial(C, red);
}

template <int KC>
__global__ __launch_bounds__(AG_FIN_WG) void pr_int_finish(PullGeom G, PrCtl C, PrF32Args A)
{
    const PrState *st = C.st;
    if (*(volatile const uint32_t *)&st->done)
        return;
    const uint32_t iter = *(volatile const uint32_t *)&st->iter;
    Red red{};
    pull_hub_finish(G, int_policy<KC>(st, iter, A), red);
    pr_close(C, iter, red);
}

// ---------------------------------------------------------------- data-driven, full pull
//
// The data-driven pull (pageRank.c:2810-2969) on a source-blocked plan, with the FP32 pull's
// kernels: every row is evaluated every iteration, and a row counts as listed when one of its
// in-neighbours took a new rank in the previous iteration. The contribution of such a
// neighbour carries the fact in its sign bit (contributions are never negative otherwise), so
// the gather reads no extra byte. This equals the worklist form exactly: an unlisted vertex's
// in-neighbours kept their contributions since its previous evaluation, so its sum and |delta|
// repeat -- 0 if it took the new rank then, below epsilon if it did not -- and it neither
// moves nor counts towards the error. Sums are ordered as in the FP32 pull.
struct PrDdPol {
    using Acc = DdAcc;
    const float *cin;     // signed contributions of this iteration
    float *cout;          // of the next one
    float *rank;
    const uint32_t *outdeg;
    float base, damp;
    double dn, eps;
    bool all_listed;      // the first iteration lists every vertex

    __device__ __forceinline__ DdAcc edge(uint32_t, uint32_t s) const
    {
        const float x = cin[s];
        return DdAcc(fabsf(x), signbit(x) ? 1.0f : 0.0f);
    }
    __device__ __forceinline__ void finish(uint32_t r, DdAcc acc, Red &red) const
    {
        const float nw = __fadd_rn(base, __fmul_rn(damp, acc.s));
        float rk = __builtin_nontemporal_load(&rank[r]);
        const double e = (double)fabsf(__fsub_rn(nw, rk));
        if (all_listed || acc.f > 0.0f)
            red.err += __ddiv_rn(e, dn);
        const bool upd = e >= eps;
        if (upd) {
            rk = nw;
            __builtin_nontemporal_store(nw, &rank[r]);
            red.active += 1u;
        }
        const uint32_t od = __builtin_nontemporal_load(&outdeg[r]);
        const float c = od ? __fdiv_rn(rk, __uint2float_rn(od)) : 0.0f;
        cout[r] = upd ? -c : c;
    }
};

__device__ __forceinline__ PrDdPol dd_policy(uint32_t iter, const PrF32Args &A)
{
    const uint32_t rd = iter & 1u, wr = rd ^ 1u;
    return PrDdPol{(const float *)A.c32[rd], (float *)A.c32[wr], A.rank, A.outdeg, A.base,
                   A.damp, A.dn, A.eps, iter == 0};
}

__global__ __launch_bounds__(AG_WG) void pr_ddf_blk_kernel(PullGeom U, BlockGeom B, PrCtl C,
                                                           PrF32Args A)
{
    __shared__ DdAcc s_vals[AG_TILE_NNZ_MAX];
    __shared__ DdAcc s_wave[AG_WG / AG_WAVE];
    __shared__ BlkLds L;
    const PrState *st = C.st;
    if (*(volatile const uint32_t *)&st->done)
        return;
    const uint32_t iter = *(volatile const uint32_t *)&st->iter;
    Red red{};
    const PrDdPol p = dd_policy(iter, A);
    pull_blocked_items(U, B, p, p, red, s_vals, s_wave, L);
    pr_store_partial(C, red);
}

__global__ __launch_bounds__(AG_WG) void pr_ddf_cmb_kernel(BlockGeom B, PrCtl C, PrF32Args A)
{
    __shared__ DdAcc s_vals[AG_TILE_NNZ_MAX];
    __shared__ DdAcc s_wave[AG_WG / AG_WAVE];
    __shared__ BlkLds L;
    const PrState *st = C.st;
    if (*(volatile const uint32_t *)&st->done)
        return;
    const uint32_t iter = *(volatile const uint32_t *)&st->iter;
    Red red{};
    pull_combine_items(B, dd_policy(iter, A), red, s_vals, s_wave, L);
    pr_store_partial(C, red);
}

__global__ __launch_bounds__(AG_FIN_WG) void pr_ddf_finish(PullGeom G, PrCtl C, PrF32Args A)
{
    const PrState *st = C.st;
    if (*(volatile const uint32_t *)&st->done)
        return;
    const uint32_t iter = *(volatile const uint32_t *)&st->iter;
    Red red{};
    pull_hub_finish(G, dd_policy(iter, A), red);
    pr_close(C, iter, red);
}

// ---------------------------------------------------------------- push direction
//
// pageRankPushGraphCSR (pageRank.c:1111-1272) and pageRankPushFixedPointGraphCSR
// (:1992-2117) keep the pull's host loop and form next[u] by scattering every source's
// contribution along its out-edges. Here the scatter is propagation-blocked (ag_push.h):
// phase 1 copies every out-edge's source contribution into its bin slot, phase 2 adds each
// bin's slots per row in LDS and applies the kind's update to every row of the bin, then the
// one-workgroup finish closes the iteration. Fixed-point sums are exact u64 sums, equal to
// the pull's bit for bit; FP32 sums take the LDS atomics' order (a hub row's bin: a fixed
// tree), which may differ from run to run (FP32 bar).

template <int KC>
struct PushVal {
    using T = typename FixT<KC>::T;
    using A = unsigned long long;
};
template <>
struct PushVal<0> {
    using T = float;
    using A = float;
};

template <int KC>
__global__ __launch_bounds__(AG_WG) void pr_push_scatter_kernel(PrCtl C, PrF32Args A, uint32_t m,
                                                                const uint32_t *fsrc,
                                                                const uint32_t *pos, void *vals)
{
    const PrState *st = C.st;
    if (*(volatile const uint32_t *)&st->done)
        return;
    using VT = typename PushVal<KC>::T;
    const VT *c = (const VT *)A.c32[(*(volatile const uint32_t *)&st->iter) & 1u];
    pb_scatter(m, pos, (VT *)vals,
               [&](uint32_t j) { return c[__builtin_nontemporal_load(&fsrc[j])]; });
}

template <int KC>
__global__ __launch_bounds__(AG_WG) void pr_push_bins_kernel(PrCtl C, PrF32Args A, PushBins P,
                                                             const void *vals)
{
    using VT = typename PushVal<KC>::T;
    using AT = typename PushVal<KC>::A;
    __shared__ AT s_acc[AG_PB_ROWS];
    __shared__ AT s_wave[AG_WG / AG_WAVE];
    const PrState *st = C.st;
    if (*(volatile const uint32_t *)&st->done)
        return;
    const uint32_t iter = *(volatile const uint32_t *)&st->iter;
    Red red{};
    if constexpr (KC == 0) {
        const uint32_t rd = iter & 1u, wr = rd ^ 1u;
        const PrF32Pol<float, float> p{(const float *)A.c32[rd], (float *)A.c32[wr], A.rank,
                                       A.outdeg, A.base, A.damp, A.dn, A.eps};
        pb_bins(P, (const VT *)vals, s_acc, s_wave,
                [&](uint32_t r, float sum) { p.finish(r, sum, red); });
    } else {
        const PrFixPol<KC> p = fix_policy<KC>(iter, A);
        pb_bins(P, (const VT *)vals, s_acc, s_wave,
                [&](uint32_t r, AT sum) { p.finish(r, (uint64_t)sum, red); });
    }
    pr_store_partial(C, red);
}

// Closes an iteration whose every row was updated by the preceding launch.
__global__ __launch_bounds__(AG_FIN_WG) void pr_plain_finish(PrCtl C)
{
    const PrState *st = C.st;
    if (*(volatile const uint32_t *)&st->done)
        return;
    const uint32_t iter = *(volatile const uint32_t *)&st->iter;
    pr_close(C, iter, Red{});
}

// ---------------------------------------------------------------- data-driven pull
//
// pageRankDataDrivenPullGraphCSR (pageRank.c:2810-2969): every vertex starts on the worklist;
// per iteration the contributions rank/out_degree are those of the iteration's start
// (:2870-2876), each listed vertex sums its in-edges, adds |delta|/N to the error, and only
// if |delta| >= epsilon takes the new rank, counts as active and lists its out-neighbours for
// the next iteration (:2888-2921); unlisted vertices keep their rank. Per iteration:
//   pr_dd_gather  one workgroup per listed hub row (in-degree > AG_DD_SEQ_MAX), one 16-lane
//                 group per other listed vertex: gather, update, mark out-neighbours
//   pr_dd_commit  contributions of the listed vertices from their (possibly new) ranks
//   pr_dd_compact next worklist from the marks (one atomic per workgroup), marks cleared
//   pr_dd_finish  iteration closed, the consumed worklist's length cleared
// Ranks change only for listed vertices, so the contribution vector stays exact between
// iterations without a full pass; the worklists' order varies, the per-vertex values do not.
// Rows of in-degree <= AG_DD_SEQ_MAX are summed in the reference's sequential order, so on
// an unblocked plan (rows keep their edge order) with no longer row every value matches the
// sequential restatement bit for bit; longer rows take a fixed lane-strided tree order.

#define AG_DD_LANES 16u
// rows up to this in-degree are summed in edge order; longer rows in a lane-strided tree
#define AG_DD_SEQ_MAX 8192u

__global__ __launch_bounds__(AG_WG) void pr_dd_gather_kernel(PrCtl C, PrF32Args A, uint32_t n,
                                                             const uint32_t *off,
                                                             const uint32_t *src,
                                                             const uint32_t *fwd_off,
                                                             const uint32_t *fwd_dst,
                                                             const uint32_t *queue,
                                                             uint8_t *flag)
{
    __shared__ float s_red[AG_WG / AG_WAVE];
    __shared__ uint32_t s_upd;
    const PrState *st = C.st;
    if (*(volatile const uint32_t *)&st->done)
        return;
    const uint32_t iter = *(volatile const uint32_t *)&st->iter;
    const uint32_t cnt = *(volatile const uint32_t *)&st->dd_count[iter & 1u];
    const uint32_t hubs = *(volatile const uint32_t *)&st->dd_hubs[iter & 1u];
    const float *c = (const float *)A.c32[0];
    Red red{};
    // hub rows (in-degree > AG_DD_SEQ_MAX), listed from the far end: one workgroup each, four
    // independent lane accumulators (a fixed order, so the sum is deterministic)
    for (uint32_t h = blockIdx.x; h < hubs; h += gridDim.x) {
        const uint32_t v = queue[n - 1u - h];
        const uint32_t b = off[v], e = off[v + 1];
        float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
        uint32_t j = b + threadIdx.x;
        for (; j + 3u * AG_WG < e; j += 4u * AG_WG) {
            a0 = __fadd_rn(a0, c[src[j]]);
            a1 = __fadd_rn(a1, c[src[j + AG_WG]]);
            a2 = __fadd_rn(a2, c[src[j + 2u * AG_WG]]);
            a3 = __fadd_rn(a3, c[src[j + 3u * AG_WG]]);
        }
        for (; j < e; j += AG_WG)
            a0 = __fadd_rn(a0, c[src[j]]);
        const float sum = block_reduce<float, AG_WG>(__fadd_rn(__fadd_rn(a0, a1),
                                                               __fadd_rn(a2, a3)), s_red);
        if (threadIdx.x == 0) {
            const float nw = __fadd_rn(A.base, __fmul_rn(A.damp, sum));
            const double err = (double)fabsf(__fsub_rn(nw, A.rank[v]));
            red.err += __ddiv_rn(err, A.dn);
            s_upd = err >= A.eps ? 1u : 0u;
            if (s_upd) {
                A.rank[v] = nw;
                red.active += 1u;
            }
        }
        __syncthreads();
        if (s_upd) {
            for (uint32_t k = fwd_off[v] + threadIdx.x; k < fwd_off[v + 1]; k += AG_WG)
                flag[fwd_dst[k]] = 1;
        }
        __syncthreads();
    }
    const uint32_t lane = threadIdx.x & (AG_DD_LANES - 1u);
    const uint32_t groups = gridDim.x * (AG_WG / AG_DD_LANES);
    for (uint32_t i = (blockIdx.x * AG_WG + threadIdx.x) / AG_DD_LANES; i < cnt; i += groups) {
        const uint32_t v = queue[i];
        const uint32_t b = off[v], e = off[v + 1];   // e - b <= AG_DD_SEQ_MAX
        // the 16 lanes load a chunk, then every lane adds it left to right: the reference's
        // sequential order (pageRank.c:2888-2894), the same sum in every lane
        float sum = 0.0f;
        for (uint32_t j0 = b; j0 < e; j0 += AG_DD_LANES) {
            const uint32_t j = j0 + lane;
            const float x = j < e ? c[src[j]] : 0.0f;
            const uint32_t k = e - j0 < AG_DD_LANES ? e - j0 : AG_DD_LANES;
            for (uint32_t l = 0; l < k; l++)
                sum = __fadd_rn(sum, __shfl(x, (int)l, (int)AG_DD_LANES));
        }
        const float nw = __fadd_rn(A.base, __fmul_rn(A.damp, sum));
        const float old = A.rank[v];
        const double err = (double)fabsf(__fsub_rn(nw, old));
        const bool upd = err >= A.eps;
        if (lane == 0) {
            red.err += __ddiv_rn(err, A.dn);
            if (upd) {
                A.rank[v] = nw;
                red.active += 1u;
            }
        }
        if (upd) {
            for (uint32_t j = fwd_off[v] + lane; j < fwd_off[v + 1]; j += AG_DD_LANES)
                flag[fwd_dst[j]] = 1;
        }
    }
    pr_store_partial(C, red);
}

__global__ __launch_bounds__(AG_WG) void pr_dd_commit_kernel(PrCtl C, PrF32Args A, uint32_t n,
                                                             const uint32_t *queue)
{
    const PrState *st = C.st;
    if (*(volatile const uint32_t *)&st->done)
        return;
    const uint32_t iter = *(volatile const uint32_t *)&st->iter;
    const uint32_t cnt = *(volatile const uint32_t *)&st->dd_count[iter & 1u];
    const uint32_t tot = cnt + *(volatile const uint32_t *)&st->dd_hubs[iter & 1u];
    float *c = (float *)A.c32[0];
    for (uint32_t i = blockIdx.x * AG_WG + threadIdx.x; i < tot; i += gridDim.x * AG_WG) {
        const uint32_t v = queue[i < cnt ? i : n - 1u - (i - cnt)];
        const uint32_t od = A.outdeg[v];
        c[v] = od ? __fdiv_rn(A.rank[v], __uint2float_rn(od)) : 0.0f;
    }
}

// Marked vertices -> worklist `nxt`: rows of in-degree <= AG_DD_SEQ_MAX from the front, hub
// rows from the far end. Each thread owns 4 consecutive marks (one u32 word); one atomic per
// list and workgroup.
__global__ __launch_bounds__(AG_WG) void pr_dd_compact_kernel(PrState *st, uint32_t n,
                                                              const uint32_t *off, uint8_t *flag,
                                                              uint32_t *queue_next, uint32_t nxt)
{
    __shared__ uint32_t s_wave[2][AG_WG / AG_WAVE];
    __shared__ uint32_t s_base[2];
    if (*(volatile const uint32_t *)&st->done)
        return;
    const uint32_t words = (n + 3u) / 4u;
    const uint32_t lane = threadIdx.x & (AG_WAVE - 1u), wave = threadIdx.x / AG_WAVE;
    uint32_t *fw = (uint32_t *)flag;
    for (uint32_t w0 = blockIdx.x * AG_WG; w0 < words; w0 += gridDim.x * AG_WG) {
        const uint32_t w = w0 + threadIdx.x;
        const uint32_t bits = w < words ? fw[w] : 0u;
        uint32_t hub = 0, k[2] = {0u, 0u};
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if ((bits >> (8 * q)) & 1u) {
                const uint32_t v = 4u * w + (uint32_t)q;
                const uint32_t h = off[v + 1] - off[v] > AG_DD_SEQ_MAX ? 1u : 0u;
                hub |= h << q;
                k[h]++;
            }
        }
        // exclusive prefixes of k[0], k[1] over the workgroup
        uint32_t inc[2] = {k[0], k[1]};
#pragma unroll
        for (uint32_t d = 1; d < AG_WAVE; d <<= 1) {
            const uint32_t t0 = __shfl_up(inc[0], d, AG_WAVE);
            const uint32_t t1 = __shfl_up(inc[1], d, AG_WAVE);
            if (lane >= d) {
                inc[0] += t0;
                inc[1] += t1;
            }
        }
        if (lane == AG_WAVE - 1u) {
            s_wave[0][wave] = inc[0];
            s_wave[1][wave] = inc[1];
        }
        __syncthreads();
        if (threadIdx.x < 2) {
            const uint32_t c = threadIdx.x;
            uint32_t tot = 0;
            for (uint32_t q = 0; q < AG_WG / AG_WAVE; q++) {
                const uint32_t t = s_wave[c][q];
                s_wave[c][q] = tot;
                tot += t;
            }
            uint32_t *ctr = c ? &st->dd_hubs[nxt] : &st->dd_count[nxt];
            s_base[c] = tot ? atomicAdd(ctr, tot) : 0u;
        }
        __syncthreads();
        uint32_t pos0 = s_base[0] + s_wave[0][wave] + inc[0] - k[0];
        uint32_t pos1 = s_base[1] + s_wave[1][wave] + inc[1] - k[1];
        if (bits) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                if ((bits >> (8 * q)) & 1u) {
                    const uint32_t v = 4u * w + (uint32_t)q;
                    if ((hub >> q) & 1u)
                        queue_next[n - 1u - pos1++] = v;
                    else
                        queue_next[pos0++] = v;
                }
            }
            fw[w] = 0u;
        }
        __syncthreads();
    }
}

// Finish of a data-driven iteration: closes it and clears the consumed worklist's length,
// which the compaction of the next iteration fills again.
__global__ __launch_bounds__(AG_FIN_WG) void pr_dd_finish(PrCtl C)
{
    PrState *st = C.st;
    if (*(volatile const uint32_t *)&st->done)
        return;
    const uint32_t iter = *(volatile const uint32_t *)&st->iter;
    pr_close(C, iter, Red{});
    if (threadIdx.x == 0) {
        st->dd_count[iter & 1u] = 0u;
        st->dd_hubs[iter & 1u] = 0u;
    }
}

// Every vertex marked (the compaction then builds the first worklist); pad bytes cleared.
__global__ __launch_bounds__(AG_WG) void pr_dd_init_kernel(uint32_t n, uint8_t *flag,
                                                           uint32_t bytes)
{
    for (uint32_t v = blockIdx.x * AG_WG + threadIdx.x; v < bytes; v += gridDim.x * AG_WG)
        flag[v] = v < n ? 1 : 0;
}

// ---------------------------------------------------------------- init / finalise

// kind: 0 = f32 contributions, 1 = bf16, 2 = f16, 64/32/16/8 = fixed point, 1xx = quantised
__global__ void pr_init_kernel(uint32_t n, uint32_t kind, uint32_t init, float base,
                               const float *user_orig, const uint32_t *d2o,
                               const uint32_t *outdeg, float *rank, void *c0)
{
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n)
        return;
    float x = (init == AG_INIT_REFERENCE) ? base : (init == AG_INIT_UNIFORM) ? 1.0f
                                                                           : user_orig[d2o[r]];
    rank[r] = x;
    uint32_t od = outdeg[r];
    float q = od ? __fdiv_rn(x, __uint2float_rn(od)) : 0.0f;
    switch (kind) {
    case 0: ((float *)c0)[r] = q; break;
    case 1: ((bf16s *)c0)[r] = st_from_f32<bf16s>(q); break;
    case 2: ((_Float16 *)c0)[r] = st_from_f32<_Float16>(q); break;
    case 64: ((uint64_t *)c0)[r] = od ? ag_dev_double_to_fixed64((double)q) : 0ull; break;
    case 32: ((uint32_t *)c0)[r] = od ? ag_dev_float_to_fixed32(q) : 0u; break;
    case 16: ((uint16_t *)c0)[r] = od ? ag_dev_float_to_fixed16(q) : (uint16_t)0; break;
    case 8: ((uint8_t *)c0)[r] = od ? ag_dev_float_to_fixed8(q) : (uint8_t)0; break;
    default: ((float *)c0)[r] = q; break;   // quantised kinds: f32, codes follow
    }
}

// out[d2o[r]] = rank[r] / N (pageRank.c:1083); per-workgroup FP64 partial sums
__global__ __launch_bounds__(AG_WG) void pr_finalize_kernel(uint32_t n, const float *rank,
                                                             const uint32_t *d2o, float *out_orig,
                                                             float *scaled_orig, double *part)
{
    __shared__ double s_wave[AG_WG / AG_WAVE];
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (r < n) {
        float x = rank[r];
        float y = __fdiv_rn(x, __uint2float_rn(n));
        out_orig[d2o[r]] = y;
        if (scaled_orig)
            scaled_orig[d2o[r]] = x;
        v = (double)y;
    }
    v = block_reduce<double, AG_WG>(v, s_wave);
    if (threadIdx.x == 0)
        part[blockIdx.x] = v;
}

__global__ void sum_partials_kernel(const double *part, uint32_t count, double *out)
{
    __shared__ double s_wave[AG_WG / AG_WAVE];
    double v = 0.0;
    for (uint32_t i = threadIdx.x; i < count; i += AG_WG)
        v += part[i];
    v = block_reduce<double, AG_WG>(v, s_wave);
    if (threadIdx.x == 0)
        *out = v;
}

__global__ void scaled_copy_kernel(uint32_t n, const float *rank, const uint32_t *d2o, float *out)
{
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n)
        out[d2o[r]] = rank[r];
}

} // namespace

// ---------------------------------------------------------------- host side

// damp = 0.85f and base_pr = 1.0f - damp, both in float (pageRank.h:13, pageRank.c:57-58)
static inline float pr_damp(void) { return 0.85f; }
static inline float pr_base(void)
{
    volatile float d = 0.85f;
    return 1.0f - d;
}



extern "C" void ag_pr_options_default(ag_pr_options *o)
{
    memset(o, 0, sizeof(*o));
    o->kind = AG_KIND_F32;
    o->max_iterations = 20;     // the reference CLI default (accel-graph.c:285-329)
    o->epsilon = 1e-4;
    o->stop_rule = AG_STOP_ACTIVE;
    o->l1_tolerance = 1e-6;
    o->init = AG_INIT_REFERENCE;
    o->storage = AG_STORE_F32;
    o->refine_l1 = 1e-3;
    o->chunk = 0;
    o->record_history = 0;
}

// Upper bound of the per-workgroup partials of one iteration (every pass, every kind): two
// passes of at most 32 workgroups per CU.
static uint32_t ag_pr_max_parts(const ag_dgraph *g)
{
    return 2u * g->dev->num_cus * 32u + AG_NXCD;
}

static void session_free(ag_pr_session *s)
{
    if (!s)
        return;
    (void)hipSetDevice(s->g->dev->ordinal);
    (void)hipStreamSynchronize(s->g->dev->stream);
    void *ptrs[] = {s->st, s->part_err, s->part_act, s->part_cmax, s->hist_err, s->hist_act,
                    s->rank, s->c[0], s->c[1], s->c16[0], s->c16[1], s->push_vals, s->dd_queue[0],
                    s->dd_queue[1], s->dd_flag, s->fin_part, s->fin_sum, s->user};
    for (void *p : ptrs)
        if (p)
            (void)hipFree(p);
    if (s->st_host)
        (void)hipHostFree(s->st_host);
    if (s->ev) {
        for (uint32_t i = 0; i < s->ev_count; i++)
            (void)hipEventDestroy(s->ev[i]);
        free(s->ev);
    }
    if (s->step_ev[0])
        (void)hipEventDestroy(s->step_ev[0]);
    if (s->step_ev[1])
        (void)hipEventDestroy(s->step_ev[1]);
    free(s);
}

static int salloc(ag_pr_session *s, void **p, size_t bytes)
{
    AG_HIP(hipMalloc(p, bytes ? bytes : 8));
    (void)s;
    return AG_OK;
}

#define AG_TRY(x)                   \
    do {                            \
        int ag_st_ = (x);           \
        if (ag_st_ != AG_OK) {      \
            session_free(s);        \
            return ag_st_;          \
        }                           \
    } while (0)

static uint32_t ag_quant_grid(const ag_dgraph *g)
{
    const uint32_t need = (g->n + AG_WG - 1) / AG_WG;
    const uint32_t cap = g->dev->num_cus * 8u;
    return need < 1u ? 1u : need < cap ? need : cap;
}

// Codes of the next iteration (quantised kinds), after the iteration's finish kernel.
static void ag_launch_quant(ag_pr_session *s)
{
    ag_dgraph *g = s->g;
    hipStream_t stream = g->dev->stream;
    const uint32_t grid = ag_quant_grid(g);
    const float *cf = (const float *)s->c[0];
    switch (s->kind_code) {
    case 132:
        pr_quant_kernel<32><<<grid, AG_WG, 0, stream>>>(g->n, s->st, cf, (uint32_t *)s->c[1]);
        break;
    case 116:
        pr_quant_kernel<16><<<grid, AG_WG, 0, stream>>>(g->n, s->st, cf, (uint16_t *)s->c[1]);
        break;
    default:
        pr_quant_kernel<8><<<grid, AG_WG, 0, stream>>>(g->n, s->st, cf, (uint8_t *)s->c[1]);
        break;
    }
}

extern "C" int ag_pr_begin(ag_dgraph *g, const ag_pr_options *opts, ag_pr_session **out)
{
    if (!g || !out)
        return ag_set_error(AG_ERR_INVALID, "graph or out is null");
    *out = NULL;
    ag_pr_options o;
    if (opts)
        o = *opts;
    else
        ag_pr_options_default(&o);
    uint32_t kc;
    size_t celem;
    switch (o.kind) {
    case AG_KIND_F32:
        if (o.storage == AG_STORE_F32) { kc = 0; celem = 4; }
        else if (o.storage == AG_STORE_BF16) { kc = 1; celem = 4; }
        else if (o.storage == AG_STORE_F16) { kc = 2; celem = 4; }
        else return ag_set_error(AG_ERR_INVALID, "unknown storage %u", o.storage);
        break;
    case AG_KIND_FIXED64: kc = 64; celem = 8; break;
    case AG_KIND_FIXED32: kc = 32; celem = 4; break;
    case AG_KIND_FIXED16: kc = 16; celem = 2; break;
    case AG_KIND_FIXED8: kc = 8; celem = 1; break;
    // quantised: c[0] holds the f32 contributions, c[1] the codes
    case AG_KIND_QUANT32: kc = 132; celem = 4; break;
    case AG_KIND_QUANT16: kc = 116; celem = 4; break;
    case AG_KIND_QUANT8: kc = 108; celem = 4; break;
    default: return ag_set_error(AG_ERR_INVALID, "unknown kind %u", o.kind);
    }
    if (o.kind != AG_KIND_F32 && o.storage != AG_STORE_F32)
        return ag_set_error(AG_ERR_UNSUPPORTED, "16-bit storage applies to the FP32 kind only");
    if (o.direction > AG_DIR_DATA_DRIVEN)
        return ag_set_error(AG_ERR_INVALID, "unknown direction %u", o.direction);
    if (o.direction == AG_DIR_PUSH && !(kc == 0 || kc == 64 || kc == 32 || kc == 16 || kc == 8))
        return ag_set_error(AG_ERR_UNSUPPORTED,
                            "the push direction takes FP32 (f32 storage) or a fixed-point kind");
    if (o.direction == AG_DIR_DATA_DRIVEN && kc != 0)
        return ag_set_error(AG_ERR_UNSUPPORTED,
                            "the data-driven pull takes FP32 with f32 storage only");
    // the data-driven pull keeps explicit worklists on unblocked plans (edge-order sums) and
    // runs as a full pull with sign-bit lists on source-blocked plans. The sign bit is free
    // only while every contribution is >= +0, i.e. from a start vector without negative
    // entries, -0 or NaN (ranks stay positive from there); other user starts keep worklists.
    bool dd_lists = o.direction == AG_DIR_DATA_DRIVEN && g->blk.bits == 0;
    if (o.direction == AG_DIR_DATA_DRIVEN && !dd_lists && o.init == AG_INIT_USER &&
        o.init_ranks) {
        for (uint32_t v = 0; v < g->n && !dd_lists; v++) {
            uint32_t bits;
            memcpy(&bits, &o.init_ranks[v], sizeof(bits));
            dd_lists = (bits >> 31) != 0u || !(o.init_ranks[v] >= 0.0f);
        }
    }
    if (o.stop_rule > AG_STOP_NONE)
        return ag_set_error(AG_ERR_INVALID, "unknown stop rule %u", o.stop_rule);
    if (o.init > AG_INIT_USER || (o.init == AG_INIT_USER && !o.init_ranks && g->n))
        return ag_set_error(AG_ERR_INVALID, "bad init mode or missing init_ranks");
    if (!(o.epsilon >= 0.0))
        return ag_set_error(AG_ERR_INVALID, "epsilon must be >= 0");
    if (o.direction == AG_DIR_PUSH || dd_lists) {
        const int fst = dd_lists ? ag_graph_forward(g) : ag_graph_push_bins(g);
        if (fst != AG_OK)
            return fst;
    }

    AG_HIP(hipSetDevice(g->dev->ordinal));
    ag_pr_session *s = (ag_pr_session *)calloc(1, sizeof(ag_pr_session));
    if (!s)
        return ag_set_error(AG_ERR_OOM, "host allocation failed");
    s->g = g;
    s->o = o;
    s->kind_code = kc;
    s->celem = celem;
    clock_gettime(CLOCK_MONOTONIC, &s->t0);
    hipStream_t stream = g->dev->stream;
    const uint32_t n = g->n;
    const uint32_t maxit = o.max_iterations ? o.max_iterations : 1;
    const uint32_t nblk = ag_pr_max_parts(g);
    const uint32_t fblk = (n + AG_WG - 1) / AG_WG ? (n + AG_WG - 1) / AG_WG : 1;

    AG_TRY(salloc(s, (void **)&s->st, sizeof(PrState)));
    {
        hipError_t e = hipHostMalloc((void **)&s->st_host, sizeof(PrState), hipHostMallocDefault);
        if (e != hipSuccess) {
            session_free(s);
            return ag_hip_fail(e, "hipHostMalloc", __FILE__, __LINE__);
        }
    }
    AG_TRY(salloc(s, (void **)&s->part_err, (size_t)nblk * sizeof(double)));
    AG_TRY(salloc(s, (void **)&s->part_act, (size_t)nblk * sizeof(uint32_t)));
    if (kc > 100)
        AG_TRY(salloc(s, (void **)&s->part_cmax, (size_t)nblk * sizeof(uint32_t)));
    AG_TRY(salloc(s, (void **)&s->hist_err, (size_t)maxit * sizeof(double)));
    AG_TRY(salloc(s, (void **)&s->hist_act, (size_t)maxit * sizeof(uint32_t)));
    AG_TRY(salloc(s, (void **)&s->rank, (size_t)n * sizeof(float)));
    AG_TRY(salloc(s, &s->c[0], (size_t)n * celem));
    AG_TRY(salloc(s, &s->c[1], (size_t)n * celem));
    if (kc == 1 || kc == 2) {
        AG_TRY(salloc(s, &s->c16[0], (size_t)n * 2));
        AG_TRY(salloc(s, &s->c16[1], (size_t)n * 2));
    }
    if (o.direction == AG_DIR_PUSH)   // one contribution per out-edge slot, stored as c[]
        AG_TRY(salloc(s, &s->push_vals, (size_t)g->m * celem));
    if (dd_lists) {
        AG_TRY(salloc(s, (void **)&s->dd_queue[0], (size_t)n * sizeof(uint32_t)));
        AG_TRY(salloc(s, (void **)&s->dd_queue[1], (size_t)n * sizeof(uint32_t)));
        AG_TRY(salloc(s, (void **)&s->dd_flag, (size_t)(n + 3u) / 4u * 4u));
    }
    AG_TRY(salloc(s, (void **)&s->fin_part, (size_t)fblk * sizeof(double)));
    AG_TRY(salloc(s, (void **)&s->fin_sum, sizeof(double)));

    PrState init_st;
    memset(&init_st, 0, sizeof(init_st));
    init_st.refine_iter = -1;
    *s->st_host = init_st;
    {
        hipError_t e = hipMemcpyAsync(s->st, s->st_host, sizeof(PrState), hipMemcpyHostToDevice, stream);
        if (e != hipSuccess) {
            session_free(s);
            return ag_hip_fail(e, "state upload", __FILE__, __LINE__);
        }
    }
    if (o.init == AG_INIT_USER && n) {
        AG_TRY(salloc(s, (void **)&s->user, (size_t)n * sizeof(float)));
        hipError_t e = hipMemcpyAsync(s->user, o.init_ranks, (size_t)n * sizeof(float),
                                      hipMemcpyHostToDevice, stream);
        if (e != hipSuccess) {
            session_free(s);
            return ag_hip_fail(e, "init ranks upload", __FILE__, __LINE__);
        }
    }
    if (n) {
        void *c0 = (kc == 1 || kc == 2) ? s->c16[0] : s->c[0];
        pr_init_kernel<<<(n + AG_WG - 1) / AG_WG, AG_WG, 0, stream>>>(
            n, kc, o.init, pr_base(), s->user, g->d2o, g->outdeg, s->rank, c0);
        if (kc > 100) {
            pr_quant_init_max<<<ag_quant_grid(g), AG_WG, 0, stream>>>(n, (const float *)s->c[0],
                                                                     s->st);
            ag_launch_quant(s);
        }
        if (s->dd_queue[0]) {
            // data-driven pull: every vertex starts listed
            pr_dd_init_kernel<<<ag_quant_grid(g), AG_WG, 0, stream>>>(n, s->dd_flag,
                                                                      (n + 3u) / 4u * 4u);
            pr_dd_compact_kernel<<<ag_quant_grid(g), AG_WG, 0, stream>>>(
                s->st, n, g->off, s->dd_flag, s->dd_queue[0], 0u);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            session_free(s);
            return ag_hip_fail(e, "pr_init_kernel launch", __FILE__, __LINE__);
        }
    }
    if (o.record_history) {
        s->ev_count = maxit + 1;
        s->ev = (hipEvent_t *)calloc(s->ev_count, sizeof(hipEvent_t));
        if (!s->ev) {
            session_free(s);
            return ag_set_error(AG_ERR_OOM, "host allocation failed");
        }
        for (uint32_t i = 0; i < s->ev_count; i++) {
            hipError_t e = hipEventCreate(&s->ev[i]);
            if (e != hipSuccess) {
                s->ev_count = i;
                session_free(s);
                return ag_hip_fail(e, "hipEventCreate", __FILE__, __LINE__);
            }
        }
    }
    {
        hipError_t e = hipEventCreate(&s->step_ev[0]);
        if (e == hipSuccess)
            e = hipEventCreate(&s->step_ev[1]);
        if (e != hipSuccess) {
            session_free(s);
            return ag_hip_fail(e, "hipEventCreate", __FILE__, __LINE__);
        }
    }
    *out = s;
    return AG_OK;
}

#define AG_PR_LAUNCH_REGULAR(MAIN, FIN)                      \
    do {                                                     \
        MAIN<<<grid, block, 0, stream>>>(G, C, A);           \
        FIN<<<1, fin, 0, stream>>>(G, C, A);                 \
    } while (0)

// Default resident workgroups per CU of the two blocked passes (plan wg_per_cu overrides
// both): pass 1 is L2-request bound and saturates at 4; the combine streams partials and
// likes 8 (measured on RMAT-25).
#ifndef AG_PASS1_PER_CU
#define AG_PASS1_PER_CU 4u
#endif
#define AG_COMBINE_PER_CU 8u

#define AG_PR_LAUNCH_BLOCKED(BLK, CMB, FIN)                                  \
    do {                                                                     \
        static const uint32_t res1 = ag_resident_per_cu(BLK);               \
        static const uint32_t res2 = ag_resident_per_cu(CMB);               \
        const uint32_t g1 = ag_blocked_grid(g, res1, AG_PASS1_PER_CU);                    \
        const uint32_t g2 = ag_blocked_grid(g, res2, AG_COMBINE_PER_CU);                    \
        C.nparts = g1 + g2;                                                  \
        C.part_base = 0;                                                     \
        BLK<<<g1, block, 0, stream>>>(G, B1, C, A);                          \
        C.part_base = g1;                                                    \
        CMB<<<g2, block, 0, stream>>>(B, C, A);                              \
        FIN<<<1, fin, 0, stream>>>(G, C, A);                                 \
    } while (0)

// The FP32 kinds (any contribution storage) first gather the hot-segment tiles from LDS
// (when the plan has hot segments); pass 1 then skips them.
#define AG_HOT_PER_CU 2u
#define AG_PR_LAUNCH_BLOCKED_HOT(HOT, BLK, CMB, FIN)                         \
    do {                                                                     \
        if (g->blk.n_hgrp) {                                                 \
            static const uint32_t resh = ag_resident_per_cu(HOT);           \
            const uint32_t gh = ag_blocked_grid(g, resh, AG_HOT_PER_CU);    \
            HOT<<<gh, block, 0, stream>>>(B, C, A);                          \
            ag_block_geom_skip_hot(g, B1);                                   \
        }                                                                    \
        AG_PR_LAUNCH_BLOCKED(BLK, CMB, FIN);                                 \
    } while (0)

int ag_pr_launch_iteration(ag_pr_session *s)
{
    ag_dgraph *g = s->g;
    hipStream_t stream = g->dev->stream;
    const bool blocked = g->blk.bits != 0;
    PullGeom G = blocked ? ag_make_unblocked_geom(g) : ag_make_geom(g);
    PrCtl C;
    C.st = s->st;
    C.part_err = s->part_err;
    C.part_act = s->part_act;
    C.part_cmax = s->kind_code > 100 ? s->part_cmax : nullptr;
    C.hist_err = s->hist_err;
    C.hist_act = s->hist_act;
    C.part_base = 0;
    C.stop_rule = s->o.stop_rule;
    C.mixed = (s->kind_code == 1 || s->kind_code == 2) ? 1u : 0u;
    C.l1_tol = s->o.l1_tolerance;
    C.refine_l1 = s->o.refine_l1;
    PrF32Args A;
    A.c32[0] = s->c[0];
    A.c32[1] = s->c[1];
    A.c16[0] = s->c16[0];
    A.c16[1] = s->c16[1];
    A.rank = s->rank;
    A.outdeg = g->outdeg;
    A.base = pr_base();
    A.damp = pr_damp();
    A.dn = (double)g->n;
    A.eps = s->o.epsilon;
    const uint32_t lds = (s->kind_code >= 8) ? AG_TILE_NNZ_MAX * 8u : AG_TILE_NNZ_MAX * 4u;
    dim3 block(AG_WG), fin(AG_FIN_WG);
    if (s->o.direction == AG_DIR_PUSH) {
        const uint32_t gs = (g->dev->num_cus * 16u + AG_NXCD - 1u) / AG_NXCD * AG_NXCD;
        const uint32_t gu = ag_quant_grid(g);
        const PushBins P{g->pb_nbins, g->pb_row, g->pb_off, g->pb_pos, g->pb_dst};
        C.nparts = gu;
        switch (s->kind_code) {
#define AG_PUSH(KC)                                                                              \
    pr_push_scatter_kernel<KC><<<gs, block, 0, stream>>>(C, A, g->m, g->fwd_src, g->pb_pos,    \
                                                         s->push_vals);                         \
    pr_push_bins_kernel<KC><<<gu, block, 0, stream>>>(C, A, P, s->push_vals);                  \
    break
        case 0: AG_PUSH(0);
        case 64: AG_PUSH(64);
        case 32: AG_PUSH(32);
        case 16: AG_PUSH(16);
        default: AG_PUSH(8);
#undef AG_PUSH
        }
        pr_plain_finish<<<1, fin, 0, stream>>>(C);
        AG_HIP(hipGetLastError());
        return AG_OK;
    }
    if (s->o.direction == AG_DIR_DATA_DRIVEN && blocked && !s->dd_queue[0]) {
        const BlockGeom B = ag_make_block_geom(g);
        const BlockGeom &B1 = B;
        AG_PR_LAUNCH_BLOCKED(pr_ddf_blk_kernel, pr_ddf_cmb_kernel, pr_ddf_finish);
        AG_HIP(hipGetLastError());
        return AG_OK;
    }
    if (s->o.direction == AG_DIR_DATA_DRIVEN) {
        const uint32_t gq = ag_quant_grid(g);
        const uint32_t it = s->enqueued;   // parity of the iteration being enqueued
        C.nparts = gq;
        pr_dd_gather_kernel<<<gq, block, 0, stream>>>(C, A, g->n, g->off, g->src, g->fwd_off,
                                                      g->fwd_dst, s->dd_queue[it & 1u],
                                                      s->dd_flag);
        pr_dd_commit_kernel<<<gq, block, 0, stream>>>(C, A, g->n, s->dd_queue[it & 1u]);
        pr_dd_compact_kernel<<<gq, block, 0, stream>>>(s->st, g->n, g->off, s->dd_flag,
                                                       s->dd_queue[(it & 1u) ^ 1u],
                                                       (it & 1u) ^ 1u);
        pr_dd_finish<<<1, fin, 0, stream>>>(C);
        AG_HIP(hipGetLastError());
        return AG_OK;
    }
    if (blocked) {
        const BlockGeom B = ag_make_block_geom(g);
        BlockGeom B1 = B;   // pass 1: past the hot tiles when the hot pass gathers them
        switch (s->kind_code) {
        case 0:
            AG_PR_LAUNCH_BLOCKED_HOT(pr_f32_hot_kernel, pr_f32_blk_kernel, pr_f32_cmb_kernel,
                                     pr_f32_finish);
            break;
        case 1:
            AG_PR_LAUNCH_BLOCKED_HOT(pr_bf16_hot_kernel, pr_bf16_blk_kernel, pr_bf16_cmb_kernel,
                                     pr_bf16_finish);
            break;
        case 2:
            AG_PR_LAUNCH_BLOCKED_HOT(pr_f16_hot_kernel, pr_f16_blk_kernel, pr_f16_cmb_kernel,
                                     pr_f16_finish);
            break;
        case 64:
            AG_PR_LAUNCH_BLOCKED(pr_int_blk_kernel<64>, pr_int_cmb_kernel<64>, pr_int_finish<64>);
            break;
        case 32:
            AG_PR_LAUNCH_BLOCKED(pr_int_blk_kernel<32>, pr_int_cmb_kernel<32>, pr_int_finish<32>);
            break;
        case 16:
            AG_PR_LAUNCH_BLOCKED(pr_int_blk_kernel<16>, pr_int_cmb_kernel<16>, pr_int_finish<16>);
            break;
        case 8:
            AG_PR_LAUNCH_BLOCKED(pr_int_blk_kernel<8>, pr_int_cmb_kernel<8>, pr_int_finish<8>);
            break;
        case 132:
            AG_PR_LAUNCH_BLOCKED(pr_int_blk_kernel<132>, pr_int_cmb_kernel<132>, pr_int_finish<132>);
            break;
        case 116:
            AG_PR_LAUNCH_BLOCKED(pr_int_blk_kernel<116>, pr_int_cmb_kernel<116>, pr_int_finish<116>);
            break;
        default:
            AG_PR_LAUNCH_BLOCKED(pr_int_blk_kernel<108>, pr_int_cmb_kernel<108>, pr_int_finish<108>);
            break;
        }
        if (s->kind_code > 100)
            ag_launch_quant(s);
        AG_HIP(hipGetLastError());
        return AG_OK;
    }
    const uint32_t grid = ag_pull_grid(g, lds);
    C.nparts = grid;
    switch (s->kind_code) {
    case 0: AG_PR_LAUNCH_REGULAR(pr_f32_kernel, pr_f32_finish); break;
    case 1: AG_PR_LAUNCH_REGULAR(pr_bf16_kernel, pr_bf16_finish); break;
    case 2: AG_PR_LAUNCH_REGULAR(pr_f16_kernel, pr_f16_finish); break;
    case 64: AG_PR_LAUNCH_REGULAR(pr_int_kernel<64>, pr_int_finish<64>); break;
    case 32: AG_PR_LAUNCH_REGULAR(pr_int_kernel<32>, pr_int_finish<32>); break;
    case 16: AG_PR_LAUNCH_REGULAR(pr_int_kernel<16>, pr_int_finish<16>); break;
    case 8: AG_PR_LAUNCH_REGULAR(pr_int_kernel<8>, pr_int_finish<8>); break;
    case 132: AG_PR_LAUNCH_REGULAR(pr_int_kernel<132>, pr_int_finish<132>); break;
    case 116: AG_PR_LAUNCH_REGULAR(pr_int_kernel<116>, pr_int_finish<116>); break;
    default: AG_PR_LAUNCH_REGULAR(pr_int_kernel<108>, pr_int_finish<108>); break;
    }
    if (s->kind_code > 100)
        ag_launch_quant(s);
    AG_HIP(hipGetLastError());
    return AG_OK;
}

static uint32_t session_maxit(const ag_pr_session *s)
{
    return s->o.max_iterations ? s->o.max_iterations : 0;
}

extern "C" int ag_pr_step(ag_pr_session *s, uint32_t iterations)
{
    if (!s)
        return ag_set_error(AG_ERR_INVALID, "session is null");
    AG_HIP(hipSetDevice(s->g->dev->ordinal));
    hipStream_t stream = s->g->dev->stream;
    const uint32_t maxit = session_maxit(s);
    s->step_launches = 0;
    if (s->g->total_blocks == 0)
        return AG_OK;
    AG_HIP(hipEventRecord(s->step_ev[0], stream));
    for (uint32_t i = 0; i < iterations && s->enqueued < maxit; i++) {
        if (s->ev)
            AG_HIP(hipEventRecord(s->ev[s->enqueued], stream));
        int st = ag_pr_launch_iteration(s);
        if (st != AG_OK)
            return st;
        s->enqueued++;
        s->step_launches++;
    }
    if (s->ev)
        AG_HIP(hipEventRecord(s->ev[s->enqueued], stream));
    AG_HIP(hipEventRecord(s->step_ev[1], stream));
    return AG_OK;
}

static void fill_result(const ag_pr_session *s, const PrState *st, ag_pr_result *r)
{
    const uint32_t maxit = session_maxit(s);
    memset(r, 0, sizeof(*r));
    if (s->g->n == 0) {
        // no vertex is ever active: the reference loop breaks in its first iteration
        bool stops = s->o.stop_rule != AG_STOP_NONE && maxit > 0;
        r->iterations = stops ? 0 : maxit;
        r->executed = stops ? 1 : maxit;
        r->refine_iteration = -1;
        return;
    }
    if (st->done) {
        r->iterations = st->stop_iter;
        r->executed = st->stop_iter + 1;
    } else {
        r->iterations = st->iter;
        r->executed = st->iter;
    }
    r->active_vertices = st->last_active;
    r->error_total = st->last_err;
    r->refine_iteration = st->refine_iter;
}

extern "C" int ag_pr_poll(ag_pr_session *s, ag_pr_result *res)
{
    if (!s)
        return ag_set_error(AG_ERR_INVALID, "session is null");
    AG_HIP(hipSetDevice(s->g->dev->ordinal));
    AG_HIP(hipMemcpyAsync(s->st_host, s->st, sizeof(PrState), hipMemcpyDeviceToHost,
                          s->g->dev->stream));
    int st = ag_wait_stream(s->g->dev);
    if (st != AG_OK)
        return st;
    if (res)
        fill_result(s, s->st_host, res);
    return AG_OK;
}

extern "C" int ag_pr_last_step_kernel_ms(ag_pr_session *s, double *avg_ms, uint32_t *launches)
{
    if (!s || !avg_ms)
        return ag_set_error(AG_ERR_INVALID, "null argument");
    *avg_ms = 0.0;
    if (launches)
        *launches = s->step_launches;
    if (s->step_launches == 0)
        return AG_OK;
    int st = ag_wait_stream(s->g->dev);
    if (st != AG_OK)
        return st;
    float ms = 0.0f;
    AG_HIP(hipEventElapsedTime(&ms, s->step_ev[0], s->step_ev[1]));
    *avg_ms = (double)ms / (double)s->step_launches;
    return AG_OK;
}

extern "C" int ag_pr_history(ag_pr_session *s, uint32_t cap, uint32_t *active, double *error,
                             float *iter_ms, uint32_t *rows)
{
    if (!s || !rows)
        return ag_set_error(AG_ERR_INVALID, "null argument");
    AG_HIP(hipSetDevice(s->g->dev->ordinal));
    ag_pr_result r;
    int st = ag_pr_poll(s, &r);
    if (st != AG_OK)
        return st;
    uint32_t k = r.executed < cap ? r.executed : cap;
    if (s->g->n == 0)
        k = 0;
    *rows = k;
    if (k == 0)
        return AG_OK;
    if (active)
        AG_HIP(hipMemcpy(active, s->hist_act, (size_t)k * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (error)
        AG_HIP(hipMemcpy(error, s->hist_err, (size_t)k * sizeof(double), hipMemcpyDeviceToHost));
    if (iter_ms) {
        for (uint32_t i = 0; i < k; i++) {
            float ms = 0.0f;
            if (s->ev && i + 1 < s->ev_count && i + 1 <= s->enqueued)
                AG_HIP(hipEventElapsedTime(&ms, s->ev[i], s->ev[i + 1]));
            iter_ms[i] = ms;
        }
    }
    return AG_OK;
}

extern "C" int ag_pr_read_scaled(ag_pr_session *s, float *scaled_out)
{
    if (!s || (!scaled_out && s->g->n))
        return ag_set_error(AG_ERR_INVALID, "null argument");
    ag_dgraph *g = s->g;
    if (g->n == 0)
        return AG_OK;
    AG_HIP(hipSetDevice(g->dev->ordinal));
    int st = ag_ensure_vec(g, (size_t)g->n * sizeof(uint64_t));
    if (st != AG_OK)
        return st;
    scaled_copy_kernel<<<(g->n + AG_WG - 1) / AG_WG, AG_WG, 0, g->dev->stream>>>(
        g->n, s->rank, g->d2o, (float *)g->vec_a);
    AG_HIP(hipGetLastError());
    AG_HIP(hipMemcpyAsync(scaled_out, g->vec_a, (size_t)g->n * sizeof(float),
                          hipMemcpyDeviceToHost, g->dev->stream));
    return ag_wait_stream(g->dev);
}

extern "C" int ag_pr_end(ag_pr_session *s, float *ranks_out, ag_pr_result *res)
{
    if (!s)
        return ag_set_error(AG_ERR_INVALID, "session is null");
    ag_dgraph *g = s->g;
    int st = AG_OK;
    hipError_t e = hipSetDevice(g->dev->ordinal);
    if (e != hipSuccess)
        st = ag_hip_fail(e, "hipSetDevice", __FILE__, __LINE__);
    ag_pr_result r;
    if (st == AG_OK)
        st = ag_pr_poll(s, &r);
    double sum = 0.0;
    if (st == AG_OK && g->n) {
        st = ag_ensure_vec(g, (size_t)g->n * sizeof(uint64_t));
        const uint32_t fblk = (g->n + AG_WG - 1) / AG_WG;
        if (st == AG_OK) {
            pr_finalize_kernel<<<fblk, AG_WG, 0, g->dev->stream>>>(
                g->n, s->rank, g->d2o, (float *)g->vec_a, NULL, s->fin_part);
            sum_partials_kernel<<<1, AG_WG, 0, g->dev->stream>>>(s->fin_part, fblk, s->fin_sum);
            e = hipGetLastError();
            if (e == hipSuccess && ranks_out)
                e = hipMemcpyAsync(ranks_out, g->vec_a, (size_t)g->n * sizeof(float),
                                   hipMemcpyDeviceToHost, g->dev->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(&s->st_host->last_err, s->fin_sum, sizeof(double),
                                   hipMemcpyDeviceToHost, g->dev->stream);
            if (e != hipSuccess)
                st = ag_hip_fail(e, "finalize", __FILE__, __LINE__);
            if (st == AG_OK)
                st = ag_wait_stream(g->dev);
            if (st == AG_OK)
                sum = s->st_host->last_err;
        }
    }
    if (st == AG_OK && res) {
        *res = r;
        res->rank_sum = sum;
        float ms = 0.0f;
        if (s->ev && s->enqueued > 0) {
            if (hipEventElapsedTime(&ms, s->ev[0], s->ev[s->enqueued]) != hipSuccess)
                ms = 0.0f;
        }
        res->loop_ms = ms;
        struct timespec t1;
        clock_gettime(CLOCK_MONOTONIC, &t1);
        res->total_ms = (double)(t1.tv_sec - s->t0.tv_sec) * 1e3 +
                        (double)(t1.tv_nsec - s->t0.tv_nsec) * 1e-6;
    }
    session_free(s);
    return st;
}

extern "C" int ag_pagerank_pull(ag_dgraph *g, const ag_pr_options *opts, float *ranks_out,
                                ag_pr_result *res)
{
    ag_pr_options o;
    if (opts)
        o = *opts;
    else
        ag_pr_options_default(&o);
    ag_pr_session *s = NULL;
    int st = ag_pr_begin(g, &o, &s);
    if (st != AG_OK)
        return st;
    const uint32_t maxit = o.max_iterations;
    // auto: graphs up to 2^24 edges iterate in tens of microseconds, about what a poll costs, so
    // they queue 32 iterations per poll (RMAT-16 to L1 <= 1e-4: 0.54 ms against 0.63 ms with 8);
    // larger graphs queue 8, which bounds the no-op launches queued past the stop
    const uint32_t chunk = o.chunk ? o.chunk : (g->m <= (1u << 24) ? 32u : 8u);
    ag_pr_result r;
    memset(&r, 0, sizeof(r));
    while (s->enqueued < maxit && g->total_blocks) {
        uint32_t k = maxit - s->enqueued < chunk ? maxit - s->enqueued : chunk;
        st = ag_pr_step(s, k);
        if (st == AG_OK)
            st = ag_pr_poll(s, &r);
        if (st != AG_OK) {
            session_free(s);
            return st;
        }
        if (s->st_host->done)
            break;
    }
    return ag_pr_end(s, ranks_out, res);
}
