// ag_pull_kernel.h -- the degree-binned pull gather shared by PageRank, SpMV and the raw
// gather (device code, included by the .hip translation units).
//
// Computes, for every device row r, acc(r) = sum over in-edges j of pol.edge(j, src[j]) and
// hands it to pol.finish(r, acc, red). The reference's gather is the FPGA sum kernel
// (03_capi_integration/accelerator_rtl/cu_control/cu_PageRank/CSR/PULL/FloatPoint/cu/
// cu_sum_kernel_control.sv:170-260; FixedPoint/cu/cu_sum_kernel_control.sv:159-187) fed by
// per-edge reads of auxiliary1[src] (cu_edge_data_read_command_control.sv:179-185), and on
// the CPU the loop shape of 02_capi_graph/src/algorithms/capi/pageRank.c:2890-2907.
//
// One launch covers every bin; blockIdx selects the bin (heaviest first):
//   HUB     one 256-lane workgroup per fixed segment of a hub row; the last segment of a row
//           to finish (ticket counter) sums the segment partials in segment order
//   BLOCK   one workgroup per row
//   WAVE    one 64-lane wavefront per row, four rows per workgroup
//   GROUP16 16 lanes per row, sixteen rows per workgroup
//   THREAD  one workgroup per tile: the tile's in-edge values are gathered cooperatively
//           (coalesced source-id reads) into LDS, then each lane sums whole rows from LDS in
//           stream order -- the same order and rounding as the reference's sequential sum
// Sums in the other bins use a fixed lane-strided order plus a fixed shuffle tree, so every
// result is deterministic from run to run.
#pragma once

#include "ag_internal.h"

#define AG_WAVE 64

struct Red {
    double err;
    uint32_t active;
};

template <typename T>
__device__ __forceinline__ T ag_shfl_xor(T v, int mask, int width)
{
    return __shfl_xor(v, mask, width);
}

template <typename Acc, int W>
__device__ __forceinline__ Acc group_reduce(Acc v)
{
#pragma unroll
    for (int m = W / 2; m >= 1; m >>= 1)
        v = v + ag_shfl_xor(v, m, W);
    return v;
}

// Sum of one Acc per lane over the workgroup; valid in thread 0.
template <typename Acc>
__device__ __forceinline__ Acc block_reduce(Acc v, Acc *s_wave)
{
    const uint32_t tid = threadIdx.x;
    v = group_reduce<Acc, AG_WAVE>(v);
    if ((tid & (AG_WAVE - 1)) == 0)
        s_wave[tid / AG_WAVE] = v;
    __syncthreads();
    Acc t = v;
    if (tid == 0) {
        t = s_wave[0];
#pragma unroll
        for (uint32_t k = 1; k < AG_WG / AG_WAVE; k++)
            t = t + s_wave[k];
    }
    __syncthreads();
    return t;
}

template <typename Acc>
__device__ __forceinline__ uint64_t acc_bits(Acc v);
template <>
__device__ __forceinline__ uint64_t acc_bits<float>(float v) { return (uint64_t)__float_as_uint(v); }
template <>
__device__ __forceinline__ uint64_t acc_bits<uint64_t>(uint64_t v) { return v; }

template <typename Acc>
__device__ __forceinline__ Acc acc_from_bits(uint64_t b);
template <>
__device__ __forceinline__ float acc_from_bits<float>(uint64_t b) { return __uint_as_float((uint32_t)b); }
template <>
__device__ __forceinline__ uint64_t acc_from_bits<uint64_t>(uint64_t b) { return b; }

template <class P>
__device__ __forceinline__ void pull_hub(const PullGeom &G, const P &pol, uint32_t s, Red &red,
                                         typename P::Acc *s_wave)
{
    using Acc = typename P::Acc;
    const uint32_t tid = threadIdx.x;
    const uint32_t row = G.seg_row[s], lo = G.seg_lo[s], hi = G.seg_hi[s];
    Acc acc = Acc(0);
    for (uint32_t j = lo + tid; j < hi; j += AG_WG)
        acc = acc + pol.edge(j, G.src[j]);
    acc = block_reduce<Acc>(acc, s_wave);
    if (tid == 0) {
        __hip_atomic_store(&G.hub_part[s], acc_bits<Acc>(acc), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t h = row - G.bin_row[AG_BIN_HUB];
        const uint32_t s0 = G.hub_seg0[h], s1 = G.hub_seg0[h + 1];
        uint32_t t = __hip_atomic_fetch_add(&G.hub_ticket[h], 1u, __ATOMIC_ACQ_REL,
                                            __HIP_MEMORY_SCOPE_AGENT);
        if (t == s1 - s0 - 1) {
            Acc total = acc_from_bits<Acc>(
                __hip_atomic_load(&G.hub_part[s0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            for (uint32_t k = s0 + 1; k < s1; k++)
                total = total + acc_from_bits<Acc>(__hip_atomic_load(
                                    &G.hub_part[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            __hip_atomic_store(&G.hub_ticket[h], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            pol.finish(row, total, red);
        }
    }
}

template <class P>
__device__ __forceinline__ void pull_block_row(const PullGeom &G, const P &pol, uint32_t row,
                                               Red &red, typename P::Acc *s_wave)
{
    using Acc = typename P::Acc;
    const uint32_t tid = threadIdx.x;
    const uint32_t lo = G.off[row], hi = G.off[row + 1];
    Acc acc = Acc(0);
    for (uint32_t j = lo + tid; j < hi; j += AG_WG)
        acc = acc + pol.edge(j, G.src[j]);
    acc = block_reduce<Acc>(acc, s_wave);
    if (tid == 0)
        pol.finish(row, acc, red);
}

template <class P, int GS>
__device__ __forceinline__ void pull_group_rows(const PullGeom &G, const P &pol, uint32_t bin,
                                                uint32_t blk, Red &red)
{
    using Acc = typename P::Acc;
    const uint32_t tid = threadIdx.x;
    const uint32_t row = G.bin_row[bin] + blk * (AG_WG / GS) + tid / GS;
    const uint32_t lane = tid % GS;
    const bool valid = row < G.bin_row[bin + 1];
    Acc acc = Acc(0);
    if (valid) {
        const uint32_t lo = G.off[row], hi = G.off[row + 1];
        for (uint32_t j = lo + lane; j < hi; j += GS)
            acc = acc + pol.edge(j, G.src[j]);
    }
    acc = group_reduce<Acc, GS>(acc);
    if (valid && lane == 0)
        pol.finish(row, acc, red);
}

template <class P>
__device__ __forceinline__ void pull_thread_tile(const PullGeom &G, const P &pol, uint32_t tile,
                                                 Red &red, typename P::Acc *s_vals)
{
    using Acc = typename P::Acc;
    const uint32_t tid = threadIdx.x;
    const uint32_t r0 = G.tile_row[tile], r1 = G.tile_row[tile + 1];
    const uint32_t e0 = G.off[r0], e1 = G.off[r1];
    const uint32_t nnz = e1 - e0;  // <= AG_TILE_NNZ_MAX, guaranteed by the plan
    for (uint32_t k = tid; k < nnz; k += AG_WG) {
        const uint32_t j = e0 + k;
        s_vals[k] = pol.edge(j, G.src[j]);
    }
    __syncthreads();
    for (uint32_t r = r0 + tid; r < r1; r += AG_WG) {
        const uint32_t b = G.off[r] - e0, e = G.off[r + 1] - e0;
        Acc acc = Acc(0);
        if (e > b) {
            acc = s_vals[b];  // the first element starts the sum, as the FPGA accumulator does
            for (uint32_t k = b + 1; k < e; k++)
                acc = acc + s_vals[k];
        }
        pol.finish(r, acc, red);
    }
}

// The bin body of one workgroup.
template <class P>
__device__ __forceinline__ void pull_body(const PullGeom &G, const P &pol, Red &red,
                                          typename P::Acc *s_vals, typename P::Acc *s_wave)
{
    const uint32_t b = blockIdx.x;
    if (b < G.slot_blk0[1])
        pull_hub(G, pol, b - G.slot_blk0[0], red, s_wave);
    else if (b < G.slot_blk0[2])
        pull_block_row(G, pol, G.bin_row[AG_BIN_BLOCK] + (b - G.slot_blk0[1]), red, s_wave);
    else if (b < G.slot_blk0[3])
        pull_group_rows<P, AG_WAVE>(G, pol, AG_BIN_WAVE, b - G.slot_blk0[2], red);
    else if (b < G.slot_blk0[4])
        pull_group_rows<P, 16>(G, pol, AG_BIN_GROUP16, b - G.slot_blk0[3], red);
    else
        pull_thread_tile(G, pol, b - G.slot_blk0[4], red, s_vals);
}

// Plain pull launch (SpMV, raw gather): no per-iteration control.
template <class P>
__global__ __launch_bounds__(AG_WG) void pull_kernel(PullGeom G, P pol)
{
    __shared__ typename P::Acc s_vals[AG_TILE_NNZ_MAX];
    __shared__ typename P::Acc s_wave[AG_WG / AG_WAVE];
    Red red{0.0, 0u};
    pull_body(G, pol, red, s_vals, s_wave);
}
