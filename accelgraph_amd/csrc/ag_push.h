// ag_push.h -- propagation-blocked push (device code shared by PageRank and SpMV push).
//
// The reference's push directions (pageRank.c:1111-1272, :1992-2117; SPMV.c:752-848,
// 1017-1136) add every source's value into next[dst] along its out-edges, with atomics when
// threaded. Random atomics over all of next[] run at the memory-side atomic rate (about
// 16-19 G/s on this chip). Propagation blocking (Beamer, Asanovic, Patterson, "Reducing
// PageRank communication via propagation blocking", IPDPS 2017) splits the scatter in two
// streaming phases:
//   phase 1  every out-edge writes its value to its slot of a bin buffer. Bins are ranges of
//            destination rows; inside a bin the slots follow the sources, so while the grid
//            walks the sources each bin's slots fill front to back, and L2 can merge the
//            stores that reach a line while it is resident. On RMAT-25 most lines leave L2
//            before they fill: a 4-byte store costs about 27 bytes of HBM writes;
//   phase 2  one workgroup per bin adds the bin's values into an LDS row array (LDS atomics)
//            and hands every row's sum to the caller (a bin of one hub row is a tree sum).
// The bins and slots are built once per graph (ag_graph_push_bins).
#pragma once

#include "ag_pull_kernel.h"

struct PushBins {
    uint32_t nbins;
    const uint32_t *row;   // [nbins + 1] first row of every bin
    const uint32_t *off;   // [nbins + 1] first slot of every bin
    const uint32_t *pos;   // [m] slot of every out-edge (forward CSR order)
    const uint16_t *dst;   // [m] row of every slot minus its bin's first row
};

// Phase 1. Workgroup b serves the b % AG_NXCD-th eighth of the out-edges (the grid is a
// multiple of AG_NXCD), so each XCD's stores into a bin stay adjacent and meet in its L2. A
// workgroup takes chunks of AG_PB_UNROLL * AG_WG consecutive out-edges; each lane loads its
// AG_PB_UNROLL slots and values before it stores, which widens the window of sources in flight
// on an XCD and with it the run of adjacent slots each bin receives while its lines are in L2.
// val(j) is the value of out-edge j.
#ifndef AG_PB_UNROLL
#define AG_PB_UNROLL 8u
#endif
template <typename VT, class F>
__device__ __forceinline__ void pb_scatter(uint32_t m, const uint32_t *pos, VT *buf, F val)
{
    const uint32_t x = blockIdx.x % AG_NXCD;
    const uint64_t W = gridDim.x / AG_NXCD, CH = (uint64_t)AG_PB_UNROLL * AG_WG;
    const uint64_t s0 = (uint64_t)m * x / AG_NXCD, s1 = (uint64_t)m * (x + 1) / AG_NXCD;
    for (uint64_t c0 = s0 + (blockIdx.x / AG_NXCD) * CH; c0 < s1; c0 += W * CH) {
        uint32_t p[AG_PB_UNROLL];
        VT v[AG_PB_UNROLL];
#pragma unroll
        for (uint32_t u = 0; u < AG_PB_UNROLL; u++) {
            const uint64_t j = c0 + u * AG_WG + threadIdx.x;
            if (j < s1) {
                p[u] = __builtin_nontemporal_load(&pos[j]);
                v[u] = val((uint32_t)j);
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < AG_PB_UNROLL; u++)
            if (c0 + u * AG_WG + threadIdx.x < s1)
                buf[p[u]] = v[u];
    }
}

// Phase 2 over the bins of a persistent grid: fin(row, sum) once per row of every bin.
// AT is the accumulator (float, or uint64_t for fixed point: exact sums). s_acc holds
// AG_PB_ROWS accumulators, s_wave AG_WG / AG_WAVE.
template <typename VT, typename AT, class FIN>
__device__ __forceinline__ void pb_bins(const PushBins &P, const VT *buf, AT *s_acc, AT *s_wave,
                                        FIN fin)
{
    const uint32_t tid = threadIdx.x;
    for (uint32_t b = blockIdx.x; b < P.nbins; b += gridDim.x) {
        const uint32_t r0 = P.row[b], nr = P.row[b + 1] - r0;
        const uint32_t e0 = P.off[b], e1 = P.off[b + 1];
        if (nr == 1) {
            // a hub row: four lane accumulators, then a fixed tree over the workgroup
            AT a0 = AT(0), a1 = AT(0), a2 = AT(0), a3 = AT(0);
            uint32_t k = e0 + tid;
            for (; k + 3u * AG_WG < e1; k += 4u * AG_WG) {
                a0 = a0 + (AT)__builtin_nontemporal_load(&buf[k]);
                a1 = a1 + (AT)__builtin_nontemporal_load(&buf[k + AG_WG]);
                a2 = a2 + (AT)__builtin_nontemporal_load(&buf[k + 2u * AG_WG]);
                a3 = a3 + (AT)__builtin_nontemporal_load(&buf[k + 3u * AG_WG]);
            }
            for (; k < e1; k += AG_WG)
                a0 = a0 + (AT)__builtin_nontemporal_load(&buf[k]);
            const AT t = block_reduce<AT, AG_WG>((a0 + a1) + (a2 + a3), s_wave);
            if (tid == 0)
                fin(r0, t);
            continue;   // block_reduce ends on a barrier
        }
        for (uint32_t i = tid; i < nr; i += AG_WG)
            s_acc[i] = AT(0);
        __syncthreads();
        for (uint32_t k = e0 + tid; k < e1; k += AG_WG)
            atomicAdd(&s_acc[__builtin_nontemporal_load(&P.dst[k])],
                      (AT)__builtin_nontemporal_load(&buf[k]));
        __syncthreads();
        for (uint32_t i = tid; i < nr; i += AG_WG)
            fin(r0 + i, s_acc[i]);
        __syncthreads();   // s_acc is cleared for the next bin
    }
}
