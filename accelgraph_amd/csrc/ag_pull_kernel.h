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
// Work items, heaviest first (the plan's launch slots):
//   HUB     a fixed segment of a hub row: one 256-lane workgroup writes the segment's partial
//           sum; the finish step adds a row's partials in segment order
//   BLOCK   one row per workgroup
//   WAVE    one row per 64-lane wavefront, four rows per item
//   GROUP16 one row per 16 lanes, sixteen rows per item
//   THREAD  a tile: its in-edge values are gathered cooperatively (coalesced source-id reads)
//           into LDS, then each lane sums whole rows from LDS in stream order -- the same order
//           and rounding as the reference's sequential sum
// The main kernel runs a persistent grid (compute units x resident workgroups); workgroup b
// takes items b, b + grid, b + 2*grid, ... so the heavy items spread over the whole chip first.
// No atomics on global memory: hub partials and per-workgroup reductions go to plain arrays
// that a one-workgroup finish kernel reads after the launch boundary, in a fixed order, so
// every result is deterministic from run to run. Lane partials use four independent
// accumulators (memory-level parallelism) combined in a fixed tree.
//
// Source-blocked graphs (the second half of this file) replace the gather by two passes over
// the plan's virtual rows -- (row, source block) pieces of in-edges -- so that the gathers of
// each XCD stay inside one L2-sized block of the contribution vector at a time: pass 1 writes
// one partial per piece, pass 2 adds each row's partials in block order and finishes the row.
#pragma once

#include "ag_internal.h"

#define AG_WAVE 64
#define AG_FIN_WG 1024

struct Red {
    double err;
    uint32_t active;
    uint32_t cmax;      // quantised PageRank: bits of the largest contribution written
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

// Sum of one value per lane over an NT-lane workgroup; valid in thread 0. s_wave holds NT/64.
template <typename T, uint32_t NT>
__device__ __forceinline__ T block_reduce(T v, T *s_wave)
{
    const uint32_t tid = threadIdx.x;
    v = group_reduce<T, AG_WAVE>(v);
    if ((tid & (AG_WAVE - 1)) == 0)
        s_wave[tid / AG_WAVE] = v;
    __syncthreads();
    T t = v;
    if (tid == 0) {
        t = s_wave[0];
#pragma unroll
        for (uint32_t k = 1; k < NT / AG_WAVE; k++)
            t = t + s_wave[k];
    }
    __syncthreads();
    return t;
}

template <uint32_t NT>
__device__ __forceinline__ Red block_reduce_red(Red r, double *s_e, uint32_t *s_a)
{
    Red out;
    out.err = block_reduce<double, NT>(r.err, s_e);
    out.active = block_reduce<uint32_t, NT>(r.active, s_a);
    out.cmax = 0u;
    return out;
}

// Largest value over an NT-lane workgroup; valid in thread 0. s_wave holds NT/64.
template <uint32_t NT>
__device__ __forceinline__ uint32_t block_max_u32(uint32_t v, uint32_t *s_wave)
{
    const uint32_t tid = threadIdx.x;
#pragma unroll
    for (int m = AG_WAVE / 2; m >= 1; m >>= 1)
        v = max(v, (uint32_t)__shfl_xor(v, m, AG_WAVE));
    if ((tid & (AG_WAVE - 1)) == 0)
        s_wave[tid / AG_WAVE] = v;
    __syncthreads();
    uint32_t t = v;
    if (tid == 0) {
        t = s_wave[0];
#pragma unroll
        for (uint32_t k = 1; k < NT / AG_WAVE; k++)
            t = max(t, s_wave[k]);
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

// Streaming loads of the index arrays: read once per iteration, so they are marked
// non-temporal to leave L2/MALL capacity to the contribution vector the gathers reuse.
__device__ __forceinline__ uint32_t ld_stream(const uint32_t *p)
{
    return __builtin_nontemporal_load(p);
}

#define AG_UNROLL 4
// Partials per lane in flight when a combine tile is staged in LDS: 8 of 8 bytes (the
// fixed-point and quantised kinds, 4 workgroups per CU for their LDS; RMAT-25 fixed-32
// iteration 2.65 -> 2.57 ms against one at a time, 4 measured the same), 4 of 4 bytes (FP32,
// 8 workgroups per CU). For 4-byte partials 1 / 2 / 4 / 6 measured RMAT-25 2.322 / 2.297 /
// 2.286 / 2.285 ms and ER-25 4.00 / 3.83 / 3.77 / 3.74 ms, and 8 was slower than 1
// (2.43 / 4.34 ms; 60 VGPRs, the same occupancy)
#define AG_CMB_UNROLL 8
#define AG_CMB_UNROLL_4B 4
// source-id loads (then gathers) in flight per lane in an LDS tile: 16 covers a full tile
// (AG_TILE_NNZ_MAX / AG_WG) in one round. Measured on RMAT-25 / ER-25 pass 1: 2 / 4 / 8 / 16
// gave 2.96 / 2.76 / 2.63 / 2.57 ms and 5.11 / 4.63 / 4.52 / 4.43 ms per iteration
#define AG_TILE_UNROLL 16

// Lane-strided partial sum of in-edges lo, lo+S, ... < hi. Each round issues the source-id
// loads of AG_UNROLL elements, then their gathers, all predicated, so a lane has up to
// AG_UNROLL independent gathers in flight and a short row costs one round trip, not one per
// element. Accumulators combine in a fixed tree.
template <class P, uint32_t S>
__device__ __forceinline__ typename P::Acc strided_sum(const PullGeom &G, const P &pol,
                                                        uint32_t lo, uint32_t hi)
{
    using Acc = typename P::Acc;
    Acc a[AG_UNROLL];
#pragma unroll
    for (int u = 0; u < AG_UNROLL; u++)
        a[u] = Acc(0);
    for (uint32_t j = lo; j < hi; j += AG_UNROLL * S) {
        uint32_t sidx[AG_UNROLL];
#pragma unroll
        for (int u = 0; u < AG_UNROLL; u++) {
            const uint32_t jj = j + u * S;
            sidx[u] = jj < hi ? ld_stream(&G.src[jj]) : 0u;
        }
#pragma unroll
        for (int u = 0; u < AG_UNROLL; u++) {
            const uint32_t jj = j + u * S;
            if (jj < hi)
                a[u] = a[u] + pol.edge(jj, sidx[u]);
        }
    }
    return (a[0] + a[1]) + (a[2] + a[3]);
}

template <class P>
__device__ __forceinline__ void pull_hub_segment(const PullGeom &G, const P &pol, uint32_t s,
                                                 typename P::Acc *s_wave)
{
    using Acc = typename P::Acc;
    const uint32_t lo = G.seg_lo[s], hi = G.seg_hi[s];
    Acc acc = strided_sum<P, AG_WG>(G, pol, lo + threadIdx.x, hi);
    acc = block_reduce<Acc, AG_WG>(acc, s_wave);
    if (threadIdx.x == 0)
        G.hub_part[s] = acc_bits<Acc>(acc);
}

// Reference order: hub segment s (the whole row) added by one lane in edge order, the first
// element starting the sum, over chunks of AG_TILE_NNZ_MAX values the workgroup gathers into
// LDS (coalesced source-id reads); the accumulator carries from chunk to chunk.
template <class P>
__device__ __forceinline__ void pull_hub_segment_seq(const PullGeom &G, const P &pol,
                                                     uint32_t s, typename P::Acc *s_vals)
{
    using Acc = typename P::Acc;
    const uint32_t lo = G.seg_lo[s], hi = G.seg_hi[s];
    Acc acc = Acc(0);
    for (uint32_t base = lo; base < hi; base += AG_TILE_NNZ_MAX) {
        const uint32_t nnz = hi - base < AG_TILE_NNZ_MAX ? hi - base : AG_TILE_NNZ_MAX;
        for (uint32_t k = threadIdx.x; k < nnz; k += AG_WG)
            s_vals[k] = pol.edge(base + k, ld_stream(&G.src[base + k]));
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t q = 0;
            if (base == lo)
                acc = s_vals[q++];
            for (; q < nnz; q++)
                acc = acc + s_vals[q];
        }
        __syncthreads();   // s_vals is refilled by the next chunk
    }
    if (threadIdx.x == 0)
        G.hub_part[s] = acc_bits<Acc>(acc);
}

template <class P>
__device__ __forceinline__ void pull_block_row(const PullGeom &G, const P &pol, uint32_t row,
                                               Red &red, typename P::Acc *s_wave)
{
    using Acc = typename P::Acc;
    const uint32_t lo = G.off[row], hi = G.off[row + 1];
    Acc acc = strided_sum<P, AG_WG>(G, pol, lo + threadIdx.x, hi);
    acc = block_reduce<Acc, AG_WG>(acc, s_wave);
    if (threadIdx.x == 0)
        pol.finish(row, acc, red);
}

// Rows [rbase + k*(AG_WG/GS) ...]: one row per GS lanes; rows at or past rend are idle.
template <class P, uint32_t GS>
__device__ __forceinline__ void group_rows(const PullGeom &G, const P &pol, uint32_t rbase,
                                           uint32_t rend, Red &red)
{
    using Acc = typename P::Acc;
    const uint32_t tid = threadIdx.x;
    const uint32_t row = rbase + tid / GS;
    const uint32_t lane = tid % GS;
    const bool valid = row < rend;
    Acc acc = Acc(0);
    if (valid)
        acc = strided_sum<P, GS>(G, pol, G.off[row] + lane, G.off[row + 1]);
    acc = group_reduce<Acc, GS>(acc);
    if (valid && lane == 0)
        pol.finish(row, acc, red);
}

template <class P, uint32_t GS>
__device__ __forceinline__ void pull_group_rows(const PullGeom &G, const P &pol, uint32_t bin,
                                                uint32_t item, Red &red)
{
    group_rows<P, GS>(G, pol, G.bin_row[bin] + item * (AG_WG / GS), G.bin_row[bin + 1], red);
}

// Rows [r0, r1) with at most AG_TILE_NNZ_MAX in-edges in total: the in-edge values are
// gathered cooperatively (coalesced source-id reads) into LDS, then each lane sums whole rows
// from LDS in stream order -- the same order and rounding as the reference's sequential sum.
template <class P>
__device__ __forceinline__ void tile_rows(const PullGeom &G, const P &pol, uint32_t r0,
                                          uint32_t r1, Red &red, typename P::Acc *s_vals)
{
    using Acc = typename P::Acc;
    const uint32_t tid = threadIdx.x;
    const uint32_t e0 = G.off[r0], e1 = G.off[r1];
    const uint32_t nnz = e1 - e0;  // <= AG_TILE_NNZ_MAX, guaranteed by the plan
    for (uint32_t k = tid; k < nnz; k += AG_TILE_UNROLL * AG_WG) {
        uint32_t sidx[AG_TILE_UNROLL];
#pragma unroll
        for (int u = 0; u < AG_TILE_UNROLL; u++) {
            const uint32_t kk = k + u * AG_WG;
            sidx[u] = kk < nnz ? ld_stream(&G.src[e0 + kk]) : 0u;
        }
#pragma unroll
        for (int u = 0; u < AG_TILE_UNROLL; u++) {
            const uint32_t kk = k + u * AG_WG;
            if (kk < nnz)
                s_vals[kk] = pol.edge(e0 + kk, sidx[u]);
        }
    }
    __syncthreads();
    for (uint32_t r = r0 + tid; r < r1; r += AG_WG) {
        const uint32_t b = G.off[r] - e0, e = G.off[r + 1] - e0;
        Acc acc = Acc(0);
        if (e > b) {
            acc = s_vals[b];  // the first element starts the sum, as the FPGA accumulator does
            for (uint32_t q = b + 1; q < e; q++)
                acc = acc + s_vals[q];
        }
        pol.finish(r, acc, red);
    }
    __syncthreads();  // s_vals is reused by this workgroup's next tile
}

template <class P>
__device__ __forceinline__ void pull_thread_tile(const PullGeom &G, const P &pol, uint32_t tile,
                                                 Red &red, typename P::Acc *s_vals)
{
    tile_rows(G, pol, G.tile_row[tile], G.tile_row[tile + 1], red, s_vals);
}

// Work item `it` of a regular schedule (launch slots of G.slot_blk0).
template <class P>
__device__ __forceinline__ void pull_one_item(const PullGeom &G, const P &pol, uint32_t it,
                                              Red &red, typename P::Acc *s_vals,
                                              typename P::Acc *s_wave)
{
    if (it < G.slot_blk0[1]) {
        if (G.seq_hubs)
            pull_hub_segment_seq(G, pol, it - G.slot_blk0[0], s_vals);
        else
            pull_hub_segment(G, pol, it - G.slot_blk0[0], s_wave);
    } else if (it < G.slot_blk0[2])
        pull_block_row(G, pol, G.bin_row[AG_BIN_BLOCK] + (it - G.slot_blk0[1]), red, s_wave);
    else if (it < G.slot_blk0[3])
        pull_group_rows<P, AG_WAVE>(G, pol, AG_BIN_WAVE, it - G.slot_blk0[2], red);
    else if (it < G.slot_blk0[4])
        pull_group_rows<P, 16>(G, pol, AG_BIN_GROUP16, it - G.slot_blk0[3], red);
    else
        pull_thread_tile(G, pol, it - G.slot_blk0[4], red, s_vals);
}

// All work items of one workgroup of the persistent grid.
template <class P>
__device__ __forceinline__ void pull_items(const PullGeom &G, const P &pol, Red &red,
                                           typename P::Acc *s_vals, typename P::Acc *s_wave)
{
    const uint32_t total = G.slot_blk0[AG_NSLOTS];
    for (uint32_t it = blockIdx.x; it < total; it += gridDim.x)
        pull_one_item(G, pol, it, red, s_vals, s_wave);
}

// ---------------------------------------------------------------- source-blocked pull

#define AG_LANE_MAX 64u    // rows of a tile with at most this many elements: a lane each
#define AG_LONG_MAX 64u    // longer rows of one tile: at most AG_TILE_NNZ_MAX / (AG_LANE_MAX + 1)
static_assert(AG_TILE_NNZ_MAX / (AG_LANE_MAX + 1u) < AG_LONG_MAX,
              "the long rows of a full tile must fit the tile's LDS list");

// First pass: a virtual row's sum is stored as its partial; the update waits for the combine.
template <class P>
struct PartPol {
    using Acc = typename P::Acc;
    P base;
    Acc *part;
    __device__ __forceinline__ Acc edge(uint32_t j, uint32_t s) const { return base.edge(j, s); }
    template <typename V>
    __device__ __forceinline__ Acc edge_val(uint32_t j, V v) const { return base.edge_val(j, v); }
    __device__ __forceinline__ void finish(uint32_t v, Acc acc, Red &) const { part[v] = acc; }
};

// Second pass: a row's "in-edges" are its partial slots, added in source-block order.
template <class P>
struct CombPol {
    using Acc = typename P::Acc;
    P base;
    const Acc *part;
    __device__ __forceinline__ Acc edge(uint32_t, uint32_t v) const { return part[v]; }
    __device__ __forceinline__ void finish(uint32_t r, Acc acc, Red &red) const
    {
        base.finish(r, acc, red);
    }
};

// Where a first-pass gather reads edge j's term from (source s): the policy's own gather
// (NoTop), or, for the sources [base, base + n) of an LDS copy of that part of the
// contribution vector, the copy (TopTab). Both give the same value.
struct NoTop {
    template <class P>
    __device__ __forceinline__ typename P::Acc term(const P &pol, uint32_t j, uint32_t s) const
    {
        return pol.edge(j, s);
    }
};
template <typename RT>
struct TopTab {
    const RT *tab;
    uint32_t base;
    template <class P>
    __device__ __forceinline__ typename P::Acc term(const P &pol, uint32_t j, uint32_t s) const
    {
        return s >= base ? pol.edge_val(j, tab[s - base]) : pol.edge(j, s);
    }
};

// Gathers in-edges [e0, e0 + nnz) (nnz <= AG_TILE_NNZ_MAX) into s_vals with coalesced
// source-id reads, AG_TILE_UNROLL loads then gathers in flight per lane.
template <class P, class TT = NoTop>
__device__ __forceinline__ void stage_edges(const uint32_t *src, uint32_t e0, uint32_t nnz,
                                            const P &pol, typename P::Acc *s_vals,
                                            const TT &tt = TT())
{
    const uint32_t tid = threadIdx.x;
    for (uint32_t k = tid; k < nnz; k += AG_TILE_UNROLL * AG_WG) {
        uint32_t sidx[AG_TILE_UNROLL];
#pragma unroll
        for (int u = 0; u < AG_TILE_UNROLL; u++) {
            const uint32_t kk = k + u * AG_WG;
            sidx[u] = kk < nnz ? ld_stream(&src[e0 + kk]) : 0u;
        }
#pragma unroll
        for (int u = 0; u < AG_TILE_UNROLL; u++) {
            const uint32_t kk = k + u * AG_WG;
            if (kk < nnz)
                s_vals[kk] = tt.term(pol, e0 + kk, sidx[u]);
        }
    }
}

// Rows [r0, r1) whose elements are s_vals[bnd[r] - base, bnd[r + 1] - base): a row of at
// most AG_LANE_MAX elements is summed by one lane in element order (the first element starts
// the sum); a longer row by a wavefront (lane-strided sums combined in a fixed tree). A row's
// value depends only on its elements, whichever lane or wavefront computes it.
// kReady: the caller has zeroed *s_nlong and passed a barrier since the values were staged.
template <class P, bool kReady = false>
__device__ __forceinline__ void mixed_reduce(const uint32_t *bnd, uint32_t base, uint32_t r0,
                                             uint32_t r1, const P &pol, Red &red,
                                             const typename P::Acc *s_vals, uint32_t *s_long,
                                             uint32_t *s_nlong)
{
    using Acc = typename P::Acc;
    const uint32_t tid = threadIdx.x;
    if (!kReady) {
        if (tid == 0)
            *s_nlong = 0u;
        __syncthreads();
    }
    for (uint32_t r = r0 + tid; r < r1; r += AG_WG) {
        const uint32_t b = bnd[r] - base, e = bnd[r + 1] - base;
        if (e - b > AG_LANE_MAX) {
            s_long[atomicAdd(s_nlong, 1u)] = r;
            continue;
        }
        Acc acc = Acc(0);
        if (e > b) {
            acc = s_vals[b];
            for (uint32_t q = b + 1; q < e; q++)
                acc = acc + s_vals[q];
        }
        pol.finish(r, acc, red);
    }
    __syncthreads();
    const uint32_t nlong = *s_nlong;
    const uint32_t lane = tid % AG_WAVE;
    for (uint32_t i = tid / AG_WAVE; i < nlong; i += AG_WG / AG_WAVE) {
        const uint32_t r = s_long[i];
        const uint32_t b = bnd[r] - base, e = bnd[r + 1] - base;
        Acc acc = Acc(0);
        for (uint32_t q = b + lane; q < e; q += AG_WAVE)
            acc = acc + s_vals[q];
        acc = group_reduce<Acc, AG_WAVE>(acc);
        if (lane == 0)
            pol.finish(r, acc, red);
    }
    __syncthreads();  // s_vals and s_long are reused by this workgroup's next tile
}

// First pass of a source-blocked iteration. Workgroup b serves slice b % AG_NXCD (the XCD it
// is dispatched to) and walks that slice's tiles, which go source block by source block, so
// the gathers of the workgroups of one XCD stay inside about one block of the contribution
// vector, which that XCD's L2 holds. `pol` gathers through virtual-row edge positions (its
// per-edge arrays, if any, are in virtual-row order); AG_ITEM_REG items run the unblocked
// rows' regular schedule U with `upol`.
// With B.dyn (never with an AG_ITEM_REG item in pass 1) the tiles are handed out in order
// through a per-slice ticket counter, so the tiles in flight on an XCD stay consecutive -- a fixed
// round-robin lets the workgroups drift apart by several blocks over an iteration, and the
// XCD's gathers then spread over more of the vector than its L2 holds. A tile's result does
// not depend on the workgroup that runs it. Each workgroup takes its next ticket while it
// runs the current tile; the last workgroup of the slice to leave resets the counters for
// the next launch.
// LDS of the blocked passes beyond the tile values, declared once per kernel (a kernel that
// instantiates a pass for several policies shares it).
struct BlkLds {
    uint32_t pre[AG_MAX_SRC_BLOCKS + 1];   // combine: the tile's partials before each block
    uint32_t start[AG_MAX_SRC_BLOCKS];     // combine: the tile's first virtual row per block
    uint32_t lng[AG_LONG_MAX];             // rows of a tile summed by a wavefront
    uint32_t nlong;
    uint32_t ticket[2];                    // pass 1 with tickets: the workgroup's next tile
};

// Partials of virtual rows [v0, v0 + nnz) that are single in-edges [e0, e0 + nnz): each is
// its edge's gathered term, written straight out (no LDS, no barrier).
template <class PP, class TT>
__device__ __forceinline__ void gather_singletons(const uint32_t *src, uint32_t e0, uint32_t v0,
                                                  uint32_t nnz, const PP &pp, Red &red,
                                                  const TT &tt)
{
    for (uint32_t k = threadIdx.x; k < nnz; k += AG_TILE_UNROLL * AG_WG) {
        uint32_t sidx[AG_TILE_UNROLL];
#pragma unroll
        for (int u = 0; u < AG_TILE_UNROLL; u++) {
            const uint32_t kk = k + u * AG_WG;
            sidx[u] = kk < nnz ? ld_stream(&src[e0 + kk]) : 0u;
        }
#pragma unroll
        for (int u = 0; u < AG_TILE_UNROLL; u++) {
            const uint32_t kk = k + u * AG_WG;
            if (kk < nnz)
                pp.finish(v0 + kk, tt.term(pp, e0 + kk, sidx[u]), red);
        }
    }
}

// One pass-1 tile, virtual rows [first, last). When they are all single edges (as many
// rows as edges) the tile is a plain gather; otherwise its edges are staged in LDS and
// reduced row by row, ending with a barrier. Either way each partial is its piece's terms
// added in element order, so the two forms agree bit for bit.
template <class PP, class TT>
__device__ __forceinline__ void blocked_tile(const BlockGeom &B, const PP &pp, Red &red,
                                             typename PP::Acc *s_vals, BlkLds &L,
                                             uint32_t first, uint32_t last, const TT &tt)
{
    const uint32_t e0 = B.V.off[first], nnz = B.V.off[last] - e0;
    if (nnz == last - first) {
        gather_singletons(B.V.src, e0, first, nnz, pp, red, tt);
        return;
    }
    stage_edges(B.V.src, e0, nnz, pp, s_vals, tt);
    __syncthreads();
    mixed_reduce(B.V.off, e0, first, last, pp, red, s_vals, L.lng, &L.nlong);
}

// A pass-1 item of a plan with sorted tiles (B.sbits): item `it` of the slice (none when
// it >= end: valid 0). Uniform across the workgroup.
struct SortedTile {
    uint32_t valid, reg, first, last, e0, nnz, base;
};

__device__ __forceinline__ SortedTile sorted_tile_at(const BlockGeom &B, uint32_t it,
                                                     uint32_t end)
{
    SortedTile t{};
    if (it >= end)
        return t;
    const uint32_t kc = B.item[2 * it + 1];
    t.valid = 1u;
    t.first = B.item[2 * it];
    if ((kc & 15u) == AG_ITEM_REG) {
        t.reg = 1u;   // nnz 0: no ids to load
        return t;
    }
    t.last = t.first + (kc >> 4);
    t.e0 = B.V.off[t.first];
    t.nnz = B.V.off[t.last] - t.e0;
    t.base = B.tbase[it];
    return t;
}

static_assert(AG_TILE_NNZ_MAX <= AG_TILE_UNROLL * AG_WG,
              "a lane holds all of its ids of a sorted tile at once");

// The packed ids of a sorted tile, AG_TILE_UNROLL per lane: the whole tile.
__device__ __forceinline__ void load_sorted_ids(const uint32_t *vsp, const SortedTile &t,
                                                uint32_t (&w)[AG_TILE_UNROLL])
{
#pragma unroll
    for (int u = 0; u < AG_TILE_UNROLL; u++) {
        const uint32_t kk = threadIdx.x + u * AG_WG;
        w[u] = kk < t.nnz ? ld_stream(&vsp[t.e0 + kk]) : 0u;
    }
}

// Pass 1 over sorted tiles, software-pipelined: word w of a tile holds the source's offset
// from the tile's block start in its low sbits bits and the edge's position in the tile
// above them, so the lanes of a wavefront gather neighbouring sources (they share more L2
// lines than the in-edges of consecutive rows do) and each term lands at its position in
// s_vals, which then holds what an unsorted tile stages; per-edge data (weights) is stored
// in the same sorted order and read at the stream index. After a tile's gathers and barrier,
// the next tile's item, offsets and ids are loaded while this tile's rows are summed, so a
// workgroup waits on one dependent round of loads per tile (the gathers) instead of four.
// Tiles go round-robin or, with B.dyn, by ticket, as in pull_blocked_items; a tile's
// partials are the same either way and the same as with unsorted ids.
template <class P, class PU, class TT>
__device__ __forceinline__ void pull_sorted_items(const PullGeom &U, const BlockGeom &B,
                                                  const P &pol, const PU &upol, Red &red,
                                                  typename P::Acc *s_vals,
                                                  typename P::Acc *s_wave, BlkLds &L,
                                                  const TT &tt)
{
    using Acc = typename P::Acc;
    const PartPol<P> pp{pol, (Acc *)B.part};
    const uint32_t x = blockIdx.x % AG_NXCD, step = gridDim.x / AG_NXCD, tid = threadIdx.x;
    const uint32_t lo = B.item_lo[x], end = B.item_off[x + 1];
    const uint32_t bits = B.sbits, mask = (1u << bits) - 1u;
    uint32_t *tick = B.tick + x, *left = B.tick + AG_NXCD + x;
    uint32_t i;
    if (B.dyn) {
        if (tid == 0)
            L.ticket[0] = atomicAdd(tick, 1u);
        __syncthreads();
        i = lo + L.ticket[0];
    } else {
        i = lo + blockIdx.x / AG_NXCD;
    }
    SortedTile cur = sorted_tile_at(B, i, end);
    uint32_t w[AG_TILE_UNROLL];
    load_sorted_ids(B.vsp, cur, w);
    uint32_t k = 0;
    while (cur.valid) {
        uint32_t next = 0;
        if (B.dyn && tid == 0)
            next = atomicAdd(tick, 1u);   // in flight during the gathers
        if (cur.reg) {
            pull_one_item(U, upol, cur.first, red, s_vals, s_wave);
        } else {
#pragma unroll
            for (int u = 0; u < AG_TILE_UNROLL; u++) {
                const uint32_t kk = tid + u * AG_WG;
                if (kk < cur.nnz) {
                    const uint32_t pos = w[u] >> bits;   // < nnz: a permutation of the tile
                    s_vals[pos] = tt.term(pp, cur.e0 + kk, cur.base + (w[u] & mask));
                }
            }
        }
        // the ticket word written here was last read before the previous tile's final
        // barrier, and so was nlong
        k ^= 1u;
        if (tid == 0) {
            L.ticket[k] = next;
            L.nlong = 0u;
        }
        __syncthreads();
        i = B.dyn ? lo + L.ticket[k] : i + step;
        const SortedTile nxt = sorted_tile_at(B, i, end);
        load_sorted_ids(B.vsp, nxt, w);   // in flight while this tile's rows are summed
        if (!cur.reg) {
            if (cur.nnz == cur.last - cur.first) {   // single edges: partials in order
                for (uint32_t q = tid; q < cur.nnz; q += AG_WG)
                    pp.finish(cur.first + q, s_vals[q], red);
                __syncthreads();   // s_vals is reused by the next tile
            } else {
                mixed_reduce<PartPol<P>, true>(B.V.off, cur.e0, cur.first, cur.last, pp, red,
                                               s_vals, L.lng, &L.nlong);
            }
        }
        cur = nxt;
    }
    if (B.dyn && tid == 0) {
        __threadfence();
        // every workgroup of the slice has taken its last ticket once `left` reaches step
        if (atomicAdd(left, 1u) == step - 1u) {
            atomicExch(tick, 0u);
            atomicExch(left, 0u);
        }
    }
}

template <class P, class PU, class TT = NoTop>
__device__ __forceinline__ void pull_blocked_items(const PullGeom &U, const BlockGeom &B,
                                                   const P &pol, const PU &upol, Red &red,
                                                   typename P::Acc *s_vals,
                                                   typename P::Acc *s_wave, BlkLds &L,
                                                   const TT &tt = TT())
{
    using Acc = typename P::Acc;
    if (B.sbits) {
        pull_sorted_items(U, B, pol, upol, red, s_vals, s_wave, L, tt);
        return;
    }
    const PartPol<P> pp{pol, (Acc *)B.part};
    const uint32_t x = blockIdx.x % AG_NXCD;
    const uint32_t step = gridDim.x / AG_NXCD;   // the grid is a multiple of AG_NXCD
    const uint32_t end = B.item_off[x + 1];
    if (B.dyn) {
        const uint32_t lo = B.item_lo[x], count = end - lo, tid = threadIdx.x;
        uint32_t *tick = B.tick + x, *left = B.tick + AG_NXCD + x;
        // the ticket of round k is read from L.ticket[k & 1] after the round's barrier, and
        // the next one written to the other word, which no lane reads before the next barrier
        uint32_t k = 0;
        if (tid == 0)
            L.ticket[0] = atomicAdd(tick, 1u);
        __syncthreads();
        uint32_t t = L.ticket[0];
        while (t < count) {
            uint32_t next = 0;
            if (tid == 0)
                next = atomicAdd(tick, 1u);   // in flight while this tile runs
            const uint32_t i = lo + t;
            const uint32_t first = B.item[2 * i], last = first + (B.item[2 * i + 1] >> 4);
            blocked_tile(B, pp, red, s_vals, L, first, last, tt);
            k ^= 1u;
            if (tid == 0)
                L.ticket[k] = next;
            __syncthreads();
            t = L.ticket[k];
        }
        if (tid == 0) {
            __threadfence();
            // every workgroup of the slice has taken its last ticket once `left` reaches step
            if (atomicAdd(left, 1u) == step - 1u) {
                atomicExch(tick, 0u);
                atomicExch(left, 0u);
            }
        }
        return;
    }
    for (uint32_t i = B.item_lo[x] + blockIdx.x / AG_NXCD; i < end; i += step) {
        const uint32_t first = B.item[2 * i], kc = B.item[2 * i + 1];
        if ((kc & 15u) == AG_ITEM_REG) {
            pull_one_item(U, upol, first, red, s_vals, s_wave);
        } else {
            blocked_tile(B, pp, red, s_vals, L, first, first + (kc >> 4), tt);
        }
    }
}

// ---------------------------------------------------------------- hot source segments

// LDS of the hot-segment table: 2^AG_HOT_SEG_BITS contributions of at most 4 bytes.
#define AG_HOT_TAB_BYTES ((1u << AG_HOT_SEG_BITS) * 4u)
static_assert(AG_HOT_NNZ <= AG_TILE_NNZ_MAX, "a hot tile is an ordinary tile too");
static_assert(AG_HOT_NNZ / (AG_LANE_MAX + 1u) < AG_LONG_MAX,
              "the long rows of a full hot tile must fit the tile's LDS list");

typedef uint32_t ag_u32x4 __attribute__((ext_vector_type(4)));

// stage_edges with the contributions read from the LDS copy `tab` of sources
// [base, base + 2^hot_bits): every source of the tile lies in that segment. The source ids
// (nnz <= AG_HOT_NNZ) arrive in 16-byte loads from the 16-byte boundary at or below e0, all
// of a lane's loads in flight at once; ids outside [e0, e0 + nnz) are dropped, and the last
// group of four is read id by id when it runs past the tile (it may run past the array).
template <class P, typename T>
__device__ __forceinline__ void stage_edges_tab(const uint32_t *src, uint32_t e0, uint32_t nnz,
                                                const P &pol, const T *tab, uint32_t base,
                                                typename P::Acc *s_vals)
{
    constexpr uint32_t NV = AG_HOT_NNZ / 4u / AG_WG + 1u;   // 4-id groups per lane, at most
    const uint32_t tid = threadIdx.x;
    const uint32_t a0 = e0 & ~3u, e1 = e0 + nnz;
    const uint32_t ngrp = (e1 - a0 + 3u) / 4u;
    ag_u32x4 q[NV];
#pragma unroll
    for (uint32_t u = 0; u < NV; u++) {
        const uint32_t v = tid + u * AG_WG, a = a0 + 4u * v;
        q[u] = ag_u32x4{base, base, base, base};
        if (v < ngrp) {
            if (a + 4u <= e1) {
                q[u] = __builtin_nontemporal_load((const ag_u32x4 *)(src + a));
            } else {
                for (uint32_t t = 0; t < 4u; t++)
                    if (a + t < e1)
                        q[u][t] = ld_stream(&src[a + t]);
            }
        }
    }
#pragma unroll
    for (uint32_t u = 0; u < NV; u++) {
        const uint32_t a = a0 + 4u * (tid + u * AG_WG);
#pragma unroll
        for (uint32_t t = 0; t < 4u; t++) {
            const uint32_t e = a + t;
            if (e >= e0 && e < e1)
                s_vals[e - e0] = pol.edge_val(e, tab[q[u][t] - base]);
        }
    }
}

// Hot pass of a source-blocked iteration (before pass 1, which then skips these tiles).
// Workgroup b serves slice b % AG_NXCD; per group it copies the group's segment of the
// contribution vector `cin` into LDS with coalesced 16-byte loads, then runs the group's
// tiles as pass 1 does, every gather an LDS read. Partials only: no row is finished here.
template <class P, typename T>
__device__ __forceinline__ void pull_hot_items(const BlockGeom &B, const P &pol, const T *cin,
                                               uint4 *s_tab, typename P::Acc *s_vals,
                                               uint32_t *s_long, uint32_t *s_nlong)
{
    static_assert(sizeof(T) <= 4 && 16 % sizeof(T) == 0,
                  "hot segments hold contributions of 1, 2 or 4 bytes");
    using Acc = typename P::Acc;
    const PartPol<P> pp{pol, (Acc *)B.part};
    Red red{};
    T *tab = (T *)s_tab;
    const uint32_t x = blockIdx.x % AG_NXCD, step = gridDim.x / AG_NXCD, tid = threadIdx.x;
    const uint32_t end = B.hgrp_off[x + 1];
    for (uint32_t gi = B.hgrp_off[x] + blockIdx.x / AG_NXCD; gi < end; gi += step) {
        const uint32_t i0 = B.hgrp[3 * gi], ni = B.hgrp[3 * gi + 1];
        const uint32_t base = B.hgrp[3 * gi + 2] << B.hot_bits;
        const uint32_t room = B.n_src - base;
        const uint32_t cnt = room < (1u << B.hot_bits) ? room : (1u << B.hot_bits);
        // base is a multiple of 2^hot_bits >= 256 sources: 16-byte aligned
        const uint32_t vec = cnt / (16u / (uint32_t)sizeof(T));
        const uint4 *g4 = (const uint4 *)(cin + base);
        for (uint32_t i = tid; i < vec; i += AG_WG)
            s_tab[i] = g4[i];
        for (uint32_t i = vec * (16u / (uint32_t)sizeof(T)) + tid; i < cnt; i += AG_WG)
            tab[i] = cin[base + i];
        __syncthreads();
        for (uint32_t it = i0; it < i0 + ni; it++) {
            const uint32_t first = B.item[2 * it], last = first + (B.item[2 * it + 1] >> 4);
            const uint32_t e0 = B.V.off[first];
            stage_edges_tab(B.V.src, e0, B.V.off[last] - e0, pol, (const T *)tab, base, s_vals);
            __syncthreads();
            // ends with a barrier: s_vals, s_long and (after the last tile) the table are free
            mixed_reduce(B.V.off, e0, first, last, pp, red, s_vals, s_long, s_nlong);
        }
    }
}

// Second pass. A combine tile (rows [r0, r1) of one slice) streams its partials block by
// block -- for each source block they are one contiguous run of virtual rows -- into LDS at
// their row-major slots (lpos, stored in this stream order, so read contiguously), then each
// row adds its partials in block order and applies
// the update. A row with more partials than a tile holds is summed by a workgroup through
// cidx. The items are spread over the whole grid: a slice's cost here is per row, and the
// slices are balanced for the first pass, not for this one.
template <class P>
__device__ __forceinline__ void pull_combine_items(const BlockGeom &B, const P &pol, Red &red,
                                                   typename P::Acc *s_vals,
                                                   typename P::Acc *s_wave, BlkLds &L)
{
    using Acc = typename P::Acc;
    constexpr int kU = sizeof(Acc) > 4 ? AG_CMB_UNROLL : AG_CMB_UNROLL_4B;
    uint32_t *s_pre = L.pre, *s_start = L.start;
    const CombPol<P> cp{pol, (const Acc *)B.part};
    const uint32_t nb = B.nblocks, tid = threadIdx.x;
    const uint32_t end = B.c_item_off[AG_NXCD];
    for (uint32_t i = blockIdx.x; i < end; i += gridDim.x) {
        const uint32_t first = B.c_item[2 * i], kind = B.c_item[2 * i + 1] & 15u;
        if (kind != AG_ITEM_CTILE) {
            pull_block_row(B.C, cp, first, red, s_wave);
            continue;
        }
        const uint32_t r0 = B.ct_row[2 * first], r1 = B.ct_row[2 * first + 1];
        const uint32_t p0 = B.C.off[r0], np = B.C.off[r1] - p0;  // np <= AG_TILE_NNZ_MAX
        const uint32_t *pre = B.ct_pre + (size_t)first * (nb + 1);
        const uint32_t *start = B.ct_start + (size_t)first * nb;
        for (uint32_t b = tid; b <= nb; b += AG_WG)
            s_pre[b] = pre[b];
        for (uint32_t b = tid; b < nb; b += AG_WG)
            s_start[b] = start[b];
        __syncthreads();
        // kU partials per lane in flight at once (their block searches are independent),
        // then their LDS writes
        for (uint32_t k0 = tid; k0 < np; k0 += kU * AG_WG) {
            Acc val[kU];
            uint32_t slot[kU];
#pragma unroll
            for (int u = 0; u < kU; u++) {
                const uint32_t k = k0 + u * AG_WG;
                if (k < np) {
                    // block of the k-th partial: s_pre[lo] <= k < s_pre[lo + 1] (s_pre[nb] = np)
                    uint32_t lo = 0, hi = nb;
                    while (hi - lo > 1) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (s_pre[mid] <= k)
                            lo = mid;
                        else
                            hi = mid;
                    }
                    val[u] = cp.part[s_start[lo] + (k - s_pre[lo])];
                    slot[u] = B.lpos[p0 + k];
                }
            }
#pragma unroll
            for (int u = 0; u < kU; u++)
                if (k0 + u * AG_WG < np)
                    s_vals[slot[u]] = val[u];
        }
        __syncthreads();
        mixed_reduce(B.C.off, p0, r0, r1, cp, red, s_vals, L.lng, &L.nlong);
    }
}

// Finish step of the hub rows: add each row's segment partials in segment order.
template <class P>
__device__ __forceinline__ void pull_hub_finish(const PullGeom &G, const P &pol, Red &red)
{
    using Acc = typename P::Acc;
    for (uint32_t h = threadIdx.x; h < G.n_hubs; h += blockDim.x) {
        const uint32_t s0 = G.hub_seg0[h], s1 = G.hub_seg0[h + 1];
        Acc total = acc_from_bits<Acc>(G.hub_part[s0]);
        for (uint32_t k = s0 + 1; k < s1; k++)
            total = total + acc_from_bits<Acc>(G.hub_part[k]);
        pol.finish(G.bin_row[AG_BIN_HUB] + h, total, red);
    }
}

// Plain pull (SpMV, raw gather): main persistent launch plus the hub finish.
template <class P>
__global__ __launch_bounds__(AG_WG) void pull_kernel(PullGeom G, P pol)
{
    __shared__ typename P::Acc s_vals[AG_TILE_NNZ_MAX];
    __shared__ typename P::Acc s_wave[AG_WG / AG_WAVE];
    Red red{};
    pull_items(G, pol, red, s_vals, s_wave);
}

// Plain source-blocked pull (SpMV): pass 1 with `pol` reading per-edge data in virtual-row
// order and `upol` in regular order for the unblocked rows, then the combine pass.
template <class P, class PU>
__global__ __launch_bounds__(AG_WG) void pull_blocked_kernel(PullGeom U, BlockGeom B, P pol,
                                                             PU upol)
{
    __shared__ typename P::Acc s_vals[AG_TILE_NNZ_MAX];
    __shared__ typename P::Acc s_wave[AG_WG / AG_WAVE];
    __shared__ BlkLds L;
    Red red{};
    pull_blocked_items(U, B, pol, upol, red, s_vals, s_wave, L);
}

template <class P>
__global__ __launch_bounds__(AG_WG) void pull_combine_kernel(BlockGeom B, P pol)
{
    __shared__ typename P::Acc s_vals[AG_TILE_NNZ_MAX];
    __shared__ typename P::Acc s_wave[AG_WG / AG_WAVE];
    __shared__ BlkLds L;
    Red red{};
    pull_combine_items(B, pol, red, s_vals, s_wave, L);
}

template <class P>
__global__ __launch_bounds__(AG_FIN_WG) void pull_finish_kernel(PullGeom G, P pol)
{
    Red red{};
    pull_hub_finish(G, pol, red);
}

// Storage switch of the 16-bit contribution phase (one thread): switch to f32 contributions
// from iteration iter + 1 once the L1 delta E is at most refine_l1, or once E has not improved
// on its best value by 1% for three iterations (the 16-bit rounding floor has been reached).
__device__ __forceinline__ void refine_decide(PrState *st, uint32_t iter, double E,
                                              double refine_l1)
{
    if (iter == 0 || E < 0.99 * st->best_err) {
        st->best_err = E;
        st->stall = 0;
    } else {
        st->stall++;
    }
    if (E <= refine_l1 || st->stall >= 3)
        st->refine_iter = (int32_t)iter + 1;
}
