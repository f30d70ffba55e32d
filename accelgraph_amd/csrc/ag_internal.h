// ag_internal.h -- shared definitions of the HIP runtime and kernels (not installed).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <time.h>

#include "accelgraph.h"
#include "../host/ag_plan.h"

#define AG_MAX_EVENTS 2
// Push bins (ag_push.h): rows per bin (the LDS row array of phase 2, and 16-bit local row
// ids) and slots per multi-row bin (one workgroup's work); a row with more in-edges than
// AG_PB_SLOTS is a bin by itself.
#ifndef AG_PB_ROWS
#define AG_PB_ROWS 4096u
#endif
#ifndef AG_PB_SLOTS
#define AG_PB_SLOTS 131072u
#endif
static_assert(AG_PB_ROWS >= 2u && AG_PB_ROWS <= 65536u, "push bins use 16-bit local row ids");

struct ag_device {
    int ordinal;
    hipStream_t stream;
    uint64_t run_timeout_ms;
    uint32_t poll_us;
    uint32_t live_graphs;   // graphs uploaded on this device and not yet freed
    uint32_t num_cus;       // compute units (persistent-grid sizing)
    void *scratch;          // this device's copy of the ag_spmv_pull_noalloc scratch array
};

// Device-resident graph: the plan arrays in HBM plus launch geometry.
struct ag_dgraph {
    ag_device *dev;
    ag_plan_params params;
    uint32_t n, m;
    uint32_t bin_row[AG_NBINS + 1];
    uint64_t bin_edges[AG_NBINS];
    uint32_t bin_blocks[AG_NBINS];
    uint32_t slot_blk0[AG_NSLOTS + 1];
    uint32_t total_blocks;
    uint32_t n_tiles, n_hubs, n_segs;
    uint64_t device_bytes;
    // HBM arrays
    uint32_t *d2o;       // [n]
    uint32_t *h_d2o;     // [n] host copy of the device-row map
    uint32_t *off;       // [n+1]
    uint32_t *src;       // [m]
    float *w;            // [m] or null
    uint32_t *wq;        // [m] FloatToFixed32(w), built on first fixed SpMV, or user codes
    int wq_user;         // wq came from ag_graph_set_fixed_weights
    uint32_t *outdeg;    // [n]
    uint32_t *tile_row;  // [n_tiles+1]
    uint32_t *seg_row, *seg_lo, *seg_hi;  // [n_segs]
    uint32_t *hub_seg0;  // [n_hubs+1]
    uint64_t *hub_part;    // [n_segs] 8-byte partial sums (float or u64 bits)
    // source blocking of the PageRank pull (blk.bits == 0: none); host copies of the
    // schedule scalars and HBM arrays of the block plan
    ag_block_plan blk;   // pointer members are device pointers here
    uint64_t *vpart;     // [blk.nv] partial sum of every virtual row (float or u64 bits)
    uint32_t *btick;     // [2 * AG_NXCD] pass-1 ticket counters (BlockGeom::tick), zeroed
    uint32_t bdyn;       // pass-1 tiles handed out by ticket (BlockGeom::dyn)
    uint32_t *vwq;       // [blk.mb] FloatToFixed32(vw), built on the first fixed SpMV
    // forward CSR in device row ids (out-edges), built on the first push / data-driven run
    uint32_t *fwd_off;   // [n+1]
    uint32_t *fwd_src;   // [m] source of every out-edge (row id of the forward CSR)
    uint32_t *fwd_dst;   // [m] destination, ascending within a row
    float *fwd_w;        // [m] weight of every out-edge (null for an unweighted graph)
    // push bins (ag_graph_push_bins): ranges of destination rows, and the slot of every
    // out-edge in the bin buffer (bins in row order, forward order inside a bin)
    uint32_t pb_nbins;
    uint32_t *pb_row;    // [pb_nbins+1] first row of every bin (one allocation with pb_off)
    uint32_t *pb_off;    // [pb_nbins+1] first slot of every bin
    uint32_t *pb_pos;    // [m] slot of every out-edge
    uint16_t *pb_dst;    // [m] row of every slot minus its bin's first row
    // scratch for permuted vectors (x, y) of SpMV / gather, sized lazily
    void *vec_a;
    void *vec_b;
    size_t vec_bytes;
    // personalised PageRank: buffers and the captured solve, kept between calls
    struct ag_ppr_ws *ppr;
};

// Frees g->ppr (ag_ppr.hip).
void ag_ppr_free_ws(ag_dgraph *g);

// Geometry passed by value to the pull kernels.
struct PullGeom {
    const uint32_t *off;
    const uint32_t *src;
    const uint32_t *tile_row;
    const uint32_t *seg_row;
    const uint32_t *seg_lo;
    const uint32_t *seg_hi;
    const uint32_t *hub_seg0;
    uint64_t *hub_part;
    uint32_t n;
    uint32_t n_hubs;
    uint32_t seq_hubs;       // plan in reference order: a hub row is one segment, added by one
                             // lane in sequence (chunks staged through LDS)
    uint32_t bin_row[AG_NBINS + 1];
    uint32_t slot_blk0[AG_NSLOTS + 1];
};

// Source-blocked pull: virtual rows, the combine schedule and the work items of each slice.
struct BlockGeom {
    PullGeom V;              // off = voff, src = vsrc (virtual rows; null with sorted tiles)
    PullGeom C;              // off = coff, src = cidx (a row's partial slots)
    const uint32_t *item;    // [2 * n_items] pass-1 items
    const uint32_t *c_item;  // [2 * c_n_items] combine items
    const uint32_t *ct_row;  // [2 * c_n_tiles]
    const uint32_t *ct_start;
    const uint32_t *ct_pre;
    const uint16_t *lpos;
    uint64_t *part;          // [nv]
    uint32_t nblocks;
    uint32_t item_off[AG_NXCD + 1];
    uint32_t c_item_off[AG_NXCD + 1];
    // pass 1 over the L2 takes the items [item_lo[x], item_off[x + 1]) of slice x: from
    // item_off[x] (every tile), or from the end of the hot-segment tiles when the hot pass
    // (groups hgrp, hgrp_off) gathers those from LDS
    uint32_t item_lo[AG_NXCD];
    const uint32_t *hgrp;    // [3 * n_hgrp] {first item, item count, segment}
    uint32_t hgrp_off[AG_NXCD + 1];
    uint32_t hot_bits;
    uint32_t n_src;          // source (vertex) count: the last segment may be short
    // pass 1 hands out its tiles through per-slice tickets (no AG_ITEM_REG item in pass 1):
    // tick[x] = next ticket of slice x, tick[AG_NXCD + x] = its workgroups that have left;
    // both zero between launches
    uint32_t dyn;
    uint32_t *tick;
    uint32_t top0;           // first source of the first-pass LDS table (n_src: none)
    // pass-1 tiles sorted by source (sbits != 0): vsp[e] = position in the tile << sbits |
    // (source - tbase[item]); the per-edge weights are then in the same order
    const uint32_t *vsp;
    const uint32_t *tbase;
    uint32_t sbits;
};

// Points pass 1 of B past the hot-segment tiles (when the hot pass runs before it).
void ag_block_geom_skip_hot(const ag_dgraph *g, BlockGeom &B);

// Persistent-grid size of a pull launch: resident workgroups per CU x CUs, capped by the
// number of work items. Per CU: the plan's wg_per_cu, else default_per_cu, never more than
// LDS allows.
uint32_t ag_pull_grid(const ag_dgraph *g, uint32_t lds_bytes, uint32_t default_per_cu = 4u);

// Device-side iteration state of a PageRank session.
struct PrState {
    uint32_t iter;          // iterations executed so far
    uint32_t done;          // set by the last block of the stopping iteration
    uint32_t stop_iter;     // index of the stopping iteration
    int32_t refine_iter;    // first iteration that writes f32 contributions (-1: none yet)
    uint32_t last_active;
    double last_err;
    double best_err;        // smallest L1 delta of the 16-bit phase so far
    uint32_t stall;         // 16-bit iterations since best_err last improved by 1%
    uint32_t qmax_bits;     // quantised kinds: bits of the largest next contribution
    float qscale;           // quantised kinds: scale of the codes being gathered
    uint32_t dd_count[2];   // data-driven pull: worklist lengths by iteration parity
    uint32_t dd_hubs[2];    // data-driven pull: hub rows listed at the worklist's far end
};

// A PageRank session (ag_pr_begin .. ag_pr_end).
struct ag_pr_session {
    ag_dgraph *g;
    ag_pr_options o;
    uint32_t kind_code;         // 0 f32, 1 bf16, 2 f16, 64/32/16/8
    size_t celem;               // bytes per stored contribution
    PrState *st;                // device
    PrState *st_host;           // pinned
    double *part_err;
    uint32_t *part_act;
    uint32_t *part_cmax;        // quantised kinds: per-workgroup largest contribution bits
    double *hist_err;
    uint32_t *hist_act;
    float *rank;
    void *c[2];
    void *c16[2];
    void *push_vals;            // push: the bin buffer, one value per out-edge slot [m]
    uint32_t *dd_queue[2];      // data-driven pull: worklists by iteration parity
    uint8_t *dd_flag;           // data-driven pull: next-worklist marks [n rounded up to 4]
    double *fin_part;
    double *fin_sum;
    float *user;
    uint32_t enqueued;
    hipEvent_t *ev;             // per-launch events when record_history (max_iterations + 1)
    hipEvent_t step_ev[2];
    uint32_t step_launches;
    uint32_t ev_count;
    struct timespec t0;  // wall clock at ag_pr_begin
};

// Enqueue one PageRank iteration of a session (main + finish kernel).
int ag_pr_launch_iteration(ag_pr_session *s);

// Builds g->fwd_off / fwd_src / fwd_dst (the forward CSR in device row ids) from the uploaded
// inverse CSR, once; later calls return at once.
int ag_graph_forward(ag_dgraph *g);

// Builds g->pb_* (push bins and slots, after the forward CSR), once; later calls return at once.
int ag_graph_push_bins(ag_dgraph *g);

int ag_set_error(int status, const char *fmt, ...);
int ag_hip_fail(hipError_t e, const char *what, const char *file, int line);
int ag_wait_stream(ag_device *dev);

#define AG_HIP(call)                                                   \
    do {                                                               \
        hipError_t ag_e_ = (call);                                     \
        if (ag_e_ != hipSuccess)                                       \
            return ag_hip_fail(ag_e_, #call, __FILE__, __LINE__);      \
    } while (0)

// graph_csr.c: parallel first touch of a new graph's host arrays (before bulk copies into them)
struct GraphCSR;
extern "C" void graphCSRPrefault(struct GraphCSR *g);

PullGeom ag_make_geom(const ag_dgraph *g);
// Regular schedule of the rows a source-blocked graph leaves unblocked (AG_ITEM_REG items).
PullGeom ag_make_unblocked_geom(const ag_dgraph *g);
BlockGeom ag_make_block_geom(const ag_dgraph *g);
// Grid of either pass of a source-blocked iteration: CUs x workgroups per CU (the plan's
// wg_per_cu, else default_per_cu, never more than resident_per_cu), a multiple of AG_NXCD. A
// persistent grid must fit on the device at once: a workgroup that waits for a free slot runs
// its whole share of the items after the others have finished. The gather pass defaults to 4
// per CU (it is bound by the rate of random L2 requests: more in flight only queue); the
// combine pass, whose reads are coalesced and whose time goes to dependent round trips,
// defaults to as many as fit, up to 8.
uint32_t ag_blocked_grid(const ag_dgraph *g, uint32_t resident_per_cu, uint32_t default_per_cu);

// Workgroups of `lanes` lanes of `kernel` that fit on one CU (registers and LDS).
template <typename K>
static inline uint32_t ag_resident_per_cu(K kernel, uint32_t lanes = AG_WG)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, (int)lanes, 0) != hipSuccess ||
        nb < 1)
        return 1u;
    return (uint32_t)nb;
}
int ag_ensure_vec(ag_dgraph *g, size_t bytes);
