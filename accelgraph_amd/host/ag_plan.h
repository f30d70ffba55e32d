/*
 * ag_plan.h -- host-side pull plan: the device layout of one graph.
 *
 * Rows (destination vertices) are grouped by in-degree into the AG_BIN_* classes and
 * renumbered so every class is a contiguous range of "device rows":
 *   [THREAD rows][GROUP16 rows][WAVE rows][BLOCK rows][HUB rows]
 * Sources inside each inverse row are renumbered with the same map and keep the reference's
 * in-row order, so a row's stream-order sum is unchanged by the renumbering. The thread bin
 * is cut into tiles of bounded rows/in-edges (staged through LDS by one workgroup); hub rows
 * are cut into fixed segments (one workgroup each, combined by the last one to finish).
 */
#ifndef AG_PLAN_H
#define AG_PLAN_H

#include <stdint.h>
#include "accelgraph.h"

#ifdef __cplusplus
extern "C" {
#endif

#define AG_TILE_ROWS_MAX 1024u
#define AG_TILE_NNZ_MAX 4096u
#define AG_WG 256u

/* launch order of the bins inside one fused launch: heaviest work first */
#define AG_NSLOTS AG_NBINS
extern const uint32_t ag_slot_bin[AG_NSLOTS];

/* Source blocking (PageRank pull and SpMV on large graphs). The blocked rows (bins >=
 * blocked_min_bin, the device-row suffix [row0, n)) are cut into AG_NXCD contiguous slices of
 * about equal in-edge counts, one per accelerator complex (XCD, which owns its own L2). Every
 * blocked row's in-edges are regrouped by source block (device id >> bits), keeping stream
 * order inside a block; each (row, block) group, split into pieces of at most
 * min(AG_PIECE_NNZ, tile_nnz) edges, is a "virtual row" whose sum is one partial. Virtual rows
 * are ordered by (slice, block, row, piece) and each slice's pass-1 tiles walk the blocks in
 * order, so at any time the gathers of one XCD fall in one block of the contribution vector.
 * The combine pass cuts each slice's rows into tiles; a tile streams its partials block by
 * block (rows are sorted inside a block, so they are contiguous), drops each into LDS at its
 * row-major position (lpos, stored in the order the tile streams its partials, so that it is
 * read contiguously), and adds every row's partials in block order. */
#define AG_NXCD 8u
#define AG_PIECE_NNZ 4096u
/* source blocks of one plan: the combine keeps two per-block words of its tile in LDS, so this
 * bounds every blocked kernel's LDS (2 KB here); the automatic block size grows to fit it */
#define AG_MAX_SRC_BLOCKS 256u
#define AG_CTILE_ROWS 4096u     /* rows of one combine tile */
#define AG_AUTO_BLOCK_BITS 19u
#define AG_AUTO_BLOCK_MIN_N (1u << 21)

/* Hot source segments: 2^AG_HOT_SEG_BITS consecutive device ids (fewer when the block bits
 * are fewer) whose in-edges are dense enough that a workgroup copies the segment's
 * contributions into LDS and gathers from there. Each hot segment is a source block of its
 * own (block ids [0, n_hot), in segment order; the ordinary blocks follow), its pieces are
 * cut at AG_HOT_NNZ edges, its pass-1 tiles hold at most AG_HOT_NNZ edges, and consecutive
 * tiles of one (slice, segment) form groups of at most AG_HOT_GROUP_NNZ edges, one LDS copy
 * of the segment each. */
#define AG_HOT_SEG_BITS 14u
#define AG_HOT_NNZ 2048u
#define AG_HOT_GROUP_NNZ 65536u
#define AG_HOT_DEFAULT 0u
/* sources of the first-pass LDS table (the last ids of the device order): 16 KB of FP32 */
#define AG_TOP_IDS 4096u

/* work items of the blocked passes: item = {first, kind | count << 4} */
enum { AG_ITEM_REG = 0,     /* pass 1: first = item of the unblocked rows' regular schedule */
       AG_ITEM_VTILE = 1,   /* pass 1: virtual rows [first, first + count) as one LDS tile  */
       AG_ITEM_CTILE = 2,   /* combine: tile `first`                                        */
       AG_ITEM_CBIG = 3 };  /* combine: row `first` (more partials than a tile holds), one
                               workgroup through cidx                                        */

typedef struct ag_block_plan {
    uint32_t bits;          /* 0: no source blocking (the fields below are empty)       */
    uint32_t nblocks;
    uint32_t row0;          /* first blocked device row                                 */
    uint32_t nv;            /* virtual rows                                             */
    uint32_t mb;            /* in-edges of the blocked rows                             */
    uint32_t slice_row[AG_NXCD + 1];
    uint32_t *voff;         /* [nv+1] offsets of the virtual rows in vsrc               */
    uint32_t *vsrc;         /* [mb]   device source ids, virtual-row order              */
    float *vw;              /* [mb]   weights in pass-1 stream order (beside vsrc, or beside
                               vsp when the tiles are sorted), or NULL                    */
    uint32_t n_items;
    uint32_t *item;         /* [2*n_items] pass-1 items                                 */
    uint32_t item_off[AG_NXCD + 1]; /* items of slice x: [item_off[x], item_off[x+1])   */
    uint32_t *coff;         /* [n+1]  partial slots of each row (empty below row0)      */
    uint32_t *cidx;         /* [nv]   virtual row of each slot, block order within a row */
    uint16_t *lpos;         /* [nv]   lpos[p0 + k]: row-major slot, relative to the tile, of
                               the k-th partial a combine tile streams (p0 = coff of its first
                               row); slots of rows combined through cidx are unused         */
    uint32_t c_n_tiles;
    uint32_t *ct_row;       /* [2*c_n_tiles] first and end row of each combine tile     */
    uint32_t *ct_start;     /* [c_n_tiles*nblocks] first virtual row of the tile in a block */
    uint32_t *ct_pre;       /* [c_n_tiles*(nblocks+1)] the tile's partials before block b */
    uint32_t c_n_items;
    uint32_t *c_item;       /* [2*c_n_items] combine items                              */
    uint32_t c_item_off[AG_NXCD + 1];
    /* regular schedule restricted to the unblocked rows (the AG_ITEM_REG items)         */
    uint32_t r_slot_blk0[AG_NSLOTS + 1];
    /* hot segments: the first items of slice x, [item_off[x], hot_end[x]), are the tiles of
     * its hot blocks; hgrp[3g .. 3g+2] = {first item, item count, segment} of group g, the
     * groups of slice x are [hgrp_off[x], hgrp_off[x+1])                                  */
    uint32_t hot_bits;      /* sources per hot segment = 2^hot_bits (0: no hot segments)  */
    uint32_t n_hot;
    uint64_t hot_edges;
    uint32_t *hot_seg;      /* [n_hot] segment (device id >> hot_bits) of hot block b     */
    uint32_t hot_end[AG_NXCD];
    uint32_t n_hgrp;
    uint32_t *hgrp;         /* [3*n_hgrp]                                                */
    uint32_t hgrp_off[AG_NXCD + 1];
    uint32_t piece;         /* most in-edges of one piece of an ordinary block            */
    /* first-pass LDS table of the sources [top0, n) (top0 == n: none)                     */
    uint32_t top0;
    uint64_t top_edges;     /* blocked in-edges whose source is >= top0                   */
    /* tile_sort (no hot segments): the ids of every pass-1 tile sorted by source; vsp[e] =
     * position of the edge in its tile << bits | (source - tbase[item]) (sorted == 0: vsp,
     * tbase NULL)                                                                         */
    uint32_t sorted;
    uint32_t *vsp;          /* [mb]      packed sorted ids, tile by tile                  */
    uint32_t *tbase;        /* [n_items] first source of the block of each pass-1 tile    */
} ag_block_plan;

typedef struct ag_plan {
    uint32_t n, m;
    ag_plan_params params;
    uint32_t *d2o;          /* [n]   device row -> original vertex              */
    uint32_t *o2d;          /* [n]   original vertex -> device row              */
    uint32_t *off;          /* [n+1] device-order inverse-CSR offsets           */
    uint32_t *src;          /* [m]   device ids of in-edge sources              */
    float *w;               /* [m]   in-edge weights, or NULL                   */
    uint32_t *outdeg;       /* [n]   forward out-degree, device order           */
    uint32_t bin_row[AG_NBINS + 1];
    uint64_t bin_edges[AG_NBINS];
    uint32_t n_tiles;
    uint32_t *tile_row;     /* [n_tiles+1] device row where each tile starts    */
    uint32_t n_hubs;
    uint32_t n_segs;
    uint32_t *seg_row;      /* [n_segs] device row of each hub segment          */
    uint32_t *seg_lo;       /* [n_segs] first in-edge of the segment            */
    uint32_t *seg_hi;       /* [n_segs] one past the last in-edge               */
    uint32_t *hub_seg0;     /* [n_hubs+1] first segment of each hub row         */
    uint32_t bin_blocks[AG_NBINS];
    uint32_t slot_blk0[AG_NSLOTS + 1]; /* first block of each launch slot       */
    uint32_t total_blocks;
    ag_block_plan blk;
} ag_plan;

/* Validates the inputs (every offset/degree in range, every source < n) and builds the
 * plan. Returns AG_OK, AG_ERR_INVALID or AG_ERR_OOM; *err receives a static message. */
int ag_plan_build(uint32_t n, uint32_t m, const uint32_t *out_degree,
                  const uint32_t *inv_edges_idx, const uint32_t *inv_degree,
                  const uint32_t *inv_src, const float *inv_weight,
                  const ag_plan_params *params, ag_plan *plan, const char **err);
void ag_plan_free(ag_plan *plan);
int ag_plan_params_check(const ag_plan_params *p, const char **err);

#ifdef __cplusplus
}
#endif

#endif
