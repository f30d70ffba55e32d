/*
 * plan.c -- builds the degree-binned pull plan (see ag_plan.h) from the reference's
 * inverse-CSR arrays (inverse_vertices->{edges_idx,out_degree},
 * inverse_sorted_edges_array->edges_array_dest; 02_capi_graph/src/capi_utils/capienv.c:
 * 858-868) and the forward out-degree the contribution step divides by
 * (02_capi_graph/src/algorithms/capi/pageRank.c:1029-1030).
 */
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#include "ag_plan.h"

const uint32_t ag_slot_bin[AG_NSLOTS] = {AG_BIN_HUB, AG_BIN_BLOCK, AG_BIN_WAVE, AG_BIN_GROUP16,
                                         AG_BIN_THREAD};

void ag_plan_params_default(ag_plan_params *p)
{
    p->thread_max_deg = 16;
    p->group16_max_deg = 128;
    p->wave_max_deg = 2048;
    p->block_max_deg = 16384;
    p->tile_rows = 1024;
    p->tile_nnz = 4096;
    p->segment_nnz = 16384;
    p->order = 0;
    p->src_block_bits = AG_SRC_BLOCK_AUTO;
    p->wg_per_cu = 0;
    p->blocked_min_bin = AG_BIN_THREAD;
    p->hot_segments = AG_HOT_DEFAULT;
    p->fp32_order = AG_ORDER_FAST;
    p->tile_order = AG_TILES_AUTO;
    p->piece_nnz = 0;
    p->top_table = AG_TOP_AUTO;
    p->tile_sort = AG_SORT_AUTO;
}

int ag_plan_params_check(const ag_plan_params *p, const char **err)
{
    if (p->tile_rows < 1 || p->tile_rows > AG_TILE_ROWS_MAX) {
        *err = "tile_rows must be in [1, 1024]";
        return AG_ERR_INVALID;
    }
    if (p->tile_nnz < 1 || p->tile_nnz > AG_TILE_NNZ_MAX) {
        *err = "tile_nnz must be in [1, 4096]";
        return AG_ERR_INVALID;
    }
    if (p->thread_max_deg > p->tile_nnz) {
        *err = "thread_max_deg must not exceed tile_nnz";
        return AG_ERR_INVALID;
    }
    if (!(p->thread_max_deg <= p->group16_max_deg && p->group16_max_deg <= p->wave_max_deg &&
          p->wave_max_deg <= p->block_max_deg)) {
        *err = "bin thresholds must be non-decreasing";
        return AG_ERR_INVALID;
    }
    if (p->segment_nnz < AG_WG || p->segment_nnz > (1u << 24)) {
        *err = "segment_nnz must be in [256, 2^24]";
        return AG_ERR_INVALID;
    }
    if (p->wg_per_cu > 32) {
        *err = "wg_per_cu must be in [0, 32]";
        return AG_ERR_INVALID;
    }
    if (p->src_block_bits != AG_SRC_BLOCK_AUTO && p->src_block_bits != AG_SRC_BLOCK_OFF &&
        (p->src_block_bits < 8 || p->src_block_bits > 30)) {
        *err = "src_block_bits must be AG_SRC_BLOCK_AUTO, AG_SRC_BLOCK_OFF or 8..30";
        return AG_ERR_INVALID;
    }
    if (p->blocked_min_bin > AG_NBINS) {
        *err = "blocked_min_bin must be in [0, 5]";
        return AG_ERR_INVALID;
    }
    if (p->order > 1) {
        *err = "order must be 0 (vertex id) or 1 (out-degree descending)";
        return AG_ERR_INVALID;
    }
    if (p->hot_segments > AG_MAX_SRC_BLOCKS) {
        *err = "hot_segments must be in [0, 256]";
        return AG_ERR_INVALID;
    }
    if (p->tile_order > AG_TILES_TICKET) {
        *err = "tile_order must be AG_TILES_AUTO, AG_TILES_ROUND_ROBIN or AG_TILES_TICKET";
        return AG_ERR_INVALID;
    }
    if (p->top_table > AG_TOP_ON) {
        *err = "top_table must be AG_TOP_AUTO, AG_TOP_OFF or AG_TOP_ON";
        return AG_ERR_INVALID;
    }
    if (p->tile_sort > AG_SORT_ON) {
        *err = "tile_sort must be AG_SORT_AUTO, AG_SORT_OFF or AG_SORT_ON";
        return AG_ERR_INVALID;
    }
    if (p->piece_nnz > p->tile_nnz) {
        *err = "piece_nnz must be 0 (auto) or in [1, tile_nnz]";
        return AG_ERR_INVALID;
    }
    if (p->fp32_order > AG_ORDER_REFERENCE) {
        *err = "fp32_order must be AG_ORDER_FAST or AG_ORDER_REFERENCE";
        return AG_ERR_INVALID;
    }
    return AG_OK;
}

void ag_plan_free(ag_plan *p)
{
    if (!p)
        return;
    free(p->d2o);
    free(p->o2d);
    free(p->off);
    free(p->src);
    free(p->w);
    free(p->outdeg);
    free(p->tile_row);
    free(p->seg_row);
    free(p->seg_lo);
    free(p->seg_hi);
    free(p->hub_seg0);
    free(p->blk.voff);
    free(p->blk.vsrc);
    free(p->blk.vw);
    free(p->blk.item);
    free(p->blk.coff);
    free(p->blk.cidx);
    free(p->blk.lpos);
    free(p->blk.ct_row);
    free(p->blk.ct_start);
    free(p->blk.ct_pre);
    free(p->blk.c_item);
    free(p->blk.hot_seg);
    free(p->blk.hgrp);
    free(p->blk.vsp);
    free(p->blk.tbase);
    memset(p, 0, sizeof(*p));
}

static int bin_of(const ag_plan_params *p, uint32_t deg)
{
    if (deg <= p->thread_max_deg)
        return AG_BIN_THREAD;
    if (deg <= p->group16_max_deg)
        return AG_BIN_GROUP16;
    if (deg <= p->wave_max_deg)
        return AG_BIN_WAVE;
    if (deg <= p->block_max_deg)
        return AG_BIN_BLOCK;
    return AG_BIN_HUB;
}

static int cmp_u64(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return (x > y) - (x < y);
}

static void *zalloc(size_t count, size_t size)
{
    return calloc(count ? count : 1, size);
}

/* Greedy cut of rows [lo, hi) into tiles of at most rows_max rows and nnz_max in-edges
 * (a single row above nnz_max is never cut; callers keep such rows out). Returns the
 * [*count + 1] tile start rows, or NULL when out of memory. */
static uint32_t *cut_tiles(const uint32_t *off, uint32_t lo, uint32_t hi, uint32_t rows_max,
                           uint32_t nnz_max, uint32_t *count)
{
    uint32_t *tr = (uint32_t *)zalloc((size_t)(hi - lo) + 2, sizeof(uint32_t));
    if (!tr)
        return NULL;
    uint32_t t = 0, r = lo;
    while (r < hi) {
        tr[t++] = r;
        uint32_t rows = 0, nnz = 0;
        while (r < hi && rows < rows_max) {
            uint32_t d = off[r + 1] - off[r];
            if (rows > 0 && nnz + d > nnz_max)
                break;
            nnz += d;
            rows++;
            r++;
        }
    }
    tr[t] = hi;
    *count = t;
    return tr;
}

/* ---------------------------------------------------------------- source blocking */

/* Effective block bits (0: no blocking). *err is set when explicit bits give too many blocks. */
static uint32_t block_bits_for(const ag_plan *p, const char **err)
{
    uint32_t b = p->params.src_block_bits;
    if (b == AG_SRC_BLOCK_OFF || p->params.blocked_min_bin >= AG_NBINS)
        return 0;
    if (b == AG_SRC_BLOCK_AUTO) {
        if (p->n <= AG_AUTO_BLOCK_MIN_N)
            return 0;
        b = AG_AUTO_BLOCK_BITS;
        while ((((uint64_t)p->n + (1ull << b) - 1) >> b) > AG_MAX_SRC_BLOCKS)
            b++;
        return b;
    }
    if ((((uint64_t)p->n + (1ull << b) - 1) >> b) > AG_MAX_SRC_BLOCKS)
        *err = "src_block_bits too small for this graph: more than 256 source blocks";
    return b;
}

/* Source block of device id s: kmap[s >> kbits] (hot segments first, then the ordinary
 * blocks; see build_block_plan). */
typedef struct {
    const uint32_t *kmap;
    uint32_t kbits;
    uint32_t n_hot;         /* blocks below this are hot segments */
    uint32_t piece;         /* piece limit of an ordinary block */
    uint32_t hot_piece;     /* piece limit of a hot segment */
} blk_key;

static inline uint32_t key_of(const blk_key *K, uint32_t s) { return K->kmap[s >> K->kbits]; }

/* Stable sort of one row's in-edges by source block into out_s/out_w. keys: scratch of d. */
static void sort_row_by_block(const uint32_t *src, const float *w, uint32_t d, const blk_key *K,
                              uint32_t *out_s, float *out_w, uint64_t *keys)
{
    if (d <= 32) {
        for (uint32_t i = 0; i < d; i++) {
            uint32_t s = src[i];
            float x = w ? w[i] : 0.0f;
            const uint32_t ks = key_of(K, s);
            uint32_t j = i;
            while (j > 0 && key_of(K, out_s[j - 1]) > ks) {
                out_s[j] = out_s[j - 1];
                if (w)
                    out_w[j] = out_w[j - 1];
                j--;
            }
            out_s[j] = s;
            if (w)
                out_w[j] = x;
        }
        return;
    }
    for (uint32_t i = 0; i < d; i++)
        keys[i] = ((uint64_t)key_of(K, src[i]) << 32) | i;
    qsort(keys, d, sizeof(uint64_t), cmp_u64);
    for (uint32_t i = 0; i < d; i++) {
        uint32_t k = (uint32_t)keys[i];
        out_s[i] = src[k];
        if (w)
            out_w[i] = w[k];
    }
}

/* The virtual rows of a block-sorted row: one per run of equal source block, split into
 * pieces of at most the block's piece limit. With st != NULL records start, length and block
 * of each. Returns the count. */
static uint32_t row_pieces(const uint32_t *ss, uint32_t d, const blk_key *K, uint32_t *st,
                           uint32_t *len, uint32_t *blk)
{
    uint32_t cnt = 0, i = 0;
    while (i < d) {
        uint32_t b = key_of(K, ss[i]), j = i + 1;
        while (j < d && key_of(K, ss[j]) == b)
            j++;
        const uint32_t piece = b < K->n_hot ? K->hot_piece : K->piece;
        for (uint32_t a = i; a < j; a += piece) {
            if (st) {
                st[cnt] = a;
                len[cnt] = (j - a > piece) ? piece : j - a;
                blk[cnt] = b;
            }
            cnt++;
        }
        i = j;
    }
    return cnt;
}

#define AG_EMIT(arr, k, first, kind, cnt)                                   \
    do {                                                                    \
        if (arr) {                                                          \
            (arr)[2 * (k)] = (first);                                       \
            (arr)[2 * (k) + 1] = (uint32_t)(kind) | ((uint32_t)(cnt) << 4); \
        }                                                                   \
        (k)++;                                                              \
    } while (0)

/* L2 lines of 32 sources (128 bytes of FP32 contributions) that a row's pieces gather: per
 * piece one plus the changes of line along it (a piece's sources are nearly in order, so this
 * is about its distinct lines). A wavefront's gather issues one L2 request per distinct line,
 * and on skewed graphs this, not the edge count, sets a tile's cost. */
static uint32_t row_lines(const uint32_t *ss, uint32_t d, const blk_key *K)
{
    uint32_t cnt = 0, i = 0;
    while (i < d) {
        uint32_t b = key_of(K, ss[i]), j = i + 1;
        while (j < d && key_of(K, ss[j]) == b)
            j++;
        const uint32_t piece = b < K->n_hot ? K->hot_piece : K->piece;
        for (uint32_t a = i; a < j; a += piece) {
            const uint32_t e = j - a > piece ? a + piece : j;
            cnt++;
            for (uint32_t k = a + 1; k < e; k++)
                cnt += (ss[k] >> 5) != (ss[k - 1] >> 5);
        }
        i = j;
    }
    return cnt;
}

/* The ids of every pass-1 tile sorted by source, stably (equal sources keep their order),
 * each packed with its position in the tile: vsp[e0 + k] = pos << bits | (src - base), base
 * = the tile's block start, kept in tbase. A tile never crosses a block key, so its sources
 * lie in one aligned block of 2^bits (checked; a tile that did not would leave the plan
 * unsorted); bits <= 20 and pos < AG_TILE_NNZ_MAX = 2^12 fill at most 32 bits. Per tile a
 * two-digit LSD counting sort of pos << 20 | (src - base) on its low 20 bits. The weights
 * (vw) are put in the same order, so pass 1 reads them beside the packed ids. */
static int sort_tile_ids(ag_block_plan *B)
{
    const uint32_t bits = B->bits, mask = (1u << bits) - 1u;
    const uint32_t n_items = B->n_items;
    B->vsp = (uint32_t *)zalloc(B->mb ? B->mb : 1, sizeof(uint32_t));
    B->tbase = (uint32_t *)zalloc(n_items ? n_items : 1, sizeof(uint32_t));
    if (!B->vsp || !B->tbase)
        return AG_ERR_OOM;
    int bad = 0, oom = 0;
    #pragma omp parallel reduction(|:bad, oom)
    {
        uint32_t *a = (uint32_t *)malloc(2 * AG_TILE_NNZ_MAX * sizeof(uint32_t));
        uint32_t *cnt = (uint32_t *)malloc(1024 * sizeof(uint32_t));
        if (!a || !cnt)
            oom = 1;
        #pragma omp for schedule(dynamic, 64)
        for (uint32_t i = 0; i < n_items; i++) {
            if (oom || (B->item[2 * i + 1] & 15u) != AG_ITEM_VTILE)
                continue;
            const uint32_t first = B->item[2 * i], last = first + (B->item[2 * i + 1] >> 4);
            const uint32_t e0 = B->voff[first], nnz = B->voff[last] - e0;
            const uint32_t base = nnz ? B->vsrc[e0] & ~mask : 0;
            uint32_t *x = a, *y = a + AG_TILE_NNZ_MAX;
            for (uint32_t k = 0; k < nnz; k++) {
                const uint32_t s = B->vsrc[e0 + k];
                bad |= (s & ~mask) != base;
                x[k] = k << 20 | (s & mask);
            }
            for (uint32_t sh = 0; sh < 20; sh += 10) {
                memset(cnt, 0, 1024 * sizeof(uint32_t));
                for (uint32_t k = 0; k < nnz; k++)
                    cnt[(x[k] >> sh) & 1023u]++;
                for (uint32_t d = 0, run = 0; d < 1024; d++) {
                    const uint32_t c = cnt[d];
                    cnt[d] = run;
                    run += c;
                }
                for (uint32_t k = 0; k < nnz; k++)
                    y[cnt[(x[k] >> sh) & 1023u]++] = x[k];
                uint32_t *t = x;
                x = y;
                y = t;
            }
            for (uint32_t k = 0; k < nnz; k++)
                B->vsp[e0 + k] = (x[k] >> 20) << bits | (x[k] & 0xFFFFFu);
            B->tbase[i] = base;
        }
        free(a);
        free(cnt);
    }
    /* the weights in the same order, once every tile is known to be sorted */
    if (!oom && !bad && B->vw) {
        #pragma omp parallel reduction(|:oom)
        {
            float *tw = (float *)malloc(AG_TILE_NNZ_MAX * sizeof(float));
            if (!tw)
                oom = 1;
            #pragma omp for schedule(dynamic, 64)
            for (uint32_t i = 0; i < n_items; i++) {
                if (oom || (B->item[2 * i + 1] & 15u) != AG_ITEM_VTILE)
                    continue;
                const uint32_t first = B->item[2 * i], last = first + (B->item[2 * i + 1] >> 4);
                const uint32_t e0 = B->voff[first], nnz = B->voff[last] - e0;
                for (uint32_t k = 0; k < nnz; k++)
                    tw[k] = B->vw[e0 + (B->vsp[e0 + k] >> bits)];
                memcpy(B->vw + e0, tw, (size_t)nnz * sizeof(float));
            }
            free(tw);
        }
    }
    if (oom)
        return AG_ERR_OOM;
    if (bad) {   /* unreachable while tiles keep to one block */
        free(B->vsp);
        free(B->tbase);
        B->vsp = B->tbase = NULL;
        return AG_OK;
    }
    B->sorted = 1;
    return AG_OK;
}

/* Pass-1 items of every slice: per block, greedy tiles of consecutive virtual rows (at most
 * tile_nnz edges, so at most that many virtual rows; at most hot_nnz in a hot segment), then
 * the slice's share of the unblocked rows' items. The hot segments' tiles come first; runs
 * of them of one segment, at most AG_HOT_GROUP_NNZ edges, form the groups hgrp (one LDS copy
 * of the segment each). item == NULL only counts them (and the groups, into *n_hgrp). */
static uint32_t emit_items(const ag_plan *p, const uint32_t *kstart, const uint32_t *vlen,
                           uint32_t reg_total, uint32_t hot_nnz, uint32_t *item,
                           uint32_t *item_off, uint32_t *hgrp, uint32_t *n_hgrp)
{
    const ag_block_plan *B = &p->blk;
    const uint32_t nb = B->nblocks, H = B->n_hot;
    uint32_t k = 0, ng = 0;
    for (uint32_t x = 0; x < AG_NXCD; x++) {
        if (item_off) {
            item_off[x] = k;
            ((ag_block_plan *)B)->hgrp_off[x] = ng;
        }
        for (uint32_t b = 0; b < nb; b++) {
            if (b == H && item_off)
                ((ag_block_plan *)B)->hot_end[x] = k;
            const uint32_t lim = b < H ? hot_nnz : p->params.tile_nnz;
            uint32_t v = kstart[x * nb + b];
            const uint32_t e = kstart[x * nb + b + 1];
            uint32_t gsz = 0, gitems = 0;
            while (v < e) {
                uint32_t t0 = v, rows = 0, nnz = 0;
                while (v < e) {
                    if (rows > 0 && nnz + vlen[v] > lim)
                        break;
                    nnz += vlen[v];
                    rows++;
                    v++;
                }
                if (b < H) {
                    if (gitems > 0 && gsz + nnz > AG_HOT_GROUP_NNZ) {
                        if (hgrp) {
                            hgrp[3 * ng] = k - gitems;
                            hgrp[3 * ng + 1] = gitems;
                            hgrp[3 * ng + 2] = B->hot_seg[b];
                        }
                        ng++;
                        gsz = gitems = 0;
                    }
                    gsz += nnz;
                    gitems++;
                }
                AG_EMIT(item, k, t0, AG_ITEM_VTILE, rows);
            }
            if (gitems > 0) {
                if (hgrp) {
                    hgrp[3 * ng] = k - gitems;
                    hgrp[3 * ng + 1] = gitems;
                    hgrp[3 * ng + 2] = B->hot_seg[b];
                }
                ng++;
            }
        }
        if (H == nb && item_off)
            ((ag_block_plan *)B)->hot_end[x] = k;
        for (uint32_t r = x; r < reg_total; r += AG_NXCD)
            AG_EMIT(item, k, r, AG_ITEM_REG, 0);
    }
    if (item_off) {
        item_off[AG_NXCD] = k;
        ((ag_block_plan *)B)->hgrp_off[AG_NXCD] = ng;
    }
    *n_hgrp = ng;
    return k;
}

/* Combine tiles and items of every slice: runs of rows with at most tile_nnz partials cut
 * into tiles of at most AG_CTILE_ROWS rows and tile_nnz partials; a row with more partials is
 * an AG_ITEM_CBIG item of its own. ct_row/item == NULL only count. */
static uint32_t emit_combine(const ag_plan *p, uint32_t *ct_row, uint32_t *n_tiles,
                             uint32_t *item, uint32_t *item_off)
{
    const ag_block_plan *B = &p->blk;
    const uint32_t *coff = B->coff;
    uint32_t k = 0, t = 0;
    for (uint32_t x = 0; x < AG_NXCD; x++) {
        if (item_off)
            item_off[x] = k;
        uint32_t r = B->slice_row[x];
        const uint32_t hi = B->slice_row[x + 1];
        while (r < hi) {
            if (coff[r + 1] - coff[r] > p->params.tile_nnz) {
                AG_EMIT(item, k, r, AG_ITEM_CBIG, 0);
                r++;
                continue;
            }
            uint32_t r0 = r, rows = 0, nnz = 0;
            while (r < hi && rows < AG_CTILE_ROWS) {
                uint32_t c = coff[r + 1] - coff[r];
                if (c > p->params.tile_nnz || (rows > 0 && nnz + c > p->params.tile_nnz))
                    break;
                nnz += c;
                rows++;
                r++;
            }
            if (ct_row) {
                ct_row[2 * t] = r0;
                ct_row[2 * t + 1] = r;
            }
            AG_EMIT(item, k, t, AG_ITEM_CTILE, 0);
            t++;
        }
    }
    if (item_off)
        item_off[AG_NXCD] = k;
    *n_tiles = t;
    return k;
}
#undef AG_EMIT

/* first index i in [lo, hi) with key[i] >= r (key sorted on the range) */
static uint32_t lower_bound_u32(const uint32_t *key, uint32_t lo, uint32_t hi, uint32_t r)
{
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (key[mid] < r)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

/* Hot segments among the sources of the blocked rows [row0, n): for every segment of 2^kb
 * device ids, its in-edges E and its pieces P (rows with at least one in-edge from it). A
 * segment qualifies when its pieces average at least two edges (2P <= E: staging it in LDS
 * turns at least half of its gathers into LDS reads) and it has at least four in-edges per
 * source (the LDS copy of the segment costs less than the gathers it serves). The hmax
 * qualifying segments with the most in-edges are numbered 0..H-1 in segment order in kmap;
 * every other entry of kmap is set to UINT32_MAX. */
static int select_hot_segments(const ag_plan *p, uint32_t row0, uint32_t kb, uint32_t nseg,
                               uint32_t hmax, uint32_t *kmap, uint32_t *H_out,
                               uint64_t *hot_edges)
{
    const uint32_t n = p->n;
    for (uint32_t sg = 0; sg < nseg; sg++)
        kmap[sg] = UINT32_MAX;
    *H_out = 0;
    *hot_edges = 0;
    if (hmax == 0)
        return AG_OK;
    const int nt = omp_get_max_threads() > 0 ? omp_get_max_threads() : 1;
    uint64_t *E = (uint64_t *)zalloc((size_t)nt * nseg, sizeof(uint64_t));
    uint32_t *P = (uint32_t *)zalloc((size_t)nt * nseg, sizeof(uint32_t));
    uint32_t *stamp = (uint32_t *)malloc((size_t)nt * nseg * sizeof(uint32_t));
    uint64_t *cand = (uint64_t *)malloc((size_t)nseg * sizeof(uint64_t));
    if (!E || !P || !stamp || !cand) {
        free(E);
        free(P);
        free(stamp);
        free(cand);
        return AG_ERR_OOM;
    }
    memset(stamp, 0xFF, (size_t)nt * nseg * sizeof(uint32_t));
    #pragma omp parallel num_threads(nt)
    {
        const int t = omp_get_thread_num();
        uint64_t *e = E + (size_t)t * nseg;
        uint32_t *pc = P + (size_t)t * nseg, *sp = stamp + (size_t)t * nseg;
        #pragma omp for schedule(dynamic, 1024)
        for (uint32_t r = row0; r < n; r++) {
            for (uint32_t j = p->off[r]; j < p->off[r + 1]; j++) {
                const uint32_t sg = p->src[j] >> kb;
                e[sg]++;
                if (sp[sg] != r) {
                    sp[sg] = r;
                    pc[sg]++;
                }
            }
        }
    }
    uint32_t nc = 0;
    for (uint32_t sg = 0; sg < nseg; sg++) {
        uint64_t es = 0, ps = 0;
        for (int t = 0; t < nt; t++) {
            es += E[(size_t)t * nseg + sg];
            ps += P[(size_t)t * nseg + sg];
        }
        if (es >= (4ull << kb) && 2 * ps <= es && es < (1ull << 32))
            cand[nc++] = (es << 32) | (uint64_t)(UINT32_MAX - sg);   /* most edges, then lowest id */
    }
    qsort(cand, nc, sizeof(uint64_t), cmp_u64);
    const uint32_t H = nc < hmax ? nc : hmax;
    uint64_t he = 0;
    for (uint32_t i = 0; i < H; i++) {
        const uint64_t c = cand[nc - 1 - i];
        kmap[UINT32_MAX - (uint32_t)(c & 0xFFFFFFFFu)] = 0;   /* mark; numbered below */
        he += c >> 32;
    }
    uint32_t id = 0;
    for (uint32_t sg = 0; sg < nseg; sg++)
        if (kmap[sg] == 0)
            kmap[sg] = id++;
    *H_out = H;
    *hot_edges = he;
    free(E);
    free(P);
    free(stamp);
    free(cand);
    return AG_OK;
}

static int build_block_plan(ag_plan *p, const char **err)
{
    ag_block_plan *B = &p->blk;
    memset(B, 0, sizeof(*B));
    const char *e = NULL;
    const uint32_t bits = block_bits_for(p, &e);
    if (e) {
        *err = e;
        return AG_ERR_INVALID;
    }
    const uint32_t n = p->n, min_bin = p->params.blocked_min_bin;
    if (!bits || min_bin >= AG_NBINS)
        return AG_OK;
    const uint32_t row0 = p->bin_row[min_bin];
    if (row0 >= n || p->off[n] == p->off[row0])
        return AG_OK;
    const uint32_t e0 = p->off[row0], mb = p->off[n] - e0;
    const uint32_t ncold = (uint32_t)((((uint64_t)n) + (1ull << bits) - 1) >> bits);
    const ag_plan_params *q = &p->params;
    const uint32_t piece = q->piece_nnz ? q->piece_nnz
                                        : (q->tile_nnz < AG_PIECE_NNZ ? q->tile_nnz : AG_PIECE_NNZ);
    const uint32_t hot_nnz = q->tile_nnz < AG_HOT_NNZ ? q->tile_nnz : AG_HOT_NNZ;
    int oom = 0;

    /* 0: source blocks. Segments of 2^kb device ids; the hot ones are blocks of their own,
     * every other segment maps to its ordinary 2^bits block. */
    const uint32_t kb = bits < AG_HOT_SEG_BITS ? bits : AG_HOT_SEG_BITS;
    const uint32_t nseg = (uint32_t)((((uint64_t)n) + (1ull << kb) - 1) >> kb);
    uint32_t *kmap = (uint32_t *)zalloc(nseg, sizeof(uint32_t));
    if (!kmap) {
        *err = "host allocation failed";
        return AG_ERR_OOM;
    }
    uint32_t H = 0;
    {
        const uint32_t hmax = q->hot_segments < AG_MAX_SRC_BLOCKS - ncold ? q->hot_segments
                                                                           : AG_MAX_SRC_BLOCKS - ncold;
        int st = select_hot_segments(p, row0, kb, nseg, hmax, kmap, &H, &B->hot_edges);
        if (st != AG_OK) {
            free(kmap);
            *err = "host allocation failed";
            return st;
        }
    }
    if (H) {
        B->hot_seg = (uint32_t *)zalloc(H, sizeof(uint32_t));
        if (!B->hot_seg) {
            free(kmap);
            *err = "host allocation failed";
            return AG_ERR_OOM;
        }
        for (uint32_t sg = 0; sg < nseg; sg++)
            if (kmap[sg] != UINT32_MAX)
                B->hot_seg[kmap[sg]] = sg;
    }
    for (uint32_t sg = 0; sg < nseg; sg++)
        if (kmap[sg] == UINT32_MAX)
            kmap[sg] = H + (sg >> (bits - kb));
    const uint32_t nb = H + ncold;
    const uint32_t nK = AG_NXCD * nb;
    blk_key K = {kmap, kb, H, piece, hot_nnz};

    B->bits = bits;
    B->nblocks = nb;
    B->row0 = row0;
    B->mb = mb;
    B->n_hot = H;
    B->hot_bits = H ? kb : 0;

    uint32_t *ss = (uint32_t *)zalloc(mb, sizeof(uint32_t));
    float *sw = p->w ? (float *)zalloc(mb, sizeof(float)) : NULL;
    B->coff = (uint32_t *)zalloc((size_t)n + 1, sizeof(uint32_t));
    uint32_t *nvr = (uint32_t *)zalloc((size_t)(n - row0), sizeof(uint32_t));
    uint32_t *lnr = (uint32_t *)zalloc((size_t)(n - row0), sizeof(uint32_t));
    if (!ss || (p->w && !sw) || !B->coff || !nvr || !lnr) {
        free(ss);
        free(sw);
        free(nvr);
        free(lnr);
        free(kmap);
        *err = "host allocation failed";
        return AG_ERR_OOM;
    }
    /* 1: sort each blocked row by source block; count its virtual rows */
    #pragma omp parallel
    {
        uint64_t *keys = NULL;
        size_t cap = 0;
        #pragma omp for schedule(dynamic, 256)
        for (uint32_t r = row0; r < n; r++) {
            const uint32_t d = p->off[r + 1] - p->off[r], o = p->off[r] - e0;
            if (d > 32 && d > cap) {
                free(keys);
                cap = d;
                keys = (uint64_t *)malloc(cap * sizeof(uint64_t));
                if (!keys) {
                    cap = 0;
                    #pragma omp atomic write
                    oom = 1;
                    continue;
                }
            }
            sort_row_by_block(p->src + p->off[r], p->w ? p->w + p->off[r] : NULL, d, &K,
                              ss + o, sw ? sw + o : NULL, keys);
            nvr[r - row0] = row_pieces(ss + o, d, &K, NULL, NULL, NULL);
            lnr[r - row0] = row_lines(ss + o, d, &K);
        }
        free(keys);
    }
    if (oom) {
        free(ss);
        free(sw);
        free(nvr);
        free(lnr);
        free(kmap);
        *err = "host allocation failed";
        return AG_ERR_OOM;
    }
    {
        uint64_t acc = 0;
        for (uint32_t r = row0; r < n; r++)
            acc += nvr[r - row0];
        /* automatic piece limit: when the pieces are nearly as many as the in-edges (sources
         * spread evenly over the blocks), every in-edge becomes a piece of its own, so that
         * the first pass is a plain gather with no per-piece offsets to read */
        if (!q->piece_nnz && K.piece > 1 && acc * 4 >= (uint64_t)mb * 3) {
            K.piece = 1;
            #pragma omp parallel for schedule(dynamic, 256)
            for (uint32_t r = row0; r < n; r++) {
                const uint32_t *rs = ss + (p->off[r] - e0), d = p->off[r + 1] - p->off[r];
                nvr[r - row0] = row_pieces(rs, d, &K, NULL, NULL, NULL);
                lnr[r - row0] = row_lines(rs, d, &K);
            }
        }
        acc = 0;
        for (uint32_t r = 0; r <= row0; r++)
            B->coff[r] = 0;
        for (uint32_t r = row0; r < n; r++) {
            acc += nvr[r - row0];
            B->coff[r + 1] = (uint32_t)acc;
        }
        B->nv = (uint32_t)acc;
        B->piece = K.piece;
    }
    free(nvr);
    const uint32_t nv = B->nv;
    /* slices: contiguous blocked rows of about equal pass-1 cost, per in-edge (its source id,
     * its gather and LDS traffic), per gathered line of a piece (an L2 request) and per
     * virtual row (its offset reads and partial write), weights fitted to the measured
     * pass-1 time of each slice of RMAT-25. Tiles read in stream order: 10 / 17 / 11 (slices
     * within 3%; an edge count alone left the slice of light rows 30% faster). Tiles sorted
     * by source (their ids then share lines across rows, so a piece's own line count no
     * longer predicts its requests): 100 / 0 / 46 -- 10 / 17 / 11 left the two slices of
     * the heaviest rows 7% slower than the rest (RMAT-25 iteration 2.357 ms; 100 / 0 / 46:
     * 2.317 ms, and 30 or 62 per virtual row 2.37 / 2.35 ms). */
    {
        uint64_t *cpre = (uint64_t *)malloc(((size_t)(n - row0) + 1) * sizeof(uint64_t));
        if (!cpre) {
            free(lnr);
            free(ss);
            free(sw);
            free(kmap);
            *err = "host allocation failed";
            return AG_ERR_OOM;
        }
        cpre[0] = 0;
        const int sorted = bits <= 20 && !H && q->tile_sort != AG_SORT_OFF;   /* as step 6 */
        const uint64_t wE = sorted ? 100 : 10, wL = sorted ? 0 : 17, wP = sorted ? 46 : 11;
        for (uint32_t r = row0; r < n; r++)
            cpre[r - row0 + 1] = cpre[r - row0] + wE * (p->off[r + 1] - p->off[r]) +
                                 wL * lnr[r - row0] + wP * (B->coff[r + 1] - B->coff[r]);
        const uint64_t total = cpre[n - row0];
        B->slice_row[0] = row0;
        for (uint32_t x = 1; x < AG_NXCD; x++) {
            const uint64_t target = total * x / AG_NXCD;
            uint32_t lo = B->slice_row[x - 1], hi = n;
            while (lo < hi) {
                uint32_t mid = lo + (hi - lo) / 2;
                if (cpre[mid - row0] < target)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            B->slice_row[x] = lo;
        }
        B->slice_row[AG_NXCD] = n;
        free(cpre);
    }
    free(lnr);
    /* first-pass LDS table: the last AG_TOP_IDS device ids, when they are gathered often */
    {
        const uint32_t t0 = n > AG_TOP_IDS ? n - AG_TOP_IDS : 0;
        uint64_t te = 0;
        #pragma omp parallel for reduction(+ : te) schedule(static)
        for (uint32_t i = 0; i < mb; i++)
            te += p->src[e0 + i] >= t0;
        B->top_edges = te;
        const uint32_t mode = q->top_table;
        B->top0 = (mode == AG_TOP_ON || (mode == AG_TOP_AUTO && te * 50 >= mb)) ? t0 : n;
        if (B->top0 == n)
            B->top_edges = 0;
    }

    /* 2: every virtual row in (row, block, piece) slot order: key, start in ss, length, row */
    uint32_t *vkey = (uint32_t *)zalloc(nv, sizeof(uint32_t));
    uint32_t *vst = (uint32_t *)zalloc(nv, sizeof(uint32_t));
    uint32_t *vln = (uint32_t *)zalloc(nv, sizeof(uint32_t));
    uint32_t *vrow = (uint32_t *)zalloc(nv, sizeof(uint32_t));   /* row of each virtual row */
    B->cidx = (uint32_t *)zalloc(nv, sizeof(uint32_t));
    B->lpos = (uint16_t *)zalloc(nv, sizeof(uint16_t));
    B->voff = (uint32_t *)zalloc((size_t)nv + 1, sizeof(uint32_t));
    B->vsrc = (uint32_t *)zalloc(mb, sizeof(uint32_t));
    if (p->w)
        B->vw = (float *)zalloc(mb, sizeof(float));
    const int nt = omp_get_max_threads() > 0 ? omp_get_max_threads() : 1;
    uint32_t *hist = (uint32_t *)zalloc((size_t)nt * nK, sizeof(uint32_t));
    uint32_t *kstart = (uint32_t *)zalloc((size_t)nK + 1, sizeof(uint32_t));
    uint32_t *vlen = (uint32_t *)zalloc(nv, sizeof(uint32_t));
    uint32_t *vfrom = (uint32_t *)zalloc(nv, sizeof(uint32_t));
    if (!vkey || !vst || !vln || !vrow || !B->cidx || !B->lpos || !B->voff || !B->vsrc ||
        (p->w && !B->vw) || !hist || !kstart || !vlen || !vfrom) {
        oom = 1;
        goto done;
    }
    #pragma omp parallel for schedule(dynamic, 256)
    for (uint32_t r = row0; r < n; r++) {
        const uint32_t d = p->off[r + 1] - p->off[r], o = p->off[r] - e0, c = B->coff[r];
        uint32_t x = 0;
        while (x + 1 < AG_NXCD && r >= B->slice_row[x + 1])
            x++;
        uint32_t cnt = row_pieces(ss + o, d, &K, vst + c, vln + c, vkey + c);
        for (uint32_t j = 0; j < cnt; j++) {
            vkey[c + j] = x * nb + vkey[c + j];
            vst[c + j] += o;
        }
    }
    /* 3: stable counting sort of the slots by (slice, block): slot -> virtual row */
    #pragma omp parallel num_threads(nt)
    {
        const int t = omp_get_thread_num(), T = omp_get_num_threads();
        const uint32_t a = (uint32_t)((uint64_t)nv * t / T), b = (uint32_t)((uint64_t)nv * (t + 1) / T);
        uint32_t *h = hist + (size_t)t * nK;
        for (uint32_t i = a; i < b; i++)
            h[vkey[i]]++;
        #pragma omp barrier
        #pragma omp single
        {
            uint32_t run = 0;
            for (uint32_t K = 0; K < nK; K++) {
                kstart[K] = run;
                for (int u = 0; u < T; u++) {
                    uint32_t c = hist[(size_t)u * nK + K];
                    hist[(size_t)u * nK + K] = run;
                    run += c;
                }
            }
            kstart[nK] = run;
        }
        /* rows of this thread's slots: the slots [a, b) belong to consecutive rows */
        uint32_t r = row0;
        {
            uint32_t lo = row0, hi = n;   /* first row with coff[r + 1] > a */
            while (lo < hi) {
                uint32_t mid = lo + (hi - lo) / 2;
                if (B->coff[mid + 1] <= a)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            r = lo;
        }
        for (uint32_t i = a; i < b; i++) {
            while (B->coff[r + 1] <= i)
                r++;
            uint32_t v = h[vkey[i]]++;
            B->cidx[i] = v;
            vlen[v] = vln[i];
            vfrom[v] = vst[i];
            vrow[v] = r;
        }
    }
    {
        uint64_t acc = 0;
        for (uint32_t v = 0; v < nv; v++) {
            B->voff[v] = (uint32_t)acc;
            acc += vlen[v];
        }
        B->voff[nv] = (uint32_t)acc;
    }
    #pragma omp parallel for schedule(dynamic, 4096)
    for (uint32_t v = 0; v < nv; v++) {
        memcpy(B->vsrc + B->voff[v], ss + vfrom[v], (size_t)vlen[v] * sizeof(uint32_t));
        if (B->vw)
            memcpy(B->vw + B->voff[v], sw + vfrom[v], (size_t)vlen[v] * sizeof(float));
    }
    {
        uint32_t rb[AG_NBINS];
        for (int k = 0; k < AG_NBINS; k++)
            rb[k] = ((uint32_t)k < min_bin) ? p->bin_blocks[k] : 0u;
        B->r_slot_blk0[0] = 0;
        for (int sl = 0; sl < AG_NSLOTS; sl++)
            B->r_slot_blk0[sl + 1] = B->r_slot_blk0[sl] + rb[ag_slot_bin[sl]];
    }

    /* 4: pass-1 items, slice by slice, block-major (hot segments first), and the groups of
     * hot-segment tiles */
    {
        uint32_t ng = 0;
        B->n_items = emit_items(p, kstart, vlen, B->r_slot_blk0[AG_NSLOTS], hot_nnz, NULL, NULL,
                                NULL, &ng);
        B->item = (uint32_t *)zalloc((size_t)B->n_items * 2, sizeof(uint32_t));
        B->hgrp = (uint32_t *)zalloc((size_t)(ng ? ng : 1) * 3, sizeof(uint32_t));
        if (!B->item || !B->hgrp) {
            oom = 1;
            goto done;
        }
        emit_items(p, kstart, vlen, B->r_slot_blk0[AG_NSLOTS], hot_nnz, B->item, B->item_off,
                   B->hgrp, &B->n_hgrp);
    }

    /* 5: combine tiles: per block, the tile's rows' virtual rows are one contiguous run of
     * the (slice, block) list; the tile streams its partials block by block, and the k-th of
     * them goes to row-major slot lpos[p0 + k] of the tile (p0 = its first slot) */
    {
        uint32_t tiles = 0;
        B->c_n_items = emit_combine(p, NULL, &tiles, NULL, NULL);
        B->c_n_tiles = tiles;
        B->ct_row = (uint32_t *)zalloc((size_t)tiles * 2, sizeof(uint32_t));
        B->c_item = (uint32_t *)zalloc((size_t)B->c_n_items * 2, sizeof(uint32_t));
        B->ct_start = (uint32_t *)zalloc((size_t)tiles * nb, sizeof(uint32_t));
        B->ct_pre = (uint32_t *)zalloc((size_t)tiles * (nb + 1), sizeof(uint32_t));
        if (!B->ct_row || !B->c_item || !B->ct_start || !B->ct_pre) {
            oom = 1;
            goto done;
        }
        emit_combine(p, B->ct_row, &tiles, B->c_item, B->c_item_off);
        /* row-major slot of each virtual row inside its tile (each tile fills its own) */
        uint16_t *vslot = (uint16_t *)zalloc(nv, sizeof(uint16_t));
        if (!vslot) {
            oom = 1;
            goto done;
        }
        #pragma omp parallel for schedule(dynamic, 64)
        for (uint32_t t = 0; t < tiles; t++) {
            const uint32_t r0 = B->ct_row[2 * t], r1 = B->ct_row[2 * t + 1];
            uint32_t x = 0;
            while (x + 1 < AG_NXCD && r0 >= B->slice_row[x + 1])
                x++;
            uint32_t run = 0;
            for (uint32_t b = 0; b < nb; b++) {
                const uint32_t lo = kstart[x * nb + b], hi = kstart[x * nb + b + 1];
                const uint32_t s = lower_bound_u32(vrow, lo, hi, r0);
                const uint32_t f = lower_bound_u32(vrow, s, hi, r1);
                B->ct_start[(size_t)t * nb + b] = s;
                B->ct_pre[(size_t)t * (nb + 1) + b] = run;
                run += f - s;
            }
            B->ct_pre[(size_t)t * (nb + 1) + nb] = run;
            const uint32_t p0 = B->coff[r0];
            for (uint32_t r = r0; r < r1; r++)
                for (uint32_t j = B->coff[r]; j < B->coff[r + 1]; j++)
                    vslot[B->cidx[j]] = (uint16_t)(j - p0);
            for (uint32_t b = 0; b < nb; b++) {
                const uint32_t s = B->ct_start[(size_t)t * nb + b];
                const uint32_t k0 = B->ct_pre[(size_t)t * (nb + 1) + b];
                const uint32_t k1 = B->ct_pre[(size_t)t * (nb + 1) + b + 1];
                for (uint32_t k = k0; k < k1; k++)
                    B->lpos[p0 + k] = vslot[s + (k - k0)];
            }
        }
        free(vslot);
    }

    /* 6: tile_sort (not with hot segments, whose pass reads vw in virtual-row order) */
    if (B->bits <= 20 && !H && q->tile_sort != AG_SORT_OFF && sort_tile_ids(B) != AG_OK) {
        oom = 1;
        goto done;
    }

done:
    free(kmap);
    free(ss);
    free(sw);
    free(vkey);
    free(vst);
    free(vln);
    free(vrow);
    free(hist);
    free(kstart);
    free(vlen);
    free(vfrom);
    if (oom) {
        *err = "host allocation failed";
        return AG_ERR_OOM;
    }
    return AG_OK;
}

int ag_plan_build(uint32_t n, uint32_t m, const uint32_t *out_degree,
                  const uint32_t *inv_edges_idx, const uint32_t *inv_degree,
                  const uint32_t *inv_src, const float *inv_weight,
                  const ag_plan_params *params, ag_plan *p, const char **err)
{
    memset(p, 0, sizeof(*p));
    if (params)
        p->params = *params;
    else
        ag_plan_params_default(&p->params);
    int st = ag_plan_params_check(&p->params, err);
    if (st != AG_OK)
        return st;
    if (p->params.fp32_order == AG_ORDER_REFERENCE) {
        /* sequential sums only: rows that fit a tile are tile rows (one lane, from LDS), every
         * longer row is one hub segment (one lane over chunks staged in LDS); no blocking */
        ag_plan_params *q = &p->params;
        q->thread_max_deg = q->group16_max_deg = q->wave_max_deg = q->block_max_deg = q->tile_nnz;
        q->src_block_bits = AG_SRC_BLOCK_OFF;
    }
    /* keeps n + 1, word-rounded mark counts and grid sizes inside 32 bits */
    if (n > 0xFFFF0000u) {
        *err = "num_vertices exceeds 2^32 - 2^16";
        return AG_ERR_INVALID;
    }
    if (n > 0 && (!out_degree || !inv_edges_idx || !inv_degree)) {
        *err = "null degree/offset array";
        return AG_ERR_INVALID;
    }
    if (m > 0 && !inv_src) {
        *err = "null inverse source array";
        return AG_ERR_INVALID;
    }
    /* validate row ranges and the total number of in-edges the rows reference */
    uint64_t total = 0;
    int bad_range = 0;
    #pragma omp parallel for reduction(+ : total) reduction(| : bad_range)
    for (uint32_t v = 0; v < n; v++) {
        uint64_t e = (uint64_t)inv_edges_idx[v] + inv_degree[v];
        if (e > m)
            bad_range |= 1;
        total += inv_degree[v];
    }
    if (bad_range) {
        *err = "inverse row extends past num_edges";
        return AG_ERR_INVALID;
    }
    if (total > 0xFFFF0000ull) {
        *err = "total in-degree exceeds 2^32 - 2^16";
        return AG_ERR_INVALID;
    }
    int bad_src = 0;
    #pragma omp parallel for schedule(dynamic, 1024) reduction(| : bad_src)
    for (uint32_t v = 0; v < n; v++) {
        uint32_t b = inv_edges_idx[v], e = b + inv_degree[v];
        for (uint32_t j = b; j < e; j++)
            if (inv_src[j] >= n)
                bad_src |= 1;
    }
    if (bad_src) {
        *err = "inverse source id >= num_vertices";
        return AG_ERR_INVALID;
    }

    p->n = n;
    p->m = (uint32_t)total;
    p->d2o = (uint32_t *)zalloc(n, sizeof(uint32_t));
    p->o2d = (uint32_t *)zalloc(n, sizeof(uint32_t));
    p->off = (uint32_t *)zalloc((size_t)n + 1, sizeof(uint32_t));
    p->src = (uint32_t *)zalloc(p->m, sizeof(uint32_t));
    p->outdeg = (uint32_t *)zalloc(n, sizeof(uint32_t));
    uint8_t *bin = (uint8_t *)zalloc(n, 1);
    if (inv_weight)
        p->w = (float *)zalloc(p->m, sizeof(float));
    if (!p->d2o || !p->o2d || !p->off || !p->src || !p->outdeg || !bin ||
        (inv_weight && !p->w)) {
        free(bin);
        ag_plan_free(p);
        *err = "host allocation failed";
        return AG_ERR_OOM;
    }

    /* bin every row; count rows and edges per bin */
    uint32_t cnt[AG_NBINS] = {0};
    for (uint32_t v = 0; v < n; v++) {
        int b = bin_of(&p->params, inv_degree[v]);
        bin[v] = (uint8_t)b;
        cnt[b]++;
        p->bin_edges[b] += inv_degree[v];
    }
    p->bin_row[0] = 0;
    for (int b = 0; b < AG_NBINS; b++)
        p->bin_row[b + 1] = p->bin_row[b] + cnt[b];

    /* device order: bin-major, then vertex id (order 0) or out-degree descending (order 1) */
    {
        uint32_t cur[AG_NBINS];
        for (int b = 0; b < AG_NBINS; b++)
            cur[b] = p->bin_row[b];
        for (uint32_t v = 0; v < n; v++)
            p->d2o[cur[bin[v]]++] = v;
        if (p->params.order == 1) {
            for (int b = 0; b < AG_NBINS; b++) {
                uint32_t lo = p->bin_row[b], len = p->bin_row[b + 1] - lo;
                uint64_t *keys = (uint64_t *)zalloc(len, sizeof(uint64_t));
                if (!keys) {
                    free(bin);
                    ag_plan_free(p);
                    *err = "host allocation failed";
                    return AG_ERR_OOM;
                }
                for (uint32_t i = 0; i < len; i++) {
                    uint32_t v = p->d2o[lo + i];
                    keys[i] = ((uint64_t)(0xFFFFFFFFu - out_degree[v]) << 32) | v;
                }
                qsort(keys, len, sizeof(uint64_t), cmp_u64);
                for (uint32_t i = 0; i < len; i++)
                    p->d2o[lo + i] = (uint32_t)keys[i];
                free(keys);
            }
        }
    }
    free(bin);
    #pragma omp parallel for
    for (uint32_t r = 0; r < n; r++) {
        p->o2d[p->d2o[r]] = r;
        p->outdeg[r] = out_degree[p->d2o[r]];
    }
    {
        uint64_t acc = 0;
        for (uint32_t r = 0; r < n; r++) {
            p->off[r] = (uint32_t)acc;
            acc += inv_degree[p->d2o[r]];
        }
        p->off[n] = (uint32_t)acc;
    }
    #pragma omp parallel for schedule(dynamic, 1024)
    for (uint32_t r = 0; r < n; r++) {
        uint32_t v = p->d2o[r];
        uint32_t b = inv_edges_idx[v], deg = inv_degree[v], o = p->off[r];
        for (uint32_t k = 0; k < deg; k++) {
            p->src[o + k] = p->o2d[inv_src[b + k]];
            if (p->w)
                p->w[o + k] = inv_weight[b + k];
        }
    }

    /* thread-bin tiles: greedy cut on rows and in-edges */
    p->tile_row = cut_tiles(p->off, p->bin_row[AG_BIN_THREAD], p->bin_row[AG_BIN_THREAD + 1],
                            p->params.tile_rows, p->params.tile_nnz, &p->n_tiles);
    if (!p->tile_row) {
        ag_plan_free(p);
        *err = "host allocation failed";
        return AG_ERR_OOM;
    }

    /* hub segments */
    {
        uint32_t lo = p->bin_row[AG_BIN_HUB], hi = p->bin_row[AG_BIN_HUB + 1];
        /* reference order: one segment per hub row, added in sequence */
        uint32_t seg = p->params.fp32_order == AG_ORDER_REFERENCE ? UINT32_MAX
                                                                  : p->params.segment_nnz;
        uint64_t nsegs = 0;
        for (uint32_t r = lo; r < hi; r++)
            nsegs += ((uint64_t)(p->off[r + 1] - p->off[r]) + seg - 1) / seg;
        p->n_hubs = hi - lo;
        p->n_segs = (uint32_t)nsegs;
        p->seg_row = (uint32_t *)zalloc(p->n_segs, sizeof(uint32_t));
        p->seg_lo = (uint32_t *)zalloc(p->n_segs, sizeof(uint32_t));
        p->seg_hi = (uint32_t *)zalloc(p->n_segs, sizeof(uint32_t));
        p->hub_seg0 = (uint32_t *)zalloc((size_t)p->n_hubs + 1, sizeof(uint32_t));
        if (!p->seg_row || !p->seg_lo || !p->seg_hi || !p->hub_seg0) {
            ag_plan_free(p);
            *err = "host allocation failed";
            return AG_ERR_OOM;
        }
        uint32_t s = 0;
        for (uint32_t r = lo; r < hi; r++) {
            p->hub_seg0[r - lo] = s;
            for (uint32_t a = p->off[r]; a < p->off[r + 1];) {
                p->seg_row[s] = r;
                p->seg_lo[s] = a;
                p->seg_hi[s] = (p->off[r + 1] - a > seg) ? a + seg : p->off[r + 1];
                a = p->seg_hi[s];
                s++;
            }
        }
        p->hub_seg0[p->n_hubs] = s;
    }

    /* blocks per bin and launch slots */
    p->bin_blocks[AG_BIN_THREAD] = p->n_tiles;
    p->bin_blocks[AG_BIN_GROUP16] = (cnt[AG_BIN_GROUP16] + 15) / 16;
    p->bin_blocks[AG_BIN_WAVE] = (cnt[AG_BIN_WAVE] + 3) / 4;
    p->bin_blocks[AG_BIN_BLOCK] = cnt[AG_BIN_BLOCK];
    p->bin_blocks[AG_BIN_HUB] = p->n_segs;
    p->slot_blk0[0] = 0;
    for (int s = 0; s < AG_NSLOTS; s++)
        p->slot_blk0[s + 1] = p->slot_blk0[s] + p->bin_blocks[ag_slot_bin[s]];
    p->total_blocks = p->slot_blk0[AG_NSLOTS];

    st = build_block_plan(p, err);
    if (st != AG_OK) {
        ag_plan_free(p);
        return st;
    }
    return AG_OK;
}
