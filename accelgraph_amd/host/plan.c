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
    p->cold_below_bin = 0;
    p->wg_per_cu = 0;
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
    if (p->cold_below_bin > AG_NBINS) {
        *err = "cold_below_bin must be in [0, 5]";
        return AG_ERR_INVALID;
    }
    if (p->order > 1) {
        *err = "order must be 0 (vertex id) or 1 (out-degree descending)";
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
    {
        uint32_t lo = p->bin_row[AG_BIN_THREAD], hi = p->bin_row[AG_BIN_THREAD + 1];
        uint32_t cap = hi - lo + 1;
        p->tile_row = (uint32_t *)zalloc((size_t)cap + 1, sizeof(uint32_t));
        if (!p->tile_row) {
            ag_plan_free(p);
            *err = "host allocation failed";
            return AG_ERR_OOM;
        }
        uint32_t t = 0, r = lo;
        while (r < hi) {
            p->tile_row[t++] = r;
            uint32_t rows = 0, nnz = 0;
            while (r < hi && rows < p->params.tile_rows) {
                uint32_t d = p->off[r + 1] - p->off[r];
                if (rows > 0 && nnz + d > p->params.tile_nnz)
                    break;
                nnz += d;
                rows++;
                r++;
            }
        }
        p->tile_row[t] = hi;
        p->n_tiles = t;
    }

    /* hub segments */
    {
        uint32_t lo = p->bin_row[AG_BIN_HUB], hi = p->bin_row[AG_BIN_HUB + 1];
        uint32_t seg = p->params.segment_nnz;
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
            for (uint32_t a = p->off[r]; a < p->off[r + 1]; a += seg) {
                p->seg_row[s] = r;
                p->seg_lo[s] = a;
                p->seg_hi[s] = (p->off[r + 1] - a > seg) ? a + seg : p->off[r + 1];
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
    return AG_OK;
}
