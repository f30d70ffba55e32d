// ag_runtime.hip -- device contexts, error reporting, the bounded wait (watchdog) and graph
// upload of the accelgraph C-ABI.
//
// Replaces the reference's CAPI runtime boundary (02_capi_graph/src/capi_utils/capienv.c:
// setupAFUGraphCSR 329-372, startAFU 375-451, waitAFU 526-638, releaseAFU 765-782,
// mapGraphCSRToWED 832-874). The FPGA read host memory coherently through a work-element
// descriptor; here the arrays the pull path reads are uploaded into HBM once, in the
// degree-binned layout built by ag_plan_build, and every later run reuses them.
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "ag_internal.h"
#include "ag_fixed_dev.h"

static thread_local char g_err[512] = "";

int ag_set_error(int status, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return status;
}

int ag_hip_fail(hipError_t e, const char *what, const char *file, int line)
{
    int st = (e == hipErrorOutOfMemory) ? AG_ERR_OOM
             : (e == hipErrorNoDevice || e == hipErrorInvalidDevice) ? AG_ERR_NO_DEVICE
                                                                    : AG_ERR_HIP;
    return ag_set_error(st, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
}

extern "C" const char *ag_last_error(void) { return g_err; }

extern "C" int ag_abi_version(void) { return AG_ABI_VERSION; }

extern "C" const char *ag_status_string(int s)
{
    switch (s) {
    case AG_OK: return "ok";
    case AG_ERR_INVALID: return "invalid argument";
    case AG_ERR_NO_DEVICE: return "no usable device";
    case AG_ERR_HIP: return "HIP runtime error";
    case AG_ERR_OOM: return "out of memory";
    case AG_ERR_TIMEOUT: return "device wait timed out";
    case AG_ERR_UNSUPPORTED: return "unsupported option";
    default: return "unknown status";
    }
}

static uint64_t env_u64(const char *name, uint64_t dflt)
{
    const char *v = getenv(name);
    if (!v || !*v)
        return dflt;
    char *end = NULL;
    unsigned long long x = strtoull(v, &end, 10);
    return (end && *end == '\0') ? (uint64_t)x : dflt;
}

extern "C" int ag_device_count(int *count)
{
    if (!count)
        return ag_set_error(AG_ERR_INVALID, "count is null");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
        *count = 0;
        return ag_hip_fail(e, "hipGetDeviceCount", __FILE__, __LINE__);
    }
    *count = c;
    return AG_OK;
}

// Scratch of the allocation-free entry point ag_spmv_pull_noalloc (permuted x and y), a
// file-scope device array: every GPU the library runs on holds its own copy at its own
// address until the process ends. Host code never uses the array's name as an address: the
// address is taken once per device with hipGetSymbolAddress (that device current) when the
// device is first opened, and kept per ordinal.
static __device__ __align__(256) unsigned char scratch[AG_NOALLOC_SCRATCH_BYTES];
#define AG_MAX_ORDINALS 64
static void *g_scratch_ptr[AG_MAX_ORDINALS];

extern "C" int ag_device_open(int ordinal, ag_device **out)
{
    if (!out)
        return ag_set_error(AG_ERR_INVALID, "out is null");
    *out = NULL;
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess || c == 0)
        return ag_set_error(AG_ERR_NO_DEVICE, "no HIP device available (%s)",
                            e == hipSuccess ? "count is 0" : hipGetErrorString(e));
    if (ordinal < 0 || ordinal >= c)
        return ag_set_error(AG_ERR_INVALID, "device ordinal %d out of range [0,%d)", ordinal, c);
    AG_HIP(hipSetDevice(ordinal));
    ag_device *d = (ag_device *)calloc(1, sizeof(ag_device));
    if (!d)
        return ag_set_error(AG_ERR_OOM, "host allocation failed");
    d->ordinal = ordinal;
    d->run_timeout_ms = env_u64("ACCELERATOR_RUN_TIMEOUT_MS", 0);
    d->poll_us = (uint32_t)env_u64("ACCELERATOR_POLL_INTERVAL_US", 50);
    int cus = 0;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ordinal);
    if (e != hipSuccess || cus <= 0) {
        free(d);
        return ag_hip_fail(e, "hipDeviceGetAttribute(CU count)", __FILE__, __LINE__);
    }
    d->num_cus = (uint32_t)cus;
    if (ordinal < AG_MAX_ORDINALS && !g_scratch_ptr[ordinal]) {
        void *p = NULL;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            free(d);
            return ag_hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
        }
        g_scratch_ptr[ordinal] = p;
    }
    d->scratch = ordinal < AG_MAX_ORDINALS ? g_scratch_ptr[ordinal] : NULL;
    e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        free(d);
        return ag_hip_fail(e, "hipStreamCreateWithFlags", __FILE__, __LINE__);
    }
    *out = d;
    return AG_OK;
}

extern "C" int ag_device_close(ag_device *d)
{
    if (!d)
        return AG_OK;
    if (d->live_graphs)
        return ag_set_error(AG_ERR_INVALID, "%u graph(s) still uploaded on this device; free "
                            "them before closing it", d->live_graphs);
    (void)hipSetDevice(d->ordinal);
    (void)hipStreamSynchronize(d->stream);
    (void)hipStreamDestroy(d->stream);
    free(d);
    return AG_OK;
}

extern "C" int ag_device_set_watchdog(ag_device *d, uint64_t run_timeout_ms, uint32_t poll_us)
{
    if (!d)
        return ag_set_error(AG_ERR_INVALID, "device is null");
    d->run_timeout_ms = run_timeout_ms;
    d->poll_us = poll_us ? poll_us : 1;
    return AG_OK;
}

static uint64_t now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (uint64_t)ts.tv_sec * 1000ull + (uint64_t)ts.tv_nsec / 1000000ull;
}

// Wait for all work on the device stream; bounded by the watchdog deadline when one is set
// (the role of waitAFU's run deadline, capienv.c:526-638).
int ag_wait_stream(ag_device *d)
{
    if (d->run_timeout_ms == 0) {
        AG_HIP(hipStreamSynchronize(d->stream));
        return AG_OK;
    }
    hipEvent_t ev;
    AG_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    hipError_t e = hipEventRecord(ev, d->stream);
    if (e != hipSuccess) {
        (void)hipEventDestroy(ev);
        return ag_hip_fail(e, "hipEventRecord", __FILE__, __LINE__);
    }
    uint64_t deadline = now_ms() + d->run_timeout_ms;
    for (;;) {
        e = hipEventQuery(ev);
        if (e == hipSuccess)
            break;
        if (e != hipErrorNotReady) {
            (void)hipEventDestroy(ev);
            return ag_hip_fail(e, "hipEventQuery", __FILE__, __LINE__);
        }
        if (now_ms() >= deadline) {
            (void)hipEventDestroy(ev);
            return ag_set_error(AG_ERR_TIMEOUT, "device work did not finish within %llu ms",
                                (unsigned long long)d->run_timeout_ms);
        }
        usleep(d->poll_us);
    }
    (void)hipEventDestroy(ev);
    return AG_OK;
}

extern "C" int ag_device_sync(ag_device *d)
{
    if (!d)
        return ag_set_error(AG_ERR_INVALID, "device is null");
    AG_HIP(hipSetDevice(d->ordinal));
    return ag_wait_stream(d);
}

// ------------------------------------------------------------------ graph upload

template <typename T>
static int dev_upload(ag_dgraph *g, T **dst, const T *src, size_t count)
{
    size_t bytes = (count ? count : 1) * sizeof(T);
    AG_HIP(hipMalloc((void **)dst, bytes));
    g->device_bytes += bytes;
    if (count)
        AG_HIP(hipMemcpyAsync(*dst, src, count * sizeof(T), hipMemcpyHostToDevice, g->dev->stream));
    return AG_OK;
}

extern "C" int ag_graph_free(ag_dgraph *g)
{
    if (!g)
        return AG_OK;
    (void)hipSetDevice(g->dev->ordinal);
    (void)hipStreamSynchronize(g->dev->stream);
    ag_ppr_free_ws(g);
    void *ptrs[] = {g->d2o, g->off, g->src, g->w, g->wq, g->outdeg, g->tile_row, g->seg_row,
                    g->seg_lo, g->seg_hi, g->hub_seg0, g->hub_part, g->vec_a, g->vec_b,
                    g->blk.voff, g->blk.vsrc, g->blk.vw, g->blk.item, g->blk.coff, g->blk.cidx,
                    g->blk.lpos, g->blk.ct_row, g->blk.ct_start, g->blk.ct_pre, g->blk.c_item,
                    g->blk.hgrp, g->blk.vsp, g->blk.tbase,
                    g->vpart, g->btick, g->vwq, g->fwd_off, g->fwd_src, g->fwd_dst,
                    g->fwd_w, g->pb_row, g->pb_pos, g->pb_dst};
    for (void *p : ptrs)
        if (p)
            (void)hipFree(p);
    free(g->h_d2o);
    if (g->dev->live_graphs)
        g->dev->live_graphs--;
    free(g);
    return AG_OK;
}

PullGeom ag_make_geom(const ag_dgraph *g)
{
    PullGeom G;
    G.off = g->off;
    G.src = g->src;
    G.tile_row = g->tile_row;
    G.seg_row = g->seg_row;
    G.seg_lo = g->seg_lo;
    G.seg_hi = g->seg_hi;
    G.hub_seg0 = g->hub_seg0;
    G.hub_part = g->hub_part;
    G.n = g->n;
    G.n_hubs = g->n_hubs;
    G.seq_hubs = g->params.fp32_order == AG_ORDER_REFERENCE ? 1u : 0u;
    for (int b = 0; b <= AG_NBINS; b++)
        G.bin_row[b] = g->bin_row[b];
    for (int s = 0; s <= AG_NSLOTS; s++)
        G.slot_blk0[s] = g->slot_blk0[s];
    return G;
}

PullGeom ag_make_unblocked_geom(const ag_dgraph *g)
{
    PullGeom G = ag_make_geom(g);
    for (int s = 0; s <= AG_NSLOTS; s++)
        G.slot_blk0[s] = g->blk.r_slot_blk0[s];
    if (g->params.blocked_min_bin <= AG_BIN_HUB)
        G.n_hubs = 0;   // hub rows are blocked: their sums come from the combine pass
    return G;
}

BlockGeom ag_make_block_geom(const ag_dgraph *g)
{
    const ag_block_plan &B = g->blk;
    BlockGeom K;
    memset(&K, 0, sizeof(K));
    K.V.off = B.voff;
    K.V.src = B.vsrc;
    K.V.n = B.nv;
    K.C.off = B.coff;
    K.C.src = B.cidx;
    K.C.n = g->n;
    K.item = B.item;
    K.c_item = B.c_item;
    K.ct_row = B.ct_row;
    K.ct_start = B.ct_start;
    K.ct_pre = B.ct_pre;
    K.lpos = B.lpos;
    K.part = g->vpart;
    K.nblocks = B.nblocks;
    for (uint32_t x = 0; x <= AG_NXCD; x++) {
        K.item_off[x] = B.item_off[x];
        K.c_item_off[x] = B.c_item_off[x];
        K.hgrp_off[x] = B.hgrp_off[x];
    }
    for (uint32_t x = 0; x < AG_NXCD; x++)
        K.item_lo[x] = B.item_off[x];
    K.hgrp = B.hgrp;
    K.hot_bits = B.hot_bits;
    K.n_src = g->n;
    K.dyn = g->bdyn;
    K.tick = g->btick;
    K.top0 = B.bits ? B.top0 : g->n;
    K.vsp = B.vsp;
    K.tbase = B.tbase;
    K.sbits = B.bits && B.sorted ? B.bits : 0u;
    return K;
}

void ag_block_geom_skip_hot(const ag_dgraph *g, BlockGeom &K)
{
    for (uint32_t x = 0; x < AG_NXCD; x++)
        K.item_lo[x] = g->blk.hot_end[x];
}

// Resident four-wave workgroups per CU: the plan's wg_per_cu, by default 4 (16 waves per
// CU), never more than 160 KB of LDS allows. The gathers are bound by the rate of random L2
// requests, and more requests in flight only lengthen the queues (8 per CU measured 5-40%
// slower than 4 on RMAT-25 and ER-25).
static uint32_t wg_per_cu(const ag_dgraph *g, uint32_t lds_bytes, uint32_t dflt = 4u)
{
    uint32_t by_lds = lds_bytes ? (160u * 1024u) / lds_bytes : 32u;
    uint32_t per_cu = g->params.wg_per_cu ? g->params.wg_per_cu : dflt;
    if (per_cu > by_lds)
        per_cu = by_lds;
    return per_cu ? per_cu : 1u;
}

uint32_t ag_blocked_grid(const ag_dgraph *g, uint32_t resident_per_cu, uint32_t default_per_cu)
{
    // workgroup i runs on XCD i % AG_NXCD (round-robin dispatch), so slice x is served by the
    // workgroups with that residue; the grid is a multiple of AG_NXCD
    uint32_t per_cu = g->params.wg_per_cu ? g->params.wg_per_cu : default_per_cu;
    if (per_cu > resident_per_cu)
        per_cu = resident_per_cu;
    uint32_t grid = g->dev->num_cus * (per_cu ? per_cu : 1u) / AG_NXCD * AG_NXCD;
    return grid ? grid : AG_NXCD;
}

uint32_t ag_pull_grid(const ag_dgraph *g, uint32_t lds_bytes, uint32_t default_per_cu)
{
    uint32_t grid = g->dev->num_cus * wg_per_cu(g, lds_bytes, default_per_cu);
    return grid < g->total_blocks ? grid : g->total_blocks;
}

// vec_a / vec_b: two scratch arrays of `bytes` each, used to permute SpMV/gather vectors
// between original and device order.
int ag_ensure_vec(ag_dgraph *g, size_t bytes)
{
    if (g->vec_bytes >= bytes)
        return AG_OK;
    if (g->vec_a)
        AG_HIP(hipFree(g->vec_a));
    if (g->vec_b)
        AG_HIP(hipFree(g->vec_b));
    g->vec_a = g->vec_b = NULL;
    g->vec_bytes = 0;
    AG_HIP(hipMalloc(&g->vec_a, bytes));
    AG_HIP(hipMalloc(&g->vec_b, bytes));
    g->vec_bytes = bytes;
    return AG_OK;
}

extern "C" int ag_graph_upload(ag_device *dev, uint32_t n, uint32_t m, const uint32_t *out_degree,
                               const uint32_t *inv_edges_idx, const uint32_t *inv_degree,
                               const uint32_t *inv_src, const float *inv_weight,
                               const ag_plan_params *params, ag_dgraph **out)
{
    if (!dev || !out)
        return ag_set_error(AG_ERR_INVALID, "device or out is null");
    *out = NULL;
    ag_plan plan;
    const char *msg = "";
    int st = ag_plan_build(n, m, out_degree, inv_edges_idx, inv_degree, inv_src, inv_weight,
                           params, &plan, &msg);
    if (st != AG_OK)
        return ag_set_error(st, "graph upload rejected: %s", msg);
    ag_dgraph *g = (ag_dgraph *)calloc(1, sizeof(ag_dgraph));
    if (!g) {
        ag_plan_free(&plan);
        return ag_set_error(AG_ERR_OOM, "host allocation failed");
    }
    g->dev = dev;
    dev->live_graphs++;
    g->params = plan.params;
    g->n = plan.n;
    g->m = plan.m;
    memcpy(g->bin_row, plan.bin_row, sizeof(g->bin_row));
    memcpy(g->bin_edges, plan.bin_edges, sizeof(g->bin_edges));
    memcpy(g->bin_blocks, plan.bin_blocks, sizeof(g->bin_blocks));
    memcpy(g->slot_blk0, plan.slot_blk0, sizeof(g->slot_blk0));
    g->total_blocks = plan.total_blocks;
    g->n_tiles = plan.n_tiles;
    g->n_hubs = plan.n_hubs;
    g->n_segs = plan.n_segs;

    hipError_t he = hipSetDevice(dev->ordinal);
    if (he != hipSuccess) {
        ag_plan_free(&plan);
        free(g);
        return ag_hip_fail(he, "hipSetDevice", __FILE__, __LINE__);
    }
    st = AG_OK;
    if (st == AG_OK) {
        g->h_d2o = (uint32_t *)malloc((size_t)(plan.n ? plan.n : 1) * sizeof(uint32_t));
        if (!g->h_d2o)
            st = ag_set_error(AG_ERR_OOM, "host allocation failed");
        else
            memcpy(g->h_d2o, plan.d2o, (size_t)plan.n * sizeof(uint32_t));
    }
    if (st == AG_OK) st = dev_upload(g, &g->d2o, plan.d2o, plan.n);
    if (st == AG_OK) st = dev_upload(g, &g->off, plan.off, (size_t)plan.n + 1);
    if (st == AG_OK) st = dev_upload(g, &g->src, plan.src, plan.m);
    if (st == AG_OK && plan.w) st = dev_upload(g, &g->w, plan.w, plan.m);
    if (st == AG_OK) st = dev_upload(g, &g->outdeg, plan.outdeg, plan.n);
    if (st == AG_OK) st = dev_upload(g, &g->tile_row, plan.tile_row, (size_t)plan.n_tiles + 1);
    if (st == AG_OK) st = dev_upload(g, &g->seg_row, plan.seg_row, plan.n_segs);
    if (st == AG_OK) st = dev_upload(g, &g->seg_lo, plan.seg_lo, plan.n_segs);
    if (st == AG_OK) st = dev_upload(g, &g->seg_hi, plan.seg_hi, plan.n_segs);
    if (st == AG_OK) st = dev_upload(g, &g->hub_seg0, plan.hub_seg0, (size_t)plan.n_hubs + 1);
    if (st == AG_OK) {
        size_t pb = (size_t)(plan.n_segs ? plan.n_segs : 1) * sizeof(uint64_t);
        he = hipMalloc((void **)&g->hub_part, pb);
        if (he != hipSuccess)
            st = ag_hip_fail(he, "hub scratch allocation", __FILE__, __LINE__);
        g->device_bytes += pb;
    }
    if (st == AG_OK && plan.blk.bits) {
        const ag_block_plan &B = plan.blk;
        g->blk = B;   // scalars; the pointers are replaced by device copies below
        g->blk.voff = g->blk.vsrc = g->blk.item = g->blk.coff = g->blk.cidx = NULL;
        g->blk.ct_row = g->blk.ct_start = g->blk.ct_pre = g->blk.c_item = NULL;
        g->blk.lpos = NULL;
        g->blk.vw = NULL;
        g->blk.hgrp = NULL;
        g->blk.vsp = g->blk.tbase = NULL;
        g->blk.hot_seg = NULL;   // host-only array of the plan, freed with it
        const size_t nt = B.c_n_tiles, nb = B.nblocks;
        st = dev_upload(g, &g->blk.voff, B.voff, (size_t)B.nv + 1);
        // with sorted tiles pass 1 reads only the packed ids (vsp): the ids in virtual-row
        // order stay on the host (4 bytes per blocked in-edge less in HBM)
        if (st == AG_OK && !B.sorted) st = dev_upload(g, &g->blk.vsrc, B.vsrc, B.mb);
        if (st == AG_OK && B.vw) st = dev_upload(g, &g->blk.vw, B.vw, B.mb);
        if (st == AG_OK) st = dev_upload(g, &g->blk.item, B.item, (size_t)B.n_items * 2);
        if (st == AG_OK) st = dev_upload(g, &g->blk.coff, B.coff, (size_t)plan.n + 1);
        if (st == AG_OK) st = dev_upload(g, &g->blk.cidx, B.cidx, B.nv);
        if (st == AG_OK) st = dev_upload(g, &g->blk.lpos, B.lpos, B.nv);
        if (st == AG_OK) st = dev_upload(g, &g->blk.ct_row, B.ct_row, nt * 2);
        if (st == AG_OK) st = dev_upload(g, &g->blk.ct_start, B.ct_start, nt * nb);
        if (st == AG_OK) st = dev_upload(g, &g->blk.ct_pre, B.ct_pre, nt * (nb + 1));
        if (st == AG_OK) st = dev_upload(g, &g->blk.c_item, B.c_item, (size_t)B.c_n_items * 2);
        if (st == AG_OK && B.n_hgrp) st = dev_upload(g, &g->blk.hgrp, B.hgrp, (size_t)B.n_hgrp * 3);
        if (st == AG_OK && B.sorted) st = dev_upload(g, &g->blk.vsp, B.vsp, B.mb);
        if (st == AG_OK && B.sorted) st = dev_upload(g, &g->blk.tbase, B.tbase, B.n_items);
        if (st == AG_OK) {
            size_t pb = (size_t)(B.nv ? B.nv : 1) * sizeof(uint64_t);
            he = hipMalloc((void **)&g->vpart, pb);
            if (he != hipSuccess)
                st = ag_hip_fail(he, "partial-sum allocation", __FILE__, __LINE__);
            g->device_bytes += pb;
        }
        if (st == AG_OK) {
            // tickets only when pass 1 holds tiles alone (no AG_ITEM_REG item, whose FP64
            // error terms go to the running workgroup's partial); by default when the pieces
            // are at least half the in-edges (sources spread evenly over the blocks: there a
            // fixed round-robin lets the XCD's workgroups drift over more blocks than its L2
            // holds; on skewed graphs it measured faster)
            uint32_t reg = 0;
            for (uint32_t i = 0; i < B.n_items; i++)
                reg |= (B.item[2 * i + 1] & 15u) == AG_ITEM_REG;
            const uint32_t order = g->params.tile_order;
            g->bdyn = !reg && (order == AG_TILES_TICKET ||
                               (order == AG_TILES_AUTO && 2ull * B.nv >= B.mb));
            const size_t tb = 2 * AG_NXCD * sizeof(uint32_t);
            he = hipMalloc((void **)&g->btick, tb);
            if (he == hipSuccess)
                he = hipMemsetAsync(g->btick, 0, tb, dev->stream);
            if (he != hipSuccess)
                st = ag_hip_fail(he, "ticket counter allocation", __FILE__, __LINE__);
            g->device_bytes += tb;
        }
    }
    // the host plan arrays must outlive the asynchronous copies
    if (st == AG_OK)
        st = ag_wait_stream(dev);
    ag_plan_free(&plan);
    if (st != AG_OK) {
        ag_graph_free(g);
        return st;
    }
    *out = g;
    return AG_OK;
}

extern "C" int ag_graph_set_fixed_weights(ag_dgraph *g, const uint32_t *inv_edges_idx,
                                          const uint32_t *inv_degree, const uint32_t *wq)
{
    if (!g || (g->n && (!inv_edges_idx || !inv_degree)) || (g->m && !wq))
        return ag_set_error(AG_ERR_INVALID, "null argument");
    uint64_t total = 0;
    for (uint32_t r = 0; r < g->n; r++)
        total += inv_degree[g->h_d2o[r]];
    if (total != g->m)
        return ag_set_error(AG_ERR_INVALID, "degrees do not match the uploaded graph");
    uint32_t *buf = (uint32_t *)malloc((size_t)(g->m ? g->m : 1) * sizeof(uint32_t));
    if (!buf)
        return ag_set_error(AG_ERR_OOM, "host allocation failed");
    uint64_t o = 0;
    for (uint32_t r = 0; r < g->n; r++) {
        uint32_t v = g->h_d2o[r];
        memcpy(buf + o, wq + inv_edges_idx[v], (size_t)inv_degree[v] * sizeof(uint32_t));
        o += inv_degree[v];
    }
    hipError_t e = hipSetDevice(g->dev->ordinal);
    if (e == hipSuccess && !g->wq) {
        e = hipMalloc((void **)&g->wq, (size_t)(g->m ? g->m : 1) * sizeof(uint32_t));
        if (e == hipSuccess)
            g->device_bytes += (size_t)g->m * sizeof(uint32_t);
    }
    if (e == hipSuccess && g->m)
        e = hipMemcpyAsync(g->wq, buf, (size_t)g->m * sizeof(uint32_t), hipMemcpyHostToDevice,
                           g->dev->stream);
    int st = (e == hipSuccess) ? ag_wait_stream(g->dev)
                               : ag_hip_fail(e, "fixed weight upload", __FILE__, __LINE__);
    free(buf);
    if (st == AG_OK)
        g->wq_user = 1;
    return st;
}

extern "C" int ag_graph_get_info(const ag_dgraph *g, ag_graph_info *info)
{
    if (!g || !info)
        return ag_set_error(AG_ERR_INVALID, "graph or info is null");
    memset(info, 0, sizeof(*info));
    info->num_vertices = g->n;
    info->num_edges = g->m;
    for (int b = 0; b < AG_NBINS; b++) {
        info->bin_rows[b] = g->bin_row[b + 1] - g->bin_row[b];
        info->bin_edges[b] = g->bin_edges[b];
        info->bin_blocks[b] = g->bin_blocks[b];
    }
    info->num_tiles = g->n_tiles;
    info->num_segments = g->n_segs;
    info->has_weights = g->w != NULL;
    info->device_bytes = g->device_bytes;
    info->params = g->params;
    info->src_block_bits = g->blk.bits;
    info->src_blocks = g->blk.nblocks;
    info->virtual_rows = g->blk.nv;
    info->blocked_row0 = g->blk.bits ? g->blk.row0 : g->n;
    info->hot_segments = g->blk.n_hot;
    info->hot_segment_bits = g->blk.hot_bits;
    info->hot_edges = g->blk.hot_edges;
    info->top_sources = g->blk.bits && g->blk.top0 < g->n ? g->n - g->blk.top0 : 0u;
    info->top_edges = info->top_sources ? g->blk.top_edges : 0u;
    info->sorted_tiles = g->blk.bits && g->blk.sorted ? 1u : 0u;
    return AG_OK;
}

// ------------------------------------------------------------------ copy bandwidth

typedef uint32_t ag_u32x4 __attribute__((ext_vector_type(4)));

__global__ void ag_copy_kernel(const ag_u32x4 *__restrict__ in, ag_u32x4 *__restrict__ out,
                               uint64_t n16)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        __builtin_nontemporal_store(__builtin_nontemporal_load(&in[i]), &out[i]);
}

extern "C" int ag_device_copy_bandwidth(ag_device *dev, uint64_t bytes, uint32_t reps, double *gbps)
{
    if (!dev || !gbps || bytes < 16 || reps == 0)
        return ag_set_error(AG_ERR_INVALID, "bad argument");
    AG_HIP(hipSetDevice(dev->ordinal));
    const uint64_t n16 = bytes / 16;
    ag_u32x4 *a = NULL, *b = NULL;
    AG_HIP(hipMalloc((void **)&a, n16 * 16));
    hipError_t e = hipMalloc((void **)&b, n16 * 16);
    hipEvent_t e0 = NULL, e1 = NULL;
    if (e == hipSuccess)
        e = hipMemsetAsync(a, 1, n16 * 16, dev->stream);
    if (e == hipSuccess)
        e = hipEventCreate(&e0);
    if (e == hipSuccess)
        e = hipEventCreate(&e1);
    const uint32_t grid = dev->num_cus * 8;
    if (e == hipSuccess) {
        ag_copy_kernel<<<grid, 256, 0, dev->stream>>>(a, b, n16);  // warm-up
        e = hipEventRecord(e0, dev->stream);
    }
    for (uint32_t r = 0; r < reps && e == hipSuccess; r++) {
        ag_copy_kernel<<<grid, 256, 0, dev->stream>>>(a, b, n16);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipEventRecord(e1, dev->stream);
    int st = (e == hipSuccess) ? ag_wait_stream(dev) : ag_hip_fail(e, "copy bandwidth", __FILE__, __LINE__);
    if (st == AG_OK) {
        float ms = 0.0f;
        e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess)
            st = ag_hip_fail(e, "hipEventElapsedTime", __FILE__, __LINE__);
        else
            *gbps = (double)(2 * n16 * 16) * reps / ((double)ms * 1e-3) / 1e9;
    }
    if (e0)
        (void)hipEventDestroy(e0);
    if (e1)
        (void)hipEventDestroy(e1);
    (void)hipFree(a);
    if (b)
        (void)hipFree(b);
    return st;
}

// ------------------------------------------------------------------ fixed-point test hooks

__global__ void fixed_encode_kernel(uint32_t bits, const float *in, uint64_t *out, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float x = in[i];
    uint64_t q;
    if (bits == 64)
        q = ag_dev_double_to_fixed64((double)x);
    else if (bits == 32)
        q = ag_dev_float_to_fixed32(x);
    else if (bits == 16)
        q = ag_dev_float_to_fixed16(x);
    else
        q = ag_dev_float_to_fixed8(x);
    out[i] = q;
}

__global__ void fixed_decode_kernel(uint32_t bits, const uint64_t *in, double *out, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint64_t q = in[i];
    double v;
    if (bits == 64)
        v = ag_dev_fixed64_to_double(q);
    else if (bits == 32)
        v = (double)ag_dev_fixed32_to_float(q);
    else if (bits == 16)
        v = (double)ag_dev_fixed16_to_float(q);
    else
        v = (double)ag_dev_fixed8_to_float(q);
    out[i] = v;
}

static int check_bits(uint32_t bits)
{
    if (bits != 64 && bits != 32 && bits != 16 && bits != 8)
        return ag_set_error(AG_ERR_INVALID, "fixed-point width must be 64, 32, 16 or 8 (got %u)", bits);
    return AG_OK;
}

extern "C" int ag_fixed_encode(ag_device *dev, uint32_t bits, const float *in, uint64_t *out, uint32_t n)
{
    if (!dev || (n && (!in || !out)))
        return ag_set_error(AG_ERR_INVALID, "null argument");
    int st = check_bits(bits);
    if (st != AG_OK || n == 0)
        return st;
    AG_HIP(hipSetDevice(dev->ordinal));
    float *din = NULL;
    uint64_t *dout = NULL;
    AG_HIP(hipMalloc((void **)&din, (size_t)n * sizeof(float)));
    hipError_t e = hipMalloc((void **)&dout, (size_t)n * sizeof(uint64_t));
    if (e == hipSuccess)
        e = hipMemcpyAsync(din, in, (size_t)n * sizeof(float), hipMemcpyHostToDevice, dev->stream);
    if (e == hipSuccess) {
        fixed_encode_kernel<<<(n + 255) / 256, 256, 0, dev->stream>>>(bits, din, dout, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, dout, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToHost, dev->stream);
    int rc = (e == hipSuccess) ? ag_wait_stream(dev) : ag_hip_fail(e, "fixed encode", __FILE__, __LINE__);
    (void)hipFree(din);
    if (dout)
        (void)hipFree(dout);
    return rc;
}

extern "C" int ag_fixed_decode(ag_device *dev, uint32_t bits, const uint64_t *in, double *out, uint32_t n)
{
    if (!dev || (n && (!in || !out)))
        return ag_set_error(AG_ERR_INVALID, "null argument");
    int st = check_bits(bits);
    if (st != AG_OK || n == 0)
        return st;
    AG_HIP(hipSetDevice(dev->ordinal));
    uint64_t *din = NULL;
    double *dout = NULL;
    AG_HIP(hipMalloc((void **)&din, (size_t)n * sizeof(uint64_t)));
    hipError_t e = hipMalloc((void **)&dout, (size_t)n * sizeof(double));
    if (e == hipSuccess)
        e = hipMemcpyAsync(din, in, (size_t)n * sizeof(uint64_t), hipMemcpyHostToDevice, dev->stream);
    if (e == hipSuccess) {
        fixed_decode_kernel<<<(n + 255) / 256, 256, 0, dev->stream>>>(bits, din, dout, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, dout, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, dev->stream);
    int rc = (e == hipSuccess) ? ag_wait_stream(dev) : ag_hip_fail(e, "fixed decode", __FILE__, __LINE__);
    (void)hipFree(din);
    if (dout)
        (void)hipFree(dout);
    return rc;
}
