// ag_build.hip -- graph construction on the GPU.
//
// The step before the path: the synthetic generators and the CSR + inverse-CSR builder of
// graph_csr.c, producing the same arrays, array for array, on the device. (The reference takes
// its graphs from OpenGraph's builder, absent here; its options are at
// 02_capi_graph/src/main/accel-graph.c:71-72,143-145.)
//
// Generators. graph_csr.c cuts the edge stream into chunks of 2^20 edges, each drawn from its
// own MT19937 stream seeded seed ^ chunk * 0x9E3779B9; RMAT draws one uniform double
// (u32 / 2^32) per level, Erdos-Renyi two u32 per edge. Here one 64-lane workgroup runs one
// chunk: the 624-word state lives in LDS and each twist runs in four dependency phases (the
// recurrence reads words ahead of i before they change and words behind i after), the 624
// tempered outputs go to a two-block ring, and every edge whose draws are all produced is
// emitted (an edge's draws span at most two consecutive blocks).
//
// Builder. Stable radix sort (rocPRIM) of (source, destination) keys with the input index as
// payload -- rows by source, destinations ascending, duplicates in input order -- then keep
// flags (no self-loop, not a repeat of the previous key) + an exclusive scan + a scatter for
// the simplification (the first occurrence of a duplicate keeps its weight), per-vertex lower
// bounds for offsets and degrees, and a second stable sort of (destination, source) keys with
// the forward index as payload for the inverse CSR (rows by destination, sources ascending).
#include <stdlib.h>
#include <string.h>

#include <thread>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "ag_internal.h"
#include "graphCSR.h"

#define MT_N 624u
#define MT_M 397u
#define GEN_CHUNK (1u << 20)
#define GEN_WG 64u
#define BUILD_WG 256u

namespace {

// ---------------------------------------------------------------- MT19937 on a workgroup

__device__ __forceinline__ uint32_t mt_temper(uint32_t y)
{
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

// One twist of the state in LDS. The sequential recurrence updates i = 0..623 in order and
// reads mt[i], mt[i + 1] and mt[(i + 397) % 624]: indices ahead of i still hold old words,
// indices behind it new ones. Phase [0, 227) reads only old words; [227, 454) and [454, 623)
// read new words 227 back (finished by the previous phase); 623 reads the new mt[0] and
// mt[396]. Each phase reads all its inputs before any lane writes.
__device__ __forceinline__ void mt_twist_lds(uint32_t *mt)
{
    const uint32_t lane = threadIdx.x;
    const uint32_t lo[4] = {0u, 227u, 454u, 623u}, hi[4] = {227u, 454u, 623u, 624u};
#pragma unroll
    for (int ph = 0; ph < 4; ph++) {
        uint32_t nv[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t i = lo[ph] + lane + (uint32_t)k * GEN_WG;
            nv[k] = 0u;
            if (i < hi[ph]) {
                const uint32_t y = (mt[i] & 0x80000000u) | (mt[(i + 1u) % MT_N] & 0x7fffffffu);
                uint32_t v = mt[(i + MT_M) % MT_N] ^ (y >> 1);
                if (y & 1u)
                    v ^= 0x9908b0dfu;
                nv[k] = v;
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t i = lo[ph] + lane + (uint32_t)k * GEN_WG;
            if (i < hi[ph])
                mt[i] = nv[k];
        }
        __syncthreads();
    }
}

struct Ring {
    const uint32_t (*r)[MT_N];
    __device__ __forceinline__ uint32_t draw(uint64_t j) const
    {
        return r[(j / MT_N) & 1u][j % MT_N];
    }
};

// RMAT edge from draws j0 .. j0 + scale - 1 (graph_csr.c graphGenerateRMATEdges)
struct RmatEdge {
    uint32_t scale;
    double a, ab, abc;
    __device__ __forceinline__ void operator()(uint64_t e, uint64_t j0, const Ring &R,
                                               uint32_t *src, uint32_t *dst) const
    {
        uint32_t s = 0u, d = 0u;
        for (uint32_t lvl = 0; lvl < scale; lvl++) {
            const double r = (double)R.draw(j0 + lvl) * (1.0 / 4294967296.0);
            const uint32_t bit = 1u << (scale - 1u - lvl);
            if (r < a) {
            } else if (r < ab) {
                d |= bit;
            } else if (r < abc) {
                s |= bit;
            } else {
                s |= bit;
                d |= bit;
            }
        }
        src[e] = s;
        dst[e] = d;
    }
};

// Erdos-Renyi edge from draws j0, j0 + 1 (graph_csr.c graphCSRGenerateER)
struct ErEdge {
    uint32_t n;
    __device__ __forceinline__ void operator()(uint64_t e, uint64_t j0, const Ring &R,
                                               uint32_t *src, uint32_t *dst) const
    {
        src[e] = (uint32_t)(((uint64_t)R.draw(j0) * n) >> 32);
        dst[e] = (uint32_t)(((uint64_t)R.draw(j0 + 1) * n) >> 32);
    }
};

// One chunk of edges per workgroup; `per_edge` draws per edge (<= 31).
template <class F>
__global__ __launch_bounds__(GEN_WG) void gen_kernel(uint64_t m, uint32_t seed, uint32_t per_edge,
                                                      F f, uint32_t *src, uint32_t *dst)
{
    __shared__ uint32_t mt[MT_N];
    __shared__ uint32_t ring[2][MT_N];
    const uint64_t ch = blockIdx.x;
    const uint64_t e0 = ch * GEN_CHUNK;
    const uint32_t ne = (uint32_t)((m - e0) < GEN_CHUNK ? (m - e0) : GEN_CHUNK);
    if (threadIdx.x == 0) {
        uint32_t x = seed ^ (uint32_t)(ch * 0x9E3779B9ull);
        mt[0] = x;
        for (uint32_t i = 1; i < MT_N; i++) {
            x = 1812433253u * (x ^ (x >> 30)) + i;
            mt[i] = x;
        }
    }
    __syncthreads();
    const Ring R{ring};
    uint32_t done = 0;
    for (uint32_t p = 0; done < ne; p++) {
        mt_twist_lds(mt);
        for (uint32_t k = threadIdx.x; k < MT_N; k += GEN_WG)
            ring[p & 1u][k] = mt_temper(mt[k]);
        __syncthreads();
        // edges whose draws are all produced: draws (p + 1) * 624 so far
        const uint64_t avail = (uint64_t)(p + 1u) * MT_N / per_edge;
        const uint32_t upto = avail < ne ? (uint32_t)avail : ne;
        for (uint32_t e = done + threadIdx.x; e < upto; e += GEN_WG)
            f(e0 + e, (uint64_t)e * per_edge, R, src, dst);
        done = upto;
        __syncthreads();  // the next twist's block overwrites the older ring slot
    }
}

// ---------------------------------------------------------------- CSR builder

__global__ void pack_keys_kernel(uint64_t m, const uint32_t *hi, const uint32_t *lo,
                                 uint32_t shift, uint64_t *key, uint32_t *val)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) {
        key[i] = ((uint64_t)hi[i] << shift) | lo[i];
        val[i] = (uint32_t)i;
    }
}

__global__ void keep_kernel(uint64_t m, const uint64_t *key, uint32_t shift, int simplify,
                            uint32_t *keep)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m)
        return;
    uint32_t k = 1u;
    if (simplify) {
        const uint64_t kk = key[i];
        const uint32_t s = (uint32_t)(kk >> shift), d = (uint32_t)(kk & ((1ull << shift) - 1ull));
        k = (s != d && (i == 0 || key[i - 1] != kk)) ? 1u : 0u;
    }
    keep[i] = k;
}

__global__ void scatter_kernel(uint64_t m, const uint64_t *key, const uint32_t *val,
                               const uint32_t *keep, const uint32_t *pos, uint32_t shift,
                               const float *w_in, uint32_t *fs, uint32_t *fd, float *fw)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m || !keep[i])
        return;
    const uint64_t kk = key[i];
    const uint32_t p = pos[i];
    fs[p] = (uint32_t)(kk >> shift);
    fd[p] = (uint32_t)(kk & ((1ull << shift) - 1ull));
    if (fw)
        fw[p] = w_in[val[i]];
}

__global__ void unpack_inverse_kernel(uint32_t m, const uint64_t *key, const uint32_t *val,
                                      uint32_t shift, const float *fw, uint32_t *is,
                                      uint32_t *id, float *iw)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m)
        return;
    const uint64_t kk = key[k];
    is[k] = (uint32_t)(kk >> shift);
    id[k] = (uint32_t)(kk & ((1ull << shift) - 1ull));
    if (iw)
        iw[k] = fw[val[k]];
}

// Row offsets and lengths of a row-sorted edge list: first index with row >= v.
__global__ void offsets_kernel(uint32_t n, uint32_t m, const uint32_t *row, uint32_t *idx,
                               uint32_t *deg)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n)
        return;
    uint32_t lo = 0, hi = m;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (row[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    uint32_t lo2 = lo, hi2 = m;
    while (lo2 < hi2) {
        const uint32_t mid = lo2 + (hi2 - lo2) / 2;
        if (row[mid] < v + 1u)
            lo2 = mid + 1;
        else
            hi2 = mid;
    }
    idx[v] = lo;
    deg[v] = lo2 - lo;
}

// Per-workgroup maximum of ids (validation) or of weights.
template <typename T>
__global__ __launch_bounds__(BUILD_WG) void max_kernel(uint64_t m, const T *a, const T *b,
                                                       T *part)
{
    __shared__ T s[BUILD_WG];
    T v = (T)0;
    for (uint64_t i = (uint64_t)blockIdx.x * BUILD_WG + threadIdx.x; i < m;
         i += (uint64_t)gridDim.x * BUILD_WG) {
        v = a[i] > v ? a[i] : v;
        if (b)
            v = b[i] > v ? b[i] : v;
    }
    s[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t w = BUILD_WG / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w && s[threadIdx.x + w] > s[threadIdx.x])
            s[threadIdx.x] = s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        part[blockIdx.x] = s[0];
}

uint32_t bits_for(uint32_t n)
{
    uint32_t b = 1;
    while (b < 32 && (1ull << b) < (uint64_t)n)
        b++;
    return b;
}

// Device allocations of one build, freed together.
struct Pool {
    void *p[24];
    int k;
    int status;
    template <typename T>
    T *get(size_t count)
    {
        if (status != AG_OK)
            return NULL;
        void *q = NULL;
        hipError_t e = hipMalloc(&q, (count ? count : 1) * sizeof(T));
        if (e != hipSuccess) {
            status = ag_hip_fail(e, "graph build allocation", __FILE__, __LINE__);
            return NULL;
        }
        p[k++] = q;
        return (T *)q;
    }
    void release(void)
    {
        for (int i = 0; i < k; i++)
            (void)hipFree(p[i]);
        k = 0;
    }
};

template <typename T>
int device_max(ag_device *dev, uint64_t m, const T *a, const T *b, T *out)
{
    const uint32_t nb = 1024;
    T *part = NULL;
    AG_HIP(hipMalloc((void **)&part, nb * sizeof(T)));
    T host[1024];
    max_kernel<T><<<nb, BUILD_WG, 0, dev->stream>>>(m, a, b, part);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(host, part, nb * sizeof(T), hipMemcpyDeviceToHost, dev->stream);
    int st = (e == hipSuccess) ? ag_wait_stream(dev) : ag_hip_fail(e, "max", __FILE__, __LINE__);
    (void)hipFree(part);
    if (st != AG_OK)
        return st;
    T v = (T)0;
    for (uint32_t i = 0; i < nb; i++)
        v = host[i] > v ? host[i] : v;
    *out = v;
    return AG_OK;
}

// Stable sort of (key, val) pairs into (key2, val2) on bits [0, end_bit).
int sort_pairs(ag_device *dev, uint64_t m, uint64_t *key, uint64_t *key2, uint32_t *val,
               uint32_t *val2, uint32_t end_bit)
{
    size_t tmp = 0;
    AG_HIP(rocprim::radix_sort_pairs(nullptr, tmp, key, key2, val, val2, (size_t)m, 0u, end_bit,
                                     dev->stream));
    void *t = NULL;
    AG_HIP(hipMalloc(&t, tmp ? tmp : 8));
    hipError_t e = rocprim::radix_sort_pairs(t, tmp, key, key2, val, val2, (size_t)m, 0u, end_bit,
                                             dev->stream);
    int st = (e == hipSuccess) ? ag_wait_stream(dev) : ag_hip_fail(e, "radix sort", __FILE__, __LINE__);
    (void)hipFree(t);
    return st;
}

int exclusive_scan_u32(ag_device *dev, uint64_t m, const uint32_t *in, uint32_t *out)
{
    size_t tmp = 0;
    AG_HIP(rocprim::exclusive_scan(nullptr, tmp, in, out, 0u, (size_t)m, rocprim::plus<uint32_t>(),
                                   dev->stream));
    void *t = NULL;
    AG_HIP(hipMalloc(&t, tmp ? tmp : 8));
    hipError_t e = rocprim::exclusive_scan(t, tmp, in, out, 0u, (size_t)m,
                                           rocprim::plus<uint32_t>(), dev->stream);
    int st = (e == hipSuccess) ? ag_wait_stream(dev) : ag_hip_fail(e, "scan", __FILE__, __LINE__);
    (void)hipFree(t);
    return st;
}

inline uint32_t grid_of(uint64_t m, uint32_t wg)
{
    const uint64_t g = (m + wg - 1) / wg;
    return (uint32_t)(g ? g : 1);
}

// The builder proper on device edge arrays src/dst[m] (and weights w[m] or NULL).
int build_on_device(ag_device *dev, uint32_t n, uint64_t m, const uint32_t *src,
                    const uint32_t *dst, const float *w, int simplify, struct GraphCSR **out)
{
    hipStream_t s = dev->stream;
    const uint32_t shift = bits_for(n);
    const uint32_t end_bit = 2u * shift;
    Pool P;
    P.k = 0;
    P.status = AG_OK;
    uint32_t mm = 0;
    uint32_t *fs = NULL, *fd = NULL;
    float *fw = NULL;
    int st = AG_OK;
    if (m > 0) {
        // forward: rows by source, destinations ascending, duplicates in input order; the
        // sort temporaries (T) are released before the inverse pass
        Pool T;
        T.k = 0;
        T.status = AG_OK;
        uint64_t *key = T.get<uint64_t>(m), *key2 = T.get<uint64_t>(m);
        uint32_t *val = T.get<uint32_t>(m), *val2 = T.get<uint32_t>(m);
        uint32_t *keep = T.get<uint32_t>(m), *pos = T.get<uint32_t>(m);
        if (T.status != AG_OK) {
            T.release();
            return T.status;
        }
        pack_keys_kernel<<<grid_of(m, BUILD_WG), BUILD_WG, 0, s>>>(m, src, dst, shift, key, val);
        st = hipGetLastError() == hipSuccess ? AG_OK : ag_set_error(AG_ERR_HIP, "pack launch");
        if (st == AG_OK)
            st = sort_pairs(dev, m, key, key2, val, val2, end_bit);
        if (st == AG_OK) {
            keep_kernel<<<grid_of(m, BUILD_WG), BUILD_WG, 0, s>>>(m, key2, shift, simplify, keep);
            st = hipGetLastError() == hipSuccess ? AG_OK : ag_set_error(AG_ERR_HIP, "keep launch");
        }
        if (st == AG_OK)
            st = exclusive_scan_u32(dev, m, keep, pos);
        uint32_t tail[2] = {0, 0};
        if (st == AG_OK) {
            hipError_t e = hipMemcpyAsync(&tail[0], pos + (m - 1), 4, hipMemcpyDeviceToHost, s);
            if (e == hipSuccess)
                e = hipMemcpyAsync(&tail[1], keep + (m - 1), 4, hipMemcpyDeviceToHost, s);
            st = (e == hipSuccess) ? ag_wait_stream(dev) : ag_hip_fail(e, "count", __FILE__, __LINE__);
        }
        if (st == AG_OK) {
            mm = tail[0] + tail[1];
            fs = P.get<uint32_t>(mm);
            fd = P.get<uint32_t>(mm);
            if (w)
                fw = P.get<float>(mm);
            st = P.status;
        }
        if (st == AG_OK) {
            scatter_kernel<<<grid_of(m, BUILD_WG), BUILD_WG, 0, s>>>(m, key2, val2, keep, pos, shift,
                                                                     w, fs, fd, fw);
            st = hipGetLastError() == hipSuccess ? AG_OK : ag_set_error(AG_ERR_HIP, "scatter launch");
        }
        if (st == AG_OK)
            st = ag_wait_stream(dev);
        T.release();
        if (st != AG_OK) {
            P.release();
            return st;
        }
    }
    struct GraphCSR *g = graphCSRAlloc(n, mm, w != NULL);
    if (!g) {
        P.release();
        return ag_set_error(AG_ERR_OOM, "host allocation failed");
    }
    // fault the host arrays in on all cores while the device finishes the graph
    std::thread prefault(graphCSRPrefault, g);
    uint32_t *vidx = P.get<uint32_t>(n), *vdeg = P.get<uint32_t>(n);
    uint32_t *iidx = P.get<uint32_t>(n), *ideg = P.get<uint32_t>(n);
    uint32_t *is = P.get<uint32_t>(mm), *id = P.get<uint32_t>(mm);
    float *iw = w ? P.get<float>(mm) : NULL;
    st = P.status;
    if (st == AG_OK && n) {
        offsets_kernel<<<grid_of(n, BUILD_WG), BUILD_WG, 0, s>>>(n, mm, fs, vidx, vdeg);
        st = hipGetLastError() == hipSuccess ? AG_OK : ag_set_error(AG_ERR_HIP, "offsets launch");
    }
    if (st == AG_OK && mm) {
        // inverse: rows by destination, sources ascending, payload = forward index
        uint64_t *key = P.get<uint64_t>(mm), *key2 = P.get<uint64_t>(mm);
        uint32_t *val = P.get<uint32_t>(mm), *val2 = P.get<uint32_t>(mm);
        st = P.status;
        if (st == AG_OK) {
            pack_keys_kernel<<<grid_of(mm, BUILD_WG), BUILD_WG, 0, s>>>(mm, fd, fs, shift, key, val);
            st = hipGetLastError() == hipSuccess ? AG_OK : ag_set_error(AG_ERR_HIP, "pack launch");
        }
        if (st == AG_OK)
            st = sort_pairs(dev, mm, key, key2, val, val2, end_bit);
        if (st == AG_OK) {
            unpack_inverse_kernel<<<grid_of(mm, BUILD_WG), BUILD_WG, 0, s>>>(mm, key2, val2, shift,
                                                                             fw, is, id, iw);
            st = hipGetLastError() == hipSuccess ? AG_OK : ag_set_error(AG_ERR_HIP, "unpack launch");
        }
    }
    if (st == AG_OK && n) {
        offsets_kernel<<<grid_of(n, BUILD_WG), BUILD_WG, 0, s>>>(n, mm, is, iidx, ideg);
        st = hipGetLastError() == hipSuccess ? AG_OK : ag_set_error(AG_ERR_HIP, "offsets launch");
    }
    float maxw = 0.0f;
    if (st == AG_OK && w && mm)
        st = device_max<float>(dev, mm, fw, NULL, &maxw);
    prefault.join();
    // results to the host structure: forward rows (source, destination, weight), inverse rows
    // (destination, source, weight); in-degree of one view = row length of the other
    if (st == AG_OK) {
        struct {
            void *h;
            const void *d;
            size_t bytes;
        } cp[] = {
            {g->sorted_edges_array->edges_array_src, fs, (size_t)mm * 4},
            {g->sorted_edges_array->edges_array_dest, fd, (size_t)mm * 4},
            {g->inverse_sorted_edges_array->edges_array_src, is, (size_t)mm * 4},
            {g->inverse_sorted_edges_array->edges_array_dest, id, (size_t)mm * 4},
            {g->vertices->edges_idx, vidx, (size_t)n * 4},
            {g->vertices->out_degree, vdeg, (size_t)n * 4},
            {g->inverse_vertices->edges_idx, iidx, (size_t)n * 4},
            {g->inverse_vertices->out_degree, ideg, (size_t)n * 4},
            {g->vertices->in_degree, ideg, (size_t)n * 4},
            {g->inverse_vertices->in_degree, vdeg, (size_t)n * 4},
            {w ? g->sorted_edges_array->edges_array_weight : NULL, fw, (size_t)mm * 4},
            {w ? g->inverse_sorted_edges_array->edges_array_weight : NULL, iw, (size_t)mm * 4},
        };
        for (auto &c : cp) {
            if (!c.h || !c.bytes || st != AG_OK)
                continue;
            hipError_t e = hipMemcpyAsync(c.h, c.d, c.bytes, hipMemcpyDeviceToHost, s);
            if (e != hipSuccess)
                st = ag_hip_fail(e, "graph download", __FILE__, __LINE__);
        }
        if (st == AG_OK)
            st = ag_wait_stream(dev);
    }
    P.release();
    if (st != AG_OK) {
        graphCSRFree(g);
        return st;
    }
    g->max_weight = maxw;
    g->sorted_edges_array->max_weight = maxw;
    g->inverse_sorted_edges_array->max_weight = maxw;
    g->avg_degree = n ? (uint32_t)(mm / n) : 0;
    *out = g;
    return AG_OK;
}

template <class F>
int generate_device(ag_device *dev, uint64_t m, uint32_t seed, uint32_t per_edge, const F &f,
                    uint32_t *d_src, uint32_t *d_dst)
{
    const uint64_t nchunks = (m + GEN_CHUNK - 1) / GEN_CHUNK;
    if (nchunks)
        gen_kernel<F><<<(uint32_t)nchunks, GEN_WG, 0, dev->stream>>>(m, seed, per_edge, f, d_src,
                                                                    d_dst);
    AG_HIP(hipGetLastError());
    return AG_OK;
}

int rmat_functor(uint32_t scale, double a, double b, double c, RmatEdge *f)
{
    if (scale < 1 || scale > 31 || !(a >= 0) || !(b >= 0) || !(c >= 0) || a + b + c > 1.0)
        return ag_set_error(AG_ERR_INVALID, "bad RMAT parameters");
    f->scale = scale;
    f->a = a;
    f->ab = a + b;
    f->abc = a + b + c;
    return AG_OK;
}

// Row id of every edge of a CSR: the last row r with off[r] <= j.
__global__ __launch_bounds__(BUILD_WG) void row_ids_kernel(uint32_t n, uint32_t m,
                                                           const uint32_t *off, uint32_t *row)
{
    for (uint32_t j = blockIdx.x * BUILD_WG + threadIdx.x; j < m; j += gridDim.x * BUILD_WG) {
        uint32_t lo = 0, hi = n;   // off[lo] <= j < off[hi]
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (off[mid] <= j)
                lo = mid;
            else
                hi = mid;
        }
        row[j] = lo;
    }
}

__global__ __launch_bounds__(BUILD_WG) void iota_kernel(uint32_t m, uint32_t *a)
{
    for (uint32_t j = blockIdx.x * BUILD_WG + threadIdx.x; j < m; j += gridDim.x * BUILD_WG)
        a[j] = j;
}

// Out-edge k came from inverse edge idx[k]: its destination is that edge's row, its weight
// that edge's weight.
__global__ __launch_bounds__(BUILD_WG) void forward_gather_kernel(uint32_t m, const uint32_t *idx,
                                                                  const uint32_t *row,
                                                                  const float *w, uint32_t *dst,
                                                                  float *fw)
{
    for (uint32_t k = blockIdx.x * BUILD_WG + threadIdx.x; k < m; k += gridDim.x * BUILD_WG) {
        const uint32_t j = idx[k];
        dst[k] = row[j];
        if (fw)
            fw[k] = w[j];
    }
}

__global__ void last_offset_kernel(uint32_t n, uint32_t m, uint32_t *off)
{
    if (blockIdx.x == 0 && threadIdx.x == 0)
        off[n] = m;
}

// Push bin of every out-edge: the last bin whose first row is <= its destination.
__global__ __launch_bounds__(BUILD_WG) void pb_key_kernel(uint32_t m, const uint32_t *fdst,
                                                          const uint32_t *brow, uint32_t nbins,
                                                          uint32_t *key)
{
    for (uint32_t j = blockIdx.x * BUILD_WG + threadIdx.x; j < m; j += gridDim.x * BUILD_WG) {
        const uint32_t d = fdst[j];
        uint32_t lo = 0, hi = nbins;   // brow[lo] <= d < brow[hi]
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (brow[mid] <= d)
                lo = mid;
            else
                hi = mid;
        }
        key[j] = lo;
    }
}

// Slot k holds out-edge order[k] (bins ascending, forward order inside a bin).
__global__ __launch_bounds__(BUILD_WG) void pb_slot_kernel(uint32_t m, const uint32_t *order,
                                                           const uint32_t *skey,
                                                           const uint32_t *fdst,
                                                           const uint32_t *brow, uint32_t *pos,
                                                           uint16_t *dst)
{
    for (uint32_t k = blockIdx.x * BUILD_WG + threadIdx.x; k < m; k += gridDim.x * BUILD_WG) {
        const uint32_t e = order[k];
        pos[e] = k;
        dst[k] = (uint16_t)(fdst[e] - brow[skey[k]]);
    }
}

// Greedy push bins over rows [0, n): a bin takes rows while it has fewer than `rows` rows and
// its in-edges stay within `slots`; a row with more in-edges than `slots` is a bin alone.
// Writes the first row of every bin to brow (when not null); returns the bin count.
uint32_t push_bins_host(uint32_t n, const uint32_t *off, uint32_t rows, uint32_t slots,
                        uint32_t *brow)
{
    uint32_t nb = 0, r = 0;
    while (r < n) {
        const uint32_t r0 = r;
        uint64_t e = off[r + 1] - off[r];
        r++;
        if (e <= slots)
            while (r < n && r - r0 < rows && e + (off[r + 1] - off[r]) <= slots) {
                e += off[r + 1] - off[r];
                r++;
            }
        if (brow)
            brow[nb] = r0;
        nb++;
    }
    return nb;
}

} // namespace

// Forward CSR of an uploaded graph in device row ids: the inverse CSR (off/src) transposed
// by one stable radix sort of (source, inverse edge index) pairs, so each out-row lists its
// destinations in ascending order; weights follow their edges. Used by the push directions
// (PageRank and SpMV) and the worklist form of the data-driven pull.
int ag_graph_forward(ag_dgraph *g)
{
    if (g->fwd_off)
        return AG_OK;
    ag_device *dev = g->dev;
    AG_HIP(hipSetDevice(dev->ordinal));
    const uint32_t n = g->n, m = g->m;
    uint32_t *off = NULL, *fsrc = NULL, *fdst = NULL;
    float *fw = NULL;
    Pool T{};
    T.status = AG_OK;
    int st = AG_OK;
    hipError_t e = hipMalloc((void **)&off, ((size_t)n + 1) * sizeof(uint32_t));
    if (e == hipSuccess)
        e = hipMalloc((void **)&fsrc, (size_t)(m ? m : 1) * sizeof(uint32_t));
    if (e == hipSuccess)
        e = hipMalloc((void **)&fdst, (size_t)(m ? m : 1) * sizeof(uint32_t));
    if (e == hipSuccess && g->w)
        e = hipMalloc((void **)&fw, (size_t)(m ? m : 1) * sizeof(float));
    if (e != hipSuccess)
        st = ag_hip_fail(e, "forward CSR allocation", __FILE__, __LINE__);
    uint32_t *row = T.get<uint32_t>(m);
    uint32_t *idx = T.get<uint32_t>(m);
    uint32_t *sidx = T.get<uint32_t>(m);
    uint32_t *deg = T.get<uint32_t>(n);
    if (st == AG_OK)
        st = T.status;
    if (st == AG_OK && m) {
        const uint32_t gm = grid_of(m, BUILD_WG) < 8192u ? grid_of(m, BUILD_WG) : 8192u;
        row_ids_kernel<<<gm, BUILD_WG, 0, dev->stream>>>(n, m, g->off, row);
        iota_kernel<<<gm, BUILD_WG, 0, dev->stream>>>(m, idx);
        e = hipGetLastError();
        size_t tmp = 0;
        if (e == hipSuccess)
            e = rocprim::radix_sort_pairs(nullptr, tmp, g->src, fsrc, idx, sidx, (size_t)m, 0u,
                                          bits_for(n), dev->stream);
        void *t = NULL;
        if (e == hipSuccess)
            e = hipMalloc(&t, tmp ? tmp : 8);
        if (e == hipSuccess)
            e = rocprim::radix_sort_pairs(t, tmp, g->src, fsrc, idx, sidx, (size_t)m, 0u,
                                          bits_for(n), dev->stream);
        if (e == hipSuccess) {
            forward_gather_kernel<<<gm, BUILD_WG, 0, dev->stream>>>(m, sidx, row, g->w, fdst, fw);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            st = ag_wait_stream(dev);
        else
            st = ag_hip_fail(e, "forward CSR sort", __FILE__, __LINE__);
        if (t)
            (void)hipFree(t);
    }
    if (st == AG_OK && n) {
        offsets_kernel<<<grid_of(n, BUILD_WG), BUILD_WG, 0, dev->stream>>>(n, m, fsrc, off, deg);
        e = hipGetLastError();
        if (e != hipSuccess)
            st = ag_hip_fail(e, "forward CSR offsets", __FILE__, __LINE__);
    }
    if (st == AG_OK) {
        last_offset_kernel<<<1, 1, 0, dev->stream>>>(n, m, off);
        e = hipGetLastError();
        st = (e == hipSuccess) ? ag_wait_stream(dev)
                               : ag_hip_fail(e, "forward CSR offsets", __FILE__, __LINE__);
    }
    T.release();
    if (st != AG_OK) {
        (void)hipFree(off);
        (void)hipFree(fsrc);
        (void)hipFree(fdst);
        (void)hipFree(fw);
        return st;
    }
    g->fwd_off = off;
    g->fwd_src = fsrc;
    g->fwd_dst = fdst;
    g->fwd_w = fw;
    g->device_bytes += ((size_t)n + 1 + (fw ? 3 : 2) * (size_t)m) * sizeof(uint32_t);
    return AG_OK;
}

// Bins and slots of the propagation-blocked push (ag_push.h), once per graph: bins are ranges
// of device rows (greedy on the host over the in-degrees); the slots of a bin are its rows'
// in-edges in forward order (sources ascending), found by one stable radix sort of the
// forward edges by the bin of their destination.
int ag_graph_push_bins(ag_dgraph *g)
{
    if (g->pb_pos)
        return AG_OK;
    int st = ag_graph_forward(g);
    if (st != AG_OK)
        return st;
    ag_device *dev = g->dev;
    AG_HIP(hipSetDevice(dev->ordinal));
    const uint32_t n = g->n, m = g->m;
    uint32_t *hoff = (uint32_t *)malloc(((size_t)n + 1) * sizeof(uint32_t));
    if (!hoff)
        return ag_set_error(AG_ERR_OOM, "push bins: host offsets");
    hipError_t e = hipMemcpy(hoff, g->off, ((size_t)n + 1) * sizeof(uint32_t),
                             hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        free(hoff);
        return ag_hip_fail(e, "push bins: offsets download", __FILE__, __LINE__);
    }
    const uint32_t nb = push_bins_host(n, hoff, AG_PB_ROWS, AG_PB_SLOTS, NULL);
    uint32_t *hb = (uint32_t *)malloc(2 * ((size_t)nb + 1) * sizeof(uint32_t));
    if (!hb) {
        free(hoff);
        return ag_set_error(AG_ERR_OOM, "push bins: host bins");
    }
    uint32_t *hrow = hb, *hboff = hb + nb + 1;
    push_bins_host(n, hoff, AG_PB_ROWS, AG_PB_SLOTS, hrow);
    hrow[nb] = n;
    for (uint32_t b = 0; b <= nb; b++)
        hboff[b] = hoff[hrow[b]];
    free(hoff);

    uint32_t *brow = NULL, *pos = NULL;
    uint16_t *dst = NULL;
    Pool T{};
    T.status = AG_OK;
    e = hipMalloc((void **)&brow, 2 * ((size_t)nb + 1) * sizeof(uint32_t));
    if (e == hipSuccess)
        e = hipMalloc((void **)&pos, (size_t)(m ? m : 1) * sizeof(uint32_t));
    if (e == hipSuccess)
        e = hipMalloc((void **)&dst, (size_t)(m ? m : 1) * sizeof(uint16_t));
    if (e == hipSuccess)
        e = hipMemcpy(brow, hb, 2 * ((size_t)nb + 1) * sizeof(uint32_t), hipMemcpyHostToDevice);
    free(hb);
    if (e != hipSuccess)
        st = ag_hip_fail(e, "push bins allocation", __FILE__, __LINE__);
    uint32_t *key = T.get<uint32_t>(m);
    uint32_t *skey = T.get<uint32_t>(m);
    uint32_t *idx = T.get<uint32_t>(m);
    uint32_t *order = T.get<uint32_t>(m);
    if (st == AG_OK)
        st = T.status;
    if (st == AG_OK && m) {
        const uint32_t gm = grid_of(m, BUILD_WG) < 8192u ? grid_of(m, BUILD_WG) : 8192u;
        pb_key_kernel<<<gm, BUILD_WG, 0, dev->stream>>>(m, g->fwd_dst, brow, nb, key);
        iota_kernel<<<gm, BUILD_WG, 0, dev->stream>>>(m, idx);
        e = hipGetLastError();
        size_t tmp = 0;
        if (e == hipSuccess)
            e = rocprim::radix_sort_pairs(nullptr, tmp, key, skey, idx, order, (size_t)m, 0u,
                                          bits_for(nb), dev->stream);
        void *t = NULL;
        if (e == hipSuccess)
            e = hipMalloc(&t, tmp ? tmp : 8);
        if (e == hipSuccess)
            e = rocprim::radix_sort_pairs(t, tmp, key, skey, idx, order, (size_t)m, 0u,
                                          bits_for(nb), dev->stream);
        if (e == hipSuccess) {
            pb_slot_kernel<<<gm, BUILD_WG, 0, dev->stream>>>(m, order, skey, g->fwd_dst, brow,
                                                             pos, dst);
            e = hipGetLastError();
        }
        st = (e == hipSuccess) ? ag_wait_stream(dev)
                               : ag_hip_fail(e, "push bins sort", __FILE__, __LINE__);
        if (t)
            (void)hipFree(t);
    }
    T.release();
    if (st != AG_OK) {
        (void)hipFree(brow);
        (void)hipFree(pos);
        (void)hipFree(dst);
        return st;
    }
    g->pb_nbins = nb;
    g->pb_row = brow;
    g->pb_off = brow + nb + 1;
    g->pb_pos = pos;
    g->pb_dst = dst;
    g->device_bytes += 2 * ((size_t)nb + 1) * sizeof(uint32_t) +
                       (size_t)m * (sizeof(uint32_t) + sizeof(uint16_t));
    return AG_OK;
}

// ---------------------------------------------------------------- C-ABI

extern "C" int ag_generate_rmat_edges(ag_device *dev, uint32_t scale, uint64_t m, double a,
                                      double b, double c, uint32_t seed, uint32_t *src_out,
                                      uint32_t *dst_out)
{
    if (!dev || (m && (!src_out || !dst_out)) || m > 0xFFFFFFFFull)
        return ag_set_error(AG_ERR_INVALID, "null argument or more than 2^32 - 1 edges");
    RmatEdge f;
    int st = rmat_functor(scale, a, b, c, &f);
    if (st != AG_OK || m == 0)
        return st;
    AG_HIP(hipSetDevice(dev->ordinal));
    uint32_t *d = NULL;
    AG_HIP(hipMalloc((void **)&d, m * 2 * sizeof(uint32_t)));
    st = generate_device(dev, m, seed, scale, f, d, d + m);
    hipError_t e = hipSuccess;
    if (st == AG_OK)
        e = hipMemcpyAsync(src_out, d, m * 4, hipMemcpyDeviceToHost, dev->stream);
    if (st == AG_OK && e == hipSuccess)
        e = hipMemcpyAsync(dst_out, d + m, m * 4, hipMemcpyDeviceToHost, dev->stream);
    if (st == AG_OK)
        st = (e == hipSuccess) ? ag_wait_stream(dev) : ag_hip_fail(e, "edge download", __FILE__, __LINE__);
    (void)hipFree(d);
    return st;
}

extern "C" int ag_generate_er_edges(ag_device *dev, uint32_t n, uint64_t m, uint32_t seed,
                                    uint32_t *src_out, uint32_t *dst_out)
{
    if (!dev || n == 0 || (m && (!src_out || !dst_out)) || m > 0xFFFFFFFFull)
        return ag_set_error(AG_ERR_INVALID, "null argument, no vertices or too many edges");
    if (m == 0)
        return AG_OK;
    AG_HIP(hipSetDevice(dev->ordinal));
    uint32_t *d = NULL;
    AG_HIP(hipMalloc((void **)&d, m * 2 * sizeof(uint32_t)));
    int st = generate_device(dev, m, seed, 2u, ErEdge{n}, d, d + m);
    hipError_t e = hipSuccess;
    if (st == AG_OK)
        e = hipMemcpyAsync(src_out, d, m * 4, hipMemcpyDeviceToHost, dev->stream);
    if (st == AG_OK && e == hipSuccess)
        e = hipMemcpyAsync(dst_out, d + m, m * 4, hipMemcpyDeviceToHost, dev->stream);
    if (st == AG_OK)
        st = (e == hipSuccess) ? ag_wait_stream(dev) : ag_hip_fail(e, "edge download", __FILE__, __LINE__);
    (void)hipFree(d);
    return st;
}

extern "C" int ag_graph_csr_build(ag_device *dev, uint32_t n, uint64_t m, const uint32_t *src,
                                  const uint32_t *dst, const float *weight, int simplify,
                                  struct GraphCSR **out)
{
    if (!dev || !out || (m && (!src || !dst)) || m > 0xFFFFFFFFull)
        return ag_set_error(AG_ERR_INVALID, "null argument or more than 2^32 - 1 edges");
    *out = NULL;
    AG_HIP(hipSetDevice(dev->ordinal));
    uint32_t *d = NULL;
    float *dw = NULL;
    AG_HIP(hipMalloc((void **)&d, (m ? m : 1) * 2 * sizeof(uint32_t)));
    int st = AG_OK;
    hipError_t e = hipSuccess;
    if (m) {
        e = hipMemcpyAsync(d, src, m * 4, hipMemcpyHostToDevice, dev->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d + m, dst, m * 4, hipMemcpyHostToDevice, dev->stream);
        if (e == hipSuccess && weight) {
            e = hipMalloc((void **)&dw, m * sizeof(float));
            if (e == hipSuccess)
                e = hipMemcpyAsync(dw, weight, m * 4, hipMemcpyHostToDevice, dev->stream);
        }
        if (e != hipSuccess)
            st = ag_hip_fail(e, "edge upload", __FILE__, __LINE__);
    }
    // vertex count: given, or 1 + the largest id; every id must be < n
    uint32_t mx = 0;
    if (st == AG_OK && m)
        st = device_max<uint32_t>(dev, m, d, d + m, &mx);
    if (st == AG_OK && n == 0)
        n = m ? mx + 1u : 0u;
    if (st == AG_OK && m && mx >= n)
        st = ag_set_error(AG_ERR_INVALID, "vertex id %u >= num_vertices %u", mx, n);
    if (st == AG_OK)
        st = build_on_device(dev, n, m, d, d + m, dw, simplify, out);
    (void)hipFree(d);
    if (dw)
        (void)hipFree(dw);
    return st;
}

extern "C" int ag_graph_csr_generate_rmat(ag_device *dev, uint32_t scale, uint32_t edge_factor,
                                          double a, double b, double c, uint32_t seed,
                                          int simplify, struct GraphCSR **out)
{
    if (!dev || !out)
        return ag_set_error(AG_ERR_INVALID, "null argument");
    *out = NULL;
    RmatEdge f;
    int st = rmat_functor(scale, a, b, c, &f);
    if (st != AG_OK)
        return st;
    const uint64_t m = (uint64_t)edge_factor << scale;
    if (m > 0xFFFFFFFFull)
        return ag_set_error(AG_ERR_INVALID, "more than 2^32 - 1 edges");
    AG_HIP(hipSetDevice(dev->ordinal));
    uint32_t *d = NULL;
    AG_HIP(hipMalloc((void **)&d, (m ? m : 1) * 2 * sizeof(uint32_t)));
    st = generate_device(dev, m, seed, scale, f, d, d + m);
    if (st == AG_OK)
        st = build_on_device(dev, 1u << scale, m, d, d + m, NULL, simplify, out);
    (void)hipFree(d);
    return st;
}

extern "C" int ag_graph_csr_generate_er(ag_device *dev, uint32_t n, uint64_t m, uint32_t seed,
                                        int simplify, struct GraphCSR **out)
{
    if (!dev || !out || n == 0 || m > 0xFFFFFFFFull)
        return ag_set_error(AG_ERR_INVALID, "null argument, no vertices or too many edges");
    *out = NULL;
    AG_HIP(hipSetDevice(dev->ordinal));
    uint32_t *d = NULL;
    AG_HIP(hipMalloc((void **)&d, (m ? m : 1) * 2 * sizeof(uint32_t)));
    int st = generate_device(dev, m, seed, 2u, ErEdge{n}, d, d + m);
    if (st == AG_OK)
        st = build_on_device(dev, n, m, d, d + m, NULL, simplify, out);
    (void)hipFree(d);
    return st;
}
