/*
 * graph_io.c -- graph files: text edge lists, binary edge lists and a serialised CSR.
 *
 * The reference's driver reads a graph in one of three formats and can convert between them
 * (02_capi_graph/src/main/accel-graph.c:51-53 `-z`: [0] text edge list, [1] binary edge list,
 * [2] graphCSR binary; accel-graph.c:115-128 `-c`/`-x`: write the loaded graph in another
 * format). The readers themselves live in the absent OpenGraph submodule; what the reference
 * tree does hold are its test files (04_test_graphs/TEST/graphbrew/): `graph.bin` (u32 LE
 * pairs), `graph.org` (the same 19 edges as text, tab- or space-separated) and `graph.bin.txt`
 * (a relabelled, symmetrised copy). These routines read those files and follow their layout;
 * the serialised-CSR layout is this library's own (OpenGraph's is not in the tree).
 *
 * Text edge list: one edge per line, `src dst [weight] [ignored...]`, fields separated by
 * spaces, tabs or commas; blank lines and lines starting with '#' or '%' are skipped; CRLF
 * line ends are accepted. Ids are decimal u32 below 2^32-1 (the vertex count is 1 + the
 * largest id, as for the binary edge list). The file is memory-mapped and parsed by all
 * OpenMP threads, each on a range of whole lines: one pass counts edges, the second writes
 * them at the offsets the counts give, so the edge order is the file's.
 *
 * Serialised CSR (little-endian):
 *   40-byte header  "AGCSR\0\0\0", u32 version (1), u32 flags (bit 0: weighted),
 *                   u32 num_vertices, u32 num_edges, f32 max_weight, u32 avg_degree,
 *                   f32 forward max_weight, f32 inverse max_weight
 *   forward view    out_degree[n], in_degree[n], edges_idx[n], src[m], dest[m], (weight[m])
 *   inverse view    the same six arrays of the inverse view
 *   trailer         u64 checksum = sum over every u32 word w_j before it of w_j * (2j + 1),
 *                   modulo 2^64 (any single-bit change alters it: the multiplier is odd)
 * A load checks the header, the size, the checksum and then the structure the kernels index
 * by (offsets are the prefix sums of the degrees, row ids and destinations are in range and
 * sorted within rows, each view's in-degrees are the other view's row lengths and the forward
 * destinations count exactly the inverse row lengths), so a damaged file is an error, never
 * a graph whose indices run off its arrays.
 */
#include <errno.h>
#include <fcntl.h>
#include <math.h>
#include <omp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "graphCSR.h"

static __thread char last_error[320];

static void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(last_error, sizeof(last_error), fmt, ap);
    va_end(ap);
}

const char *graphIOLastError(void)
{
    return last_error;
}

/* ------------------------------------------------------------------ text edge list */

#define MAX_ID 0xFFFFFFFEu

static int is_sep(char c)
{
    return c == ' ' || c == '\t' || c == ',';
}

static int is_blank(char c)
{
    return c == ' ' || c == '\t' || c == '\r';
}

/* Decimal u32 at *p (< q); advances *p. 0 on success. */
static int parse_id(const char **p, const char *q, uint32_t *out)
{
    const char *s = *p;
    uint64_t v = 0;
    if (s >= q || *s < '0' || *s > '9')
        return -1;
    while (s < q && *s >= '0' && *s <= '9') {
        v = v * 10 + (uint64_t)(*s - '0');
        if (v > MAX_ID)
            return -1;
        s++;
    }
    *p = s;
    *out = (uint32_t)v;
    return 0;
}

/* A float token ending at a separator, a blank or q; advances *p. 0 on success. */
static int parse_weight(const char **p, const char *q, float *out)
{
    const char *s = *p, *e = s;
    while (e < q && !is_sep(*e) && !is_blank(*e))
        e++;
    char tok[64];
    size_t len = (size_t)(e - s);
    if (len == 0 || len >= sizeof(tok))
        return -1;
    memcpy(tok, s, len);
    tok[len] = '\0';
    char *end = NULL;
    errno = 0;
    float w = strtof(tok, &end);
    if (end != tok + len || errno == ERANGE || !isfinite(w))
        return -1;
    *p = e;
    *out = w;
    return 0;
}

/* One line [p, q). Returns 1 for an edge, 0 for a blank or comment line, -1 for an error. */
static int parse_line(const char *p, const char *q, int weighted, uint32_t *s, uint32_t *d,
                      float *w)
{
    while (p < q && is_blank(*p))
        p++;
    if (p == q || *p == '#' || *p == '%')
        return 0;
    if (parse_id(&p, q, s) != 0)
        return -1;
    if (p == q || !is_sep(*p))
        return -1;
    while (p < q && is_sep(*p))
        p++;
    if (parse_id(&p, q, d) != 0)
        return -1;
    if (weighted) {
        if (p == q || !is_sep(*p))
            return -1;
        while (p < q && is_sep(*p))
            p++;
        if (parse_weight(&p, q, w) != 0)
            return -1;
    }
    /* anything further must start with a separator (extra columns are ignored) */
    if (p < q && !is_sep(*p) && !is_blank(*p))
        return -1;
    return 1;
}

/* Parses the lines of [b, e); writes edges from index `at` when src != NULL. Returns the
 * edge count, or -1 with *bad set to the offending line's first byte. */
static int64_t parse_range(const char *buf, size_t b, size_t e, int weighted, uint32_t *src,
                           uint32_t *dst, float *wt, uint64_t at, uint32_t *maxid, size_t *bad)
{
    int64_t count = 0;
    uint32_t mx = 0;
    size_t pos = b;
    while (pos < e) {
        const char *nl = (const char *)memchr(buf + pos, '\n', e - pos);
        size_t le = nl ? (size_t)(nl - buf) : e;
        uint32_t s = 0, d = 0;
        float w = 0.0f;
        int r = parse_line(buf + pos, buf + le, weighted, &s, &d, &w);
        if (r < 0) {
            *bad = pos;
            return -1;
        }
        if (r > 0) {
            if (src) {
                src[at + (uint64_t)count] = s;
                dst[at + (uint64_t)count] = d;
                if (wt)
                    wt[at + (uint64_t)count] = w;
            }
            mx = s > mx ? s : mx;
            mx = d > mx ? d : mx;
            count++;
        }
        pos = le + 1;
    }
    *maxid = mx;
    return count;
}

struct GraphCSR *graphCSRLoadTextEdgeList(const char *path, int weighted, int simplify)
{
    int fd = open(path, O_RDONLY);
    if (fd < 0) {
        set_error("%s: cannot open (%s)", path, strerror(errno));
        return NULL;
    }
    struct stat st;
    if (fstat(fd, &st) != 0) {
        set_error("%s: cannot stat (%s)", path, strerror(errno));
        close(fd);
        return NULL;
    }
    const size_t size = (size_t)st.st_size;
    const char *buf = NULL;
    if (size > 0) {
        void *map = mmap(NULL, size, PROT_READ, MAP_PRIVATE, fd, 0);
        if (map == MAP_FAILED) {
            set_error("%s: cannot map (%s)", path, strerror(errno));
            close(fd);
            return NULL;
        }
        madvise(map, size, MADV_SEQUENTIAL);
        buf = (const char *)map;
    }
    close(fd);

    /* thread ranges start at line beginnings */
    int nt = size < (1u << 20) ? 1 : omp_get_max_threads();
    size_t *begin = (size_t *)calloc((size_t)nt + 1, sizeof(size_t));
    int64_t *cnt = (int64_t *)calloc((size_t)nt, sizeof(int64_t));
    size_t *bad = (size_t *)calloc((size_t)nt, sizeof(size_t));
    uint32_t *mx = (uint32_t *)calloc((size_t)nt, sizeof(uint32_t));
    uint32_t *src = NULL, *dst = NULL;
    float *wt = NULL;
    struct GraphCSR *g = NULL;
    if (!begin || !cnt || !bad || !mx) {
        set_error("%s: out of memory", path);
        goto done;
    }
    for (int t = 0; t <= nt; t++) {
        size_t p = (size_t)((unsigned __int128)size * (unsigned)t / (unsigned)nt);
        if (p > 0 && p < size && buf[p - 1] != '\n') {
            const char *nl = (const char *)memchr(buf + p, '\n', size - p);
            p = nl ? (size_t)(nl - buf) + 1 : size;
        }
        begin[t] = p;
    }

    for (int pass = 0; pass < 2; pass++) {
        uint64_t total = 0;
        if (pass == 1) {
            for (int t = 0; t < nt; t++) {
                const int64_t c = cnt[t];
                cnt[t] = (int64_t)total;
                total += (uint64_t)c;
            }
            if (total > 0xFFFFFFFFull) {
                set_error("%s: %llu edges exceed the 32-bit edge count", path,
                          (unsigned long long)total);
                goto done;
            }
            src = (uint32_t *)malloc((total ? total : 1) * sizeof(uint32_t));
            dst = (uint32_t *)malloc((total ? total : 1) * sizeof(uint32_t));
            wt = weighted ? (float *)malloc((total ? total : 1) * sizeof(float)) : NULL;
            if (!src || !dst || (weighted && !wt)) {
                set_error("%s: out of memory for %llu edges", path, (unsigned long long)total);
                goto done;
            }
        }
        /* pass 0 finds every malformed line; pass 1 re-reads the same bytes */
        if (pass == 0) {
            #pragma omp parallel for num_threads(nt) schedule(static, 1)
            for (int t = 0; t < nt; t++) {
                bad[t] = (size_t)-1;
                cnt[t] = parse_range(buf, begin[t], begin[t + 1], weighted, NULL, NULL, NULL, 0,
                                     &mx[t], &bad[t]);
            }
        } else {
            #pragma omp parallel for num_threads(nt) schedule(static, 1)
            for (int t = 0; t < nt; t++)
                parse_range(buf, begin[t], begin[t + 1], weighted, src, dst, wt,
                            (uint64_t)cnt[t], &mx[t], &bad[t]);
        }
        size_t first = (size_t)-1;
        for (int t = 0; pass == 0 && t < nt && first == (size_t)-1; t++)
            if (cnt[t] < 0)
                first = bad[t];
        if (first != (size_t)-1) {
            /* the first malformed line, with its 1-based number */
            uint64_t line = 1;
            for (size_t i = 0; i < first; i++)
                line += buf[i] == '\n';
            set_error("%s:%llu: expected `src dst%s` (decimal ids below 4294967295)", path,
                      (unsigned long long)line, weighted ? " weight" : "");
            goto done;
        }
        if (pass == 1) {
            uint32_t n = 0;
            for (int t = 0; t < nt; t++) {
                const uint64_t lo = (uint64_t)cnt[t];
                const uint64_t hi = t + 1 < nt ? (uint64_t)cnt[t + 1] : total;
                if (hi > lo && mx[t] + 1 > n)
                    n = mx[t] + 1;
            }
            g = graphCSRFromEdgeList(n, total, src, dst, wt, simplify);
            if (!g)
                set_error("%s: out of memory building the CSR", path);
        }
    }

done:
    if (buf)
        munmap((void *)buf, size);
    free(begin);
    free(cnt);
    free(bad);
    free(mx);
    free(src);
    free(dst);
    free(wt);
    return g;
}

/* ------------------------------------------------------------------ writers */

static int write_all(FILE *f, const void *p, size_t bytes)
{
    return bytes == 0 || fwrite(p, 1, bytes, f) == bytes ? 0 : -1;
}

/* Decimal u32 into `o`, returns the length. */
static int put_u32(char *o, uint32_t v)
{
    char t[10];
    int k = 0;
    do {
        t[k++] = (char)('0' + v % 10);
        v /= 10;
    } while (v);
    for (int i = 0; i < k; i++)
        o[i] = t[k - 1 - i];
    return k;
}

int graphCSRSaveTextEdgeList(const struct GraphCSR *g, const char *path)
{
    if (!g || !path) {
        set_error("null graph or path");
        return -1;
    }
    FILE *f = fopen(path, "wb");
    if (!f) {
        set_error("%s: cannot create (%s)", path, strerror(errno));
        return -1;
    }
    const uint64_t m = g->num_edges;
    const uint32_t *s = g->sorted_edges_array->edges_array_src;
    const uint32_t *d = g->sorted_edges_array->edges_array_dest;
    const float *w = g->sorted_edges_array->edges_array_weight;
    /* per line: two ids (10 digits), a tab, a newline, and "\t%.9g" (at most 16 chars) */
    const size_t per_line = 22 + (w ? 17 : 0);
    const uint64_t step = 1u << 22;          /* edges formatted per round */
    const int nt = omp_get_max_threads();
    char **out = (char **)calloc((size_t)nt, sizeof(char *));
    size_t *len = (size_t *)calloc((size_t)nt, sizeof(size_t));
    int rc = out && len ? 0 : -1;
    const uint64_t per_t = (step + (uint64_t)nt - 1) / (uint64_t)nt;
    for (int t = 0; rc == 0 && t < nt; t++) {
        out[t] = (char *)malloc(per_t * per_line + 1);
        if (!out[t])
            rc = -1;
    }
    for (uint64_t base = 0; rc == 0 && base < m; base += step) {
        const uint64_t end = base + step < m ? base + step : m;
        #pragma omp parallel for num_threads(nt) schedule(static, 1)
        for (int t = 0; t < nt; t++) {
            uint64_t lo = base + per_t * (uint64_t)t, hi = lo + per_t;
            lo = lo < end ? lo : end;
            hi = hi < end ? hi : end;
            char *o = out[t];
            for (uint64_t i = lo; i < hi; i++) {
                o += put_u32(o, s[i]);
                *o++ = '\t';
                o += put_u32(o, d[i]);
                if (w) {
                    char tmp[32];
                    int k = snprintf(tmp, sizeof(tmp), "\t%.9g", (double)w[i]);
                    memcpy(o, tmp, (size_t)k);
                    o += k;
                }
                *o++ = '\n';
            }
            len[t] = (size_t)(o - out[t]);
        }
        for (int t = 0; rc == 0 && t < nt; t++)
            rc = write_all(f, out[t], len[t]);
    }
    if (out)
        for (int t = 0; t < nt; t++)
            free(out[t]);
    free(out);
    free(len);
    if (fclose(f) != 0)
        rc = -1;
    if (rc != 0)
        set_error("%s: write failed", path);
    return rc;
}

int graphCSRSaveBinaryEdgeList(const struct GraphCSR *g, const char *path)
{
    if (!g || !path) {
        set_error("null graph or path");
        return -1;
    }
    FILE *f = fopen(path, "wb");
    if (!f) {
        set_error("%s: cannot create (%s)", path, strerror(errno));
        return -1;
    }
    const uint64_t m = g->num_edges;
    const uint32_t *s = g->sorted_edges_array->edges_array_src;
    const uint32_t *d = g->sorted_edges_array->edges_array_dest;
    const uint64_t step = 1u << 22;
    uint8_t *buf = (uint8_t *)malloc(step * 8);
    int rc = buf ? 0 : -1;
    for (uint64_t base = 0; rc == 0 && base < m; base += step) {
        const uint64_t end = base + step < m ? base + step : m;
        #pragma omp parallel for schedule(static)
        for (uint64_t i = base; i < end; i++) {
            uint8_t *p = buf + (i - base) * 8;
            for (int k = 0; k < 4; k++) {
                p[k] = (uint8_t)(s[i] >> (8 * k));
                p[4 + k] = (uint8_t)(d[i] >> (8 * k));
            }
        }
        rc = write_all(f, buf, (size_t)(end - base) * 8);
    }
    free(buf);
    if (fclose(f) != 0)
        rc = -1;
    if (rc != 0)
        set_error("%s: write failed", path);
    return rc;
}

/* ------------------------------------------------------------------ serialised CSR */

#define SER_VERSION 1u

struct ser_header {
    char magic[8];
    uint32_t version;
    uint32_t flags;
    uint32_t num_vertices;
    uint32_t num_edges;
    float max_weight;
    uint32_t avg_degree;
    float fwd_max_weight;
    float inv_max_weight;
};

static const char ser_magic[8] = {'A', 'G', 'C', 'S', 'R', 0, 0, 0};

/* sum of w[i] * (2 * (base + i) + 1) mod 2^64 */
static uint64_t word_sum(const void *p, size_t words, uint64_t base)
{
    const uint32_t *w = (const uint32_t *)p;
    uint64_t s = 0;
    #pragma omp parallel for schedule(static) reduction(+:s) if (words > (1u << 16))
    for (size_t i = 0; i < words; i++)
        s += (uint64_t)w[i] * (2 * (base + (uint64_t)i) + 1);
    return s;
}

struct ser_array {
    void *p;
    size_t words;
};

/* The serialised arrays in file order; returns their count. */
static int ser_arrays(const struct GraphCSR *g, int weighted, struct ser_array *a)
{
    const size_t n = g->num_vertices, m = g->num_edges;
    const struct Vertex *vs[2] = {g->vertices, g->inverse_vertices};
    const struct EdgeList *es[2] = {g->sorted_edges_array, g->inverse_sorted_edges_array};
    int k = 0;
    for (int v = 0; v < 2; v++) {
        a[k].p = vs[v]->out_degree;        a[k++].words = n;
        a[k].p = vs[v]->in_degree;         a[k++].words = n;
        a[k].p = vs[v]->edges_idx;         a[k++].words = n;
        a[k].p = es[v]->edges_array_src;   a[k++].words = m;
        a[k].p = es[v]->edges_array_dest;  a[k++].words = m;
        if (weighted) {
            a[k].p = es[v]->edges_array_weight;
            a[k++].words = m;
        }
    }
    return k;
}

int graphCSRSaveSerialized(const struct GraphCSR *g, const char *path)
{
    if (!g || !path) {
        set_error("null graph or path");
        return -1;
    }
    const int weighted = g->sorted_edges_array->edges_array_weight != NULL &&
                         g->inverse_sorted_edges_array->edges_array_weight != NULL;
    struct ser_header h;
    memset(&h, 0, sizeof(h));
    memcpy(h.magic, ser_magic, sizeof(h.magic));
    h.version = SER_VERSION;
    h.flags = weighted ? 1u : 0u;
    h.num_vertices = g->num_vertices;
    h.num_edges = g->num_edges;
    h.max_weight = g->max_weight;
    h.avg_degree = g->avg_degree;
    h.fwd_max_weight = g->sorted_edges_array->max_weight;
    h.inv_max_weight = g->inverse_sorted_edges_array->max_weight;

    FILE *f = fopen(path, "wb");
    if (!f) {
        set_error("%s: cannot create (%s)", path, strerror(errno));
        return -1;
    }
    struct ser_array a[12];
    const int na = ser_arrays(g, weighted, a);
    uint64_t word = sizeof(h) / 4;
    uint64_t sum = word_sum(&h, word, 0);
    int rc = write_all(f, &h, sizeof(h));
    for (int i = 0; rc == 0 && i < na; i++) {
        sum += word_sum(a[i].p, a[i].words, word);
        word += a[i].words;
        rc = write_all(f, a[i].p, a[i].words * 4);
    }
    if (rc == 0)
        rc = write_all(f, &sum, sizeof(sum));
    if (fclose(f) != 0)
        rc = -1;
    if (rc != 0)
        set_error("%s: write failed", path);
    return rc;
}

/* Structure of one view: offsets are the prefix sums of the row lengths, every row holds its
 * own id in `src`, destinations are below n and non-decreasing within a row. */
static int check_view(uint32_t n, uint64_t m, const struct Vertex *v, const struct EdgeList *e)
{
    uint64_t acc = 0;
    for (uint32_t r = 0; r < n; r++) {
        if (v->edges_idx[r] != acc)
            return -1;
        acc += v->out_degree[r];
        if (acc > m)
            return -1;
    }
    if (acc != m)
        return -1;
    int bad = 0;
    #pragma omp parallel for schedule(dynamic, 4096) reduction(|:bad)
    for (int64_t r = 0; r < (int64_t)n; r++) {
        const uint64_t b = v->edges_idx[r], d = v->out_degree[r];
        for (uint64_t j = b; j < b + d; j++) {
            if (e->edges_array_src[j] != (uint32_t)r || e->edges_array_dest[j] >= n ||
                (j > b && e->edges_array_dest[j] < e->edges_array_dest[j - 1]))
                bad = 1;
        }
    }
    return bad ? -1 : 0;
}

static int check_graph(const struct GraphCSR *g)
{
    const uint32_t n = g->num_vertices;
    const uint64_t m = g->num_edges;
    if (check_view(n, m, g->vertices, g->sorted_edges_array) != 0 ||
        check_view(n, m, g->inverse_vertices, g->inverse_sorted_edges_array) != 0)
        return -1;
    uint32_t *count = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    if (!count)
        return -2;
    const uint32_t *d = g->sorted_edges_array->edges_array_dest;
    #pragma omp parallel for schedule(static)
    for (uint64_t j = 0; j < m; j++) {
        #pragma omp atomic
        count[d[j]]++;
    }
    int bad = 0;
    #pragma omp parallel for schedule(static) reduction(|:bad)
    for (int64_t r = 0; r < (int64_t)n; r++) {
        if (count[r] != g->inverse_vertices->out_degree[r] ||
            g->vertices->in_degree[r] != g->inverse_vertices->out_degree[r] ||
            g->inverse_vertices->in_degree[r] != g->vertices->out_degree[r])
            bad = 1;
    }
    free(count);
    return bad ? -1 : 0;
}

struct GraphCSR *graphCSRLoadSerialized(const char *path)
{
    FILE *f = fopen(path, "rb");
    if (!f) {
        set_error("%s: cannot open (%s)", path, strerror(errno));
        return NULL;
    }
    struct ser_header h;
    struct GraphCSR *g = NULL;
    if (fread(&h, sizeof(h), 1, f) != 1 || memcmp(h.magic, ser_magic, sizeof(ser_magic)) != 0) {
        set_error("%s: not a serialised CSR file", path);
        goto fail;
    }
    if (h.version != SER_VERSION || (h.flags & ~1u) != 0) {
        set_error("%s: unsupported version %u / flags %#x", path, h.version, h.flags);
        goto fail;
    }
    const int weighted = (int)(h.flags & 1u);
    const uint64_t n = h.num_vertices, m = h.num_edges;
    const uint64_t expect = sizeof(h) + 4 * (6 * n + (uint64_t)(4 + 2 * weighted) * m) + 8;
    struct stat st;
    if (fstat(fileno(f), &st) != 0 || (uint64_t)st.st_size != expect) {
        set_error("%s: size %lld, header implies %llu", path, (long long)st.st_size,
                  (unsigned long long)expect);
        goto fail;
    }
    g = graphCSRAlloc(h.num_vertices, h.num_edges, weighted);
    if (!g) {
        set_error("%s: out of memory for %llu vertices, %llu edges", path,
                  (unsigned long long)n, (unsigned long long)m);
        goto fail;
    }
    struct ser_array a[12];
    const int na = ser_arrays(g, weighted, a);
    uint64_t word = sizeof(h) / 4;
    uint64_t sum = word_sum(&h, word, 0);
    for (int i = 0; i < na; i++) {
        if (a[i].words && fread(a[i].p, 4, a[i].words, f) != a[i].words) {
            set_error("%s: short read", path);
            goto fail;
        }
        sum += word_sum(a[i].p, a[i].words, word);
        word += a[i].words;
    }
    uint64_t stored = 0;
    if (fread(&stored, sizeof(stored), 1, f) != 1 || stored != sum) {
        set_error("%s: checksum mismatch (file damaged)", path);
        goto fail;
    }
    int chk = check_graph(g);
    if (chk != 0) {
        set_error(chk == -2 ? "%s: out of memory checking the graph"
                            : "%s: inconsistent CSR arrays (offsets, ids or the two views)",
                  path);
        goto fail;
    }
    g->max_weight = h.max_weight;
    g->avg_degree = h.avg_degree;
    g->sorted_edges_array->max_weight = h.fwd_max_weight;
    g->inverse_sorted_edges_array->max_weight = h.inv_max_weight;
    fclose(f);
    return g;
fail:
    graphCSRFree(g);
    fclose(f);
    return NULL;
}

/* ------------------------------------------------------------------ by format code */

struct GraphCSR *graphCSRLoadFile(const char *path, int file_format, int simplify)
{
    switch (file_format) {
    case GRAPH_FILE_TEXT_EDGE_LIST:
        return graphCSRLoadTextEdgeList(path, 0, simplify);
    case GRAPH_FILE_BINARY_EDGE_LIST: {
        struct GraphCSR *g = graphCSRLoadBinaryEdgeList(path, simplify);
        if (!g)
            set_error("%s: cannot read a binary edge list (missing, or size not a multiple of "
                      "8 bytes)", path);
        return g;
    }
    case GRAPH_FILE_CSR_SERIALIZED:
        return graphCSRLoadSerialized(path);
    default:
        set_error("file format %d: expected 0 (text edge list), 1 (binary edge list) or 2 "
                  "(serialised CSR)", file_format);
        return NULL;
    }
}

int graphCSRSaveFile(const struct GraphCSR *g, const char *path, int file_format)
{
    switch (file_format) {
    case GRAPH_FILE_TEXT_EDGE_LIST:
        return graphCSRSaveTextEdgeList(g, path);
    case GRAPH_FILE_BINARY_EDGE_LIST:
        return graphCSRSaveBinaryEdgeList(g, path);
    case GRAPH_FILE_CSR_SERIALIZED:
        return graphCSRSaveSerialized(g, path);
    default:
        set_error("file format %d: expected 0, 1 or 2", file_format);
        return -1;
    }
}
