/* Graph builders and file readers under AddressSanitizer + UndefinedBehaviorSanitizer (host
 * code only). Exits non-zero on any unexpected accept or reject; the sanitizers abort on any
 * memory or undefined-behaviour error. argv[1]: the directory of graphbrew_graph.bin,
 * argv[2]: a scratch directory. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "graphCSR.h"

static char scratch[4000];
static int failures;

static const char *tmp(const char *name)
{
    static char path[4096];
    snprintf(path, sizeof path, "%s/%s", scratch, name);
    return path;
}

static void write_bytes(const char *path, const void *data, size_t len)
{
    FILE *f = fopen(path, "wb");
    if (!f || fwrite(data, 1, len, f) != len) {
        fprintf(stderr, "cannot write %s\n", path);
        exit(2);
    }
    fclose(f);
}

/* expect: 1 accepted (with n, m), 0 rejected */
static void text_case(const char *text, int weighted, int expect, uint32_t n, uint32_t m)
{
    const char *p = tmp("case.txt");
    write_bytes(p, text, strlen(text));
    struct GraphCSR *g = graphCSRLoadTextEdgeList(p, weighted, 1);
    const int ok = g != NULL;
    if (ok != expect || (ok && (g->num_vertices != n || g->num_edges != m))) {
        printf("text case %s(weighted=%d): got %s\n", text, weighted,
               ok ? "accepted" : graphIOLastError());
        failures++;
    }
    graphCSRFree(g);
}

int main(int argc, char **argv)
{
    if (argc < 3)
        return 2;
    snprintf(scratch, sizeof scratch, "%s", argv[2]);
    char path[4096];
    snprintf(path, sizeof path, "%s/graphbrew_graph.bin", argv[1]);
    struct GraphCSR *g = graphCSRLoadBinaryEdgeList(path, 1);
    if (!g)
        return 3;
    if (graphCSRSaveTextEdgeList(g, tmp("rt.txt")) || graphCSRSaveBinaryEdgeList(g, tmp("rt.bin")) ||
        graphCSRSaveSerialized(g, tmp("rt.csr")))
        return 4;
    struct GraphCSR *h = graphCSRLoadSerialized(tmp("rt.csr"));
    struct GraphCSR *t = graphCSRLoadTextEdgeList(tmp("rt.txt"), 0, 1);
    struct GraphCSR *b = graphCSRLoadBinaryEdgeList(tmp("rt.bin"), 1);
    if (!h || !t || !b || h->num_edges != g->num_edges || t->num_edges != g->num_edges ||
        b->num_edges != g->num_edges)
        return 5;
    graphCSRFree(h);
    graphCSRFree(t);
    graphCSRFree(b);
    graphCSRFree(g);

    text_case("", 0, 1, 0, 0);
    text_case("\n\n", 0, 1, 0, 0);
    text_case("# c\n% c\n", 1, 1, 0, 0);
    text_case("0 1\n1 2\n2 0\n", 0, 1, 3, 3);
    text_case("0 1\n1 2\n2 0\n", 1, 0, 0, 0);          /* weighted needs a third column */
    text_case("0 1", 0, 1, 2, 1);                      /* no final newline */
    text_case("0 1\n1\n", 0, 0, 0, 0);
    text_case("0 x\n", 0, 0, 0, 0);
    text_case("0 1 2\n", 0, 1, 2, 1);                  /* further columns ignored */
    text_case("0,1\n1\t2\n", 0, 1, 3, 2);
    text_case("4294967295 0\n", 0, 0, 0, 0);
    text_case("-1 2\n", 0, 0, 0, 0);
    text_case("0 1 0.5\n1 2 0.25\n", 1, 1, 3, 2);
    text_case("0 1 0.5\n1 2 1e40\n", 1, 0, 0, 0);      /* weight beyond float */
    text_case("0 1 nan\n", 1, 0, 0, 0);
    text_case("   0    1   \r\n1 0\r\n", 0, 1, 2, 2);
    text_case("99999999999 1\n", 0, 0, 0, 0);

    /* every truncation and a sweep of single-byte corruptions of a serialised file */
    g = graphCSRGenerateRMAT(8, 8, 0.57, 0.19, 0.19, 7u, 1);
    if (graphCSRSaveSerialized(g, tmp("s.csr")))
        return 6;
    FILE *f = fopen(tmp("s.csr"), "rb");
    fseek(f, 0, SEEK_END);
    const long len = ftell(f);
    fseek(f, 0, SEEK_SET);
    unsigned char *buf = (unsigned char *)malloc((size_t)len);
    if (fread(buf, 1, (size_t)len, f) != (size_t)len)
        return 7;
    fclose(f);
    for (long cut = 0; cut < len; cut += len / 61 + 1) {
        write_bytes(tmp("t.csr"), buf, (size_t)cut);
        struct GraphCSR *x = graphCSRLoadSerialized(tmp("t.csr"));
        if (x) {
            printf("truncated file of %ld bytes accepted\n", cut);
            failures++;
            graphCSRFree(x);
        }
    }
    for (long pos = 0; pos < len; pos += len / 97 + 1) {
        buf[pos] ^= 0x5A;
        write_bytes(tmp("t.csr"), buf, (size_t)len);
        struct GraphCSR *x = graphCSRLoadSerialized(tmp("t.csr"));
        if (x) {
            printf("corrupted byte %ld accepted\n", pos);
            failures++;
            graphCSRFree(x);
        }
        buf[pos] ^= 0x5A;
    }
    free(buf);
    graphCSRFree(g);
    printf("failures %d\n", failures);
    return failures ? 1 : 0;
}
