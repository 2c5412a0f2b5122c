/* Hooks linked into the reference harness (test infrastructure only).
 *
 * __wrap_glsc3_ is reached through `-Wl,--wrap=glsc3_`: every call the
 * reference makes to glsc3 (src/math.f:603-617) from another object file
 * (src/cg.f:42,52,59,66) goes through here, is forwarded unchanged to the
 * reference's own glsc3, and its result is appended to the file named by
 * $HK_GLSC3 (raw doubles).  This records the full scalar history of a CG
 * solve (rnorm^2, then rtz1, pap, rtr per iteration) without touching the
 * reference's sources.
 */
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

double __real_glsc3_(double *a, double *b, double *m, int *n);

static FILE *g_rec;
static int g_rec_tried;

double __wrap_glsc3_(double *a, double *b, double *m, int *n)
{
    double v = __real_glsc3_(a, b, m, n);
    if (!g_rec_tried) {
        const char *path = getenv("HK_GLSC3");
        g_rec_tried = 1;
        if (path && *path) g_rec = fopen(path, "wb");
    }
    if (g_rec) fwrite(&v, sizeof v, 1, g_rec);
    return v;
}

void hk_flush_(void)
{
    if (g_rec) fflush(g_rec);
}

double hk_clock_(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}
