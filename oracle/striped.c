/* striped.c -- CPU baseline: the reference's dispatched AVX2 path for
 * score-only SW, NW and SG, restated (Farrar striped query profile, one
 * profile built per call, 8-bit lanes first and 16-bit on saturation, then
 * 32-bit through the serial recurrence: the `_striped_sat` chain of
 * src/satcheck.c:41520-41537 over src/{sw,sg,nw}_striped_avx2_256_{8,16}.c).
 *
 * TEST INFRASTRUCTURE ONLY: bench.py times it on the host cores as the
 * reported CPU baseline and tests/ check it against orc_align.  The
 * product library never links it.  Needs AVX2 at run time; without it the
 * entry points return -1. */
#define _POSIX_C_SOURCE 199309L
#define _ISOC11_SOURCE
#include <immintrin.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

#define E_BITS 8
#include "striped_tmpl.h"
#undef E_BITS
#define E_BITS 16
#include "striped_tmpl.h"
#undef E_BITS

int orc_striped_available(void) { return __builtin_cpu_supports("avx2") ? 1 : 0; }

/* One pair through the chain from `start` bits (8, or 16 for the _16
 * functions); returns the width that held (8, 16, 32) or -1.  s1_beg.. are
 * the SG end flags (ignored for SW and NW). */
int orc_striped_chain(int alg, int s1_beg, int s1_end, int s2_beg, int s2_end, const int *q,
                      int m, const int *t, int n, int open, int gap, const orc_matrix_t *mx,
                      orc_result_t *out, int start)
{
    if (!orc_striped_available() || m <= 0 || n <= 0 || mx->type != 0) return -1;
    memset(out, 0, sizeof(*out));
    int rc;
    if (start <= 8) {
        rc = st_align8(alg, s1_beg, s1_end, s2_beg, s2_end, q, m, t, n, open, gap, mx, out);
        if (rc == 0) return 8;
        if (rc < 0) return -1;
    }
    rc = st_align16(alg, s1_beg, s1_end, s2_beg, s2_end, q, m, t, n, open, gap, mx, out);
    if (rc == 0) return 16;
    if (rc < 0) return -1;
    if (orc_align(alg, ORC_SCORE, s1_beg, s1_end, s2_beg, s2_end, q, m, t, n, open, gap, mx, out,
                  NULL) != 0)
        return -1;
    return 32;
}

int orc_striped_sat(int alg, int s1_beg, int s1_end, int s2_beg, int s2_end, const int *q,
                    int m, const int *t, int n, int open, int gap, const orc_matrix_t *mx,
                    orc_result_t *out)
{
    return orc_striped_chain(alg, s1_beg, s1_end, s2_beg, s2_end, q, m, t, n, open, gap, mx, out,
                             8);
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

/* Every pair through orc_striped_chain from `start` bits, OpenMP over pairs with guided
 * scheduling (tests/test_openmp.c:353-369); width[p] (may be NULL) gets
 * the width that held.  Returns the wall seconds, or -1 when a pair
 * failed. */
double orc_striped_batch(int alg, int s1_beg, int s1_end, int s2_beg, int s2_end, int npairs,
                         const int *q, const int64_t *qoff, const int *qlen, const int *t,
                         const int64_t *toff, const int *tlen, int open, int gap,
                         const orc_matrix_t *mx, orc_result_t *out, int *width, int threads,
                         int start)
{
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    int failed = 0;
    const double t0 = now_s();
#ifdef _OPENMP
#pragma omp parallel for schedule(guided) reduction(| : failed)
#endif
    for (int p = 0; p < npairs; ++p) {
        const int w = orc_striped_chain(alg, s1_beg, s1_end, s2_beg, s2_end, q + qoff[p],
                                        qlen[p], t + toff[p], tlen[p], open, gap, mx, &out[p],
                                        start);
        if (width) width[p] = w;
        failed |= w < 0;
    }
    const double secs = now_s() - t0;
    return failed ? -1.0 : secs;
}
