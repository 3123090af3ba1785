/* percall_omp.c -- times a per-pair profile function the way parasail's own
 * OpenMP driver calls it (tests/test_openmp.c:353-369): one profile of the
 * query, then every database sequence aligned in a `parallel for` with
 * guided scheduling, each result freed at once.  bench.py runs it for the
 * per-call figures of its JSON line; it is no part of the library.
 *
 *   percall_omp FUNCTION QUERIES.fa DATABASE.fa OPEN GAP THREADS...
 *
 * FUNCTION is a pfunction name (e.g. sw_striped_profile_hip_64_16); the
 * first record of QUERIES.fa is the query.  For each thread count it prints
 * one line: threads, best of three wall times (s), GCUPS, and the number of
 * saturated results. */
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../../include/parasail_hip.h"

int main(int argc, char **argv)
{
    if (argc < 7) {
        fprintf(stderr, "usage: %s FUNCTION QUERIES.fa DATABASE.fa OPEN GAP THREADS...\n", argv[0]);
        return 2;
    }
    parasail_pfunction_t *pf = parasail_lookup_pfunction(argv[1]);
    parasail_pcreator_t *pc = parasail_lookup_pcreator(argv[1]);
    parasail_sequences_t *qs = parasail_sequences_from_file(argv[2]);
    parasail_sequences_t *db = parasail_sequences_from_file(argv[3]);
    const parasail_matrix_t *mx = parasail_matrix_lookup("blosum62");
    if (!pf || !pc || !qs || !db || qs->l == 0 || db->l == 0 || !mx) {
        fprintf(stderr, "%s: bad function, sequence files or matrix\n", argv[0]);
        return 1;
    }
    const int open = atoi(argv[4]), gap = atoi(argv[5]);
    const char *q = qs->seqs[0].seq.s;
    const int m = (int)qs->seqs[0].seq.l;
    parasail_profile_t *prof = pc(q, m, mx);
    if (!prof) {
        fprintf(stderr, "%s: profile creation failed\n", argv[0]);
        return 1;
    }
    const long n = (long)db->l;
    double cells = 0;
    for (long k = 0; k < n; ++k) cells += (double)m * db->seqs[k].seq.l;
    for (long k = 0; k < n && k < 20; ++k)
        parasail_result_free(pf(prof, db->seqs[k].seq.s, (int)db->seqs[k].seq.l, open, gap));
    for (int a = 6; a < argc; ++a) {
        const int threads = atoi(argv[a]);
        omp_set_num_threads(threads > 0 ? threads : 1);
        double best = 1e30;
        long sat = 0;
        for (int rep = 0; rep < 3; ++rep) {
            long s = 0, bad = 0;
            const double t0 = omp_get_wtime();
#pragma omp parallel for schedule(guided) reduction(+ : s, bad)
            for (long k = 0; k < n; ++k) {
                parasail_result_t *r = pf(prof, db->seqs[k].seq.s, (int)db->seqs[k].seq.l, open, gap);
                if (!r) {
                    ++bad;
                    continue;
                }
                s += parasail_result_is_saturated(r);
                parasail_result_free(r);
            }
            const double dt = omp_get_wtime() - t0;
            if (bad) {
                fprintf(stderr, "%s: %ld calls failed\n", argv[0], bad);
                return 1;
            }
            if (dt < best) best = dt;
            sat = s;
        }
        printf("%d %.6f %.4f %ld\n", threads, best, cells / best / 1e9, sat);
        fflush(stdout);
    }
    parasail_profile_free(prof);
    parasail_sequences_free(qs);
    parasail_sequences_free(db);
    return 0;
}
