/*
 * oracle.h -- CPU restatement of parasail's serial alignment recurrences.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU path is
 * compared against; it is linked by tests/, by __graft_entry__.smoke() and by
 * the cpu_baseline leg of bench.py, never by the product library
 * (parasail_amd/libparasail_hip.so), which fails loudly instead of falling back
 * to it.
 *
 * Every routine restates one reference function (file:line cited at each
 * definition in oracle.c) in plain C99 int32 arithmetic.  The reference's own
 * sources cannot be compiled here (every src/ C file includes the CMake-generated
 * "config.h"), so this restatement is pinned by the
 * reference's known-answer tests, brute-force enumeration and invariants in
 * tests/test_oracle.py.
 *
 * Conventions (all of them the reference's):
 *   s1 / m  : query  (rows,    DP index i), mapped residue codes, NULL for PSSM
 *   s2 / n  : target (columns, DP index j), mapped residue codes
 *   matrix  : row-major int32 scores, `size` columns; for a PSSM row r is
 *             query position r (parasail.h:147-162)
 *   gaps    : a gap of length k costs open + (k-1)*gap (README.md:55)
 */
#ifndef PARASAIL_ORACLE_H
#define PARASAIL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NEG_INF (INT32_MIN / 2) /* src/sw.c:16 */

/* trace codes, parasail.h:31-41 */
#define ORC_ZERO_MASK 120
#define ORC_INS 1
#define ORC_DEL 2
#define ORC_DIAG 4
#define ORC_DIAG_E 8
#define ORC_INS_E 16
#define ORC_DIAG_F 32
#define ORC_DEL_F 64

enum { ORC_SW = 0, ORC_NW = 1, ORC_SG = 2 };
enum { ORC_SCORE = 0, ORC_STATS = 1, ORC_TRACE = 2 };

typedef struct orc_result {
    int score;
    int end_query;
    int end_ref;
    int matches;
    int similar;
    int length;
} orc_result_t;

typedef struct orc_matrix {
    int size;          /* columns of `matrix` (alphabet + 1)          */
    int length;        /* rows: == size (square) or PSSM length         */
    int type;          /* 0 square, 1 pssm                              */
    int max, min;
    int mapper[256];
    int *matrix;       /* owned, length*size                            */
    char alphabet[257];
} orc_matrix_t;

/* ---- matrices (src/memory.c) ---- */
int  orc_matrix_create(orc_matrix_t *out, const char *alphabet, int match,
                       int mismatch, int case_sensitive);
int  orc_matrix_pssm_create(orc_matrix_t *out, const char *alphabet,
                            const int *values, int length, int case_sensitive);
int  orc_matrix_blosum62(orc_matrix_t *out);
int  orc_matrix_square_from_values(orc_matrix_t *out, const char *alphabet,
                                   const int *values, int size, int max, int min);
int  orc_matrix_convert_square_to_pssm(orc_matrix_t *out,
                                       const orc_matrix_t *sq,
                                       const char *s1, int s1Len);
void orc_matrix_free(orc_matrix_t *m);
void orc_map(const orc_matrix_t *mx, const char *s, int len, int *out);

/* ---- serial DP (src/{sw,nw,sg}{,_stats,_trace}.c) ----
 * alg: ORC_SW/NW/SG; kind: ORC_SCORE/STATS/TRACE; sg flags used for SG only.
 * trace: int8 row-major m*n table (LOC_NOVEC, cigar.c:156) when kind==TRACE.
 * Returns 0 on success, -1 on invalid arguments (the reference returns NULL).
 */
int orc_align(int alg, int kind, int s1_beg, int s1_end, int s2_beg, int s2_end,
              const int *s1, int m, const int *s2, int n, int open, int gap,
              const orc_matrix_t *mx, orc_result_t *res, int8_t *trace);
/* orc_align plus the DP tables: tabs[0] every cell's H, tabs[1..3] (stats)
 * its matches, similar, length; m*n ints each, row-major */
int orc_align_tables(int alg, int kind, int s1_beg, int s1_end, int s2_beg, int s2_end,
                     const int *s1, int m, const int *s2, int n, int open, int gap,
                     const orc_matrix_t *mx, orc_result_t *res, int8_t *trace, int **tabs);

/* src/nw_banded.c:29-189 */
int orc_nw_banded(const int *s1, int m, const int *s2, int n, int open, int gap,
                  int k, const orc_matrix_t *mx, orc_result_t *res);

/* ---- CIGAR (src/cigar.c, src/cigar_template.c) ----
 * Walks a row-major int8 trace (LOC_NOVEC) or a striped table of `bits`
 * width with `lanes` segments (LOC_STRIPED).  Writes up to cap ops into seq,
 * returns the op count, or -1 where the reference returns NULL.
 */
int orc_cigar(const char *seqA, int lena, const char *seqB, int lenb,
              int is_sg, int end_query, int end_ref, const void *trace,
              int striped, int bits, int lanes, int case_sensitive,
              const char *aliases, uint32_t *seq, int cap, int *beg_query,
              int *beg_ref);

/* striped trace -> row-major (src/striped_unwind.c:5-26) */
void orc_striped_unwind(const void *striped, int bits, int lanes, int m, int n,
                        int *out);

/* ---- CPU baseline harness (tests/test_openmp.c:383-441 pattern) ----
 * Aligns npairs (s1[k], s2[k]) with the serial oracle, OpenMP over pairs.
 * Sequences are concatenated residue codes with offsets.  Returns seconds.
 */
double orc_batch(int alg, int kind, int s1_beg, int s1_end, int s2_beg,
                 int s2_end, int npairs, const int *q, const int64_t *qoff,
                 const int *qlen, const int *t, const int64_t *toff,
                 const int *tlen, int open, int gap, const orc_matrix_t *mx,
                 orc_result_t *out, int threads);

/* ---- AVX2 striped baseline (striped.c) ----
 * Score-only SW/NW/SG through the reference's dispatched `_striped_sat`
 * chain restated: Farrar striped query profile in AVX2, 8-bit lanes, then
 * 16-bit, then the serial recurrence.  Square matrices only.
 * orc_striped_chain starts at `start` bits (8, or 16 as the _16 functions
 * do) and returns the width that held (8, 16, 32) or -1 (no AVX2, bad
 * arguments); orc_striped_sat is the chain from 8 bits; orc_striped_batch
 * returns the wall seconds (-1 on failure). */
int orc_striped_available(void);
int orc_striped_chain(int alg, int s1_beg, int s1_end, int s2_beg, int s2_end, const int *q,
                      int m, const int *t, int n, int open, int gap, const orc_matrix_t *mx,
                      orc_result_t *out, int start);
int orc_striped_sat(int alg, int s1_beg, int s1_end, int s2_beg, int s2_end, const int *q,
                    int m, const int *t, int n, int open, int gap, const orc_matrix_t *mx,
                    orc_result_t *out);
double orc_striped_batch(int alg, int s1_beg, int s1_end, int s2_beg, int s2_end, int npairs,
                         const int *q, const int64_t *qoff, const int *qlen, const int *t,
                         const int64_t *toff, const int *tlen, int open, int gap,
                         const orc_matrix_t *mx, orc_result_t *out, int *width, int threads,
                         int start);

#ifdef __cplusplus
}
#endif
#endif
