/*
 * parasail_hip.h -- the "hip" ISA family of parasail for AMD Instinct
 * MI355X (gfx950), as a C ABI.
 *
 * Drop-in boundary.  Every entry point below has the exact signature of the
 * reference interface it replaces:
 *   parasail_function_t   parasail.h:184-188  (s1, s1Len, s2, s2Len, open, gap, matrix)
 *   parasail_pfunction_t  parasail.h:206-209  (profile, s2, s2Len, open, gap)
 *   parasail_pcreator_t   parasail.h:211-213  (s1, s1Len, matrix)
 *   parasail_nw_banded    parasail.h:293-297
 * and returns the reference's own object types (parasail_result_t,
 * parasail_cigar_t, parasail_profile_t, parasail_matrix_t; parasail.h:85-230)
 * allocated with the reference's allocators, so parasail_result_free() of
 * either library releases them.  Names follow the reference grammar
 * alg[_stats|_trace]_{striped|scan}[_profile]_{isa}_{bits}_{width}
 * (README.md:246-281) with isa "hip" and bits "64" (one wave64).
 *
 * The library also exports the data-model half of parasail's API (result
 * getters, matrices, profiles, CIGAR, function lookup) with the reference's
 * names and semantics, so a program that uses only the alignment path links
 * against this library alone.
 */
#ifndef PARASAIL_HIP_H
#define PARASAIL_HIP_H

#include <stdint.h>
#include <stdio.h>
#include <sys/types.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef _PARASAIL_H_
/* ---- types and flags, ABI-identical to parasail.h:31-230 ---- */
#define PARASAIL_VERSION_MAJOR 2
#define PARASAIL_VERSION_MINOR 6
#define PARASAIL_VERSION_PATCH 2

#define PARASAIL_ZERO_MASK 120
#define PARASAIL_E_MASK 103
#define PARASAIL_F_MASK 31
#define PARASAIL_ZERO 0
#define PARASAIL_INS 1
#define PARASAIL_DEL 2
#define PARASAIL_DIAG 4
#define PARASAIL_DIAG_E 8
#define PARASAIL_INS_E 16
#define PARASAIL_DIAG_F 32
#define PARASAIL_DEL_F 64

#define PARASAIL_FLAG_NW (1 << 0)
#define PARASAIL_FLAG_SG (1 << 1)
#define PARASAIL_FLAG_SW (1 << 2)
#define PARASAIL_FLAG_SG_S1_BEG (1 << 3)
#define PARASAIL_FLAG_SG_S1_END (1 << 4)
#define PARASAIL_FLAG_SATURATED (1 << 6)
#define PARASAIL_FLAG_BANDED (1 << 7)
#define PARASAIL_FLAG_NOVEC (1 << 8)
#define PARASAIL_FLAG_NOVEC_SCAN (1 << 9)
#define PARASAIL_FLAG_SCAN (1 << 10)
#define PARASAIL_FLAG_STRIPED (1 << 11)
#define PARASAIL_FLAG_DIAG (1 << 12)
#define PARASAIL_FLAG_BLOCKED (1 << 13)
#define PARASAIL_FLAG_SG_S2_BEG (1 << 14)
#define PARASAIL_FLAG_SG_S2_END (1 << 15)
#define PARASAIL_FLAG_STATS (1 << 16)
#define PARASAIL_FLAG_TABLE (1 << 17)
#define PARASAIL_FLAG_ROWCOL (1 << 18)
#define PARASAIL_FLAG_TRACE (1 << 19)
#define PARASAIL_FLAG_BITS_8 (1 << 20)
#define PARASAIL_FLAG_BITS_16 (1 << 21)
#define PARASAIL_FLAG_BITS_32 (1 << 22)
#define PARASAIL_FLAG_BITS_64 (1 << 23)
#define PARASAIL_FLAG_LANES_1 (1 << 24)
#define PARASAIL_FLAG_LANES_2 (1 << 25)
#define PARASAIL_FLAG_LANES_4 (1 << 26)
#define PARASAIL_FLAG_LANES_8 (1 << 27)
#define PARASAIL_FLAG_LANES_16 (1 << 28)
#define PARASAIL_FLAG_LANES_32 (1 << 29)
#define PARASAIL_FLAG_LANES_64 (1 << 30)
#define PARASAIL_FLAG_INVALID 0x80000020

#define PARASAIL_MATRIX_TYPE_SQUARE 0
#define PARASAIL_MATRIX_TYPE_PSSM 1

typedef struct parasail_result_extra_stats_tables {
    int *score_table, *matches_table, *similar_table, *length_table;
} parasail_result_extra_stats_tables_t;

typedef struct parasail_result_extra_stats_rowcols {
    int *score_row, *matches_row, *similar_row, *length_row;
    int *score_col, *matches_col, *similar_col, *length_col;
} parasail_result_extra_stats_rowcols_t;

typedef struct parasail_result_extra_stats {
    int matches;
    int similar;
    int length;
    union {
        void *extra;
        parasail_result_extra_stats_tables_t *tables;
        parasail_result_extra_stats_rowcols_t *rowcols;
    };
} parasail_result_extra_stats_t;

typedef struct parasail_result_extra_tables {
    int *score_table;
} parasail_result_extra_tables_t;

typedef struct parasail_result_extra_rowcols {
    int *score_row;
    int *score_col;
} parasail_result_extra_rowcols_t;

typedef struct parasail_result_extra_trace {
    void *trace_table;
    void *trace_ins_table;
    void *trace_del_table;
} parasail_result_extra_trace_t;

typedef struct parasail_result {
    int score;
    int end_query;
    int end_ref;
    int flag;
    union {
        void *extra;
        parasail_result_extra_stats_t *stats;
        parasail_result_extra_tables_t *tables;
        parasail_result_extra_rowcols_t *rowcols;
        parasail_result_extra_trace_t *trace;
    };
} parasail_result_t;

typedef struct parasail_matrix {
    const char *name;
    const int *matrix;
    const int *mapper;
    int size;
    int max;
    int min;
    int *user_matrix;
    int type;
    int length;
    const char *alphabet;
    const char *query;
} parasail_matrix_t;

typedef struct parasail_profile_data {
    void *score;
    void *matches;
    void *similar;
} parasail_profile_data_t;

typedef struct parasail_profile {
    const char *s1;
    int s1Len;
    const parasail_matrix_t *matrix;
    struct parasail_profile_data profile8;
    struct parasail_profile_data profile16;
    struct parasail_profile_data profile32;
    struct parasail_profile_data profile64;
    void (*free)(void *profile);
    int stop;
} parasail_profile_t;

typedef parasail_result_t *parasail_function_t(const char *const s1, const int s1Len,
                                               const char *const s2, const int s2Len,
                                               const int open, const int gap,
                                               const parasail_matrix_t *matrix);

typedef struct parasail_function_info {
    parasail_function_t *pointer;
    const char *name;
    const char *alg;
    const char *type;
    const char *isa;
    const char *bits;
    const char *width;
    int lanes;
    char is_table;
    char is_rowcol;
    char is_trace;
    char is_stats;
    char is_ref;
} parasail_function_info_t;

typedef parasail_result_t *parasail_pfunction_t(const parasail_profile_t *const profile,
                                                const char *const s2, const int s2Len,
                                                const int open, const int gap);

typedef parasail_profile_t *parasail_pcreator_t(const char *const s1, const int s1Len,
                                                const parasail_matrix_t *matrix);

typedef struct parasail_pfunction_info {
    parasail_pfunction_t *pointer;
    parasail_pcreator_t *creator;
    const char *name;
    const char *alg;
    const char *type;
    const char *isa;
    const char *bits;
    const char *width;
    int lanes;
    char is_table;
    char is_rowcol;
    char is_trace;
    char is_stats;
    char is_ref;
} parasail_pfunction_info_t;

typedef struct parasail_cigar_ {
    uint32_t *seq;
    int len;
    int beg_query;
    int beg_ref;
} parasail_cigar_t;

/* ---- data model (reference names; parasail.h:232-518, src/memory.c) ---- */
void parasail_version(int *major, int *minor, int *patch);
double parasail_time(void);
void parasail_result_free(parasail_result_t *result);
void parasail_profile_free(parasail_profile_t *profile);
parasail_function_t *parasail_lookup_function(const char *funcname);
parasail_pfunction_t *parasail_lookup_pfunction(const char *funcname);
parasail_pcreator_t *parasail_lookup_pcreator(const char *funcname);
const parasail_function_info_t *parasail_lookup_function_info(const char *funcname);
const parasail_pfunction_info_t *parasail_lookup_pfunction_info(const char *funcname);
const parasail_matrix_t *parasail_matrix_lookup(const char *matrixname);
parasail_matrix_t *parasail_matrix_create(const char *alphabet, const int match, const int mismatch);
parasail_matrix_t *parasail_matrix_create_case_sensitive(const char *alphabet, const int match,
                                                         const int mismatch);
parasail_matrix_t *parasail_matrix_pssm_create(const char *alphabet, const int *values,
                                               const int length);
parasail_matrix_t *parasail_matrix_pssm_create_case_sensitive(const char *alphabet,
                                                              const int *values, const int length);
void parasail_matrix_free(parasail_matrix_t *matrix);
parasail_matrix_t *parasail_matrix_copy(const parasail_matrix_t *original);
parasail_matrix_t *parasail_matrix_convert_square_to_pssm(const parasail_matrix_t *original,
                                                          const char *seq, int len);
void parasail_matrix_set_value(parasail_matrix_t *matrix, int row, int col, int value);

uint32_t parasail_cigar_encode(uint32_t length, char op_letter);
parasail_cigar_t *parasail_cigar_encode_string(const char *cigar);
char parasail_cigar_decode_op(uint32_t cigar_int);
uint32_t parasail_cigar_decode_len(uint32_t cigar_int);
char *parasail_cigar_decode(parasail_cigar_t *cigar);
parasail_cigar_t *parasail_result_get_cigar(parasail_result_t *result, const char *seqA, int lena,
                                            const char *seqB, int lenb,
                                            const parasail_matrix_t *matrix);
parasail_cigar_t *parasail_result_get_cigar_extra(parasail_result_t *result, const char *seqA,
                                                  int lena, const char *seqB, int lenb,
                                                  const parasail_matrix_t *matrix,
                                                  int case_sensitive,
                                                  const char *alphabet_aliases);
void parasail_cigar_free(parasail_cigar_t *cigar);

int parasail_result_is_nw(const parasail_result_t *const result);
int parasail_result_is_sg(const parasail_result_t *const result);
int parasail_result_is_sw(const parasail_result_t *const result);
int parasail_result_is_saturated(const parasail_result_t *const result);
int parasail_result_is_banded(const parasail_result_t *const result);
int parasail_result_is_scan(const parasail_result_t *const result);
int parasail_result_is_striped(const parasail_result_t *const result);
int parasail_result_is_diag(const parasail_result_t *const result);
int parasail_result_is_blocked(const parasail_result_t *const result);
int parasail_result_is_stats(const parasail_result_t *const result);
int parasail_result_is_stats_table(const parasail_result_t *const result);
int parasail_result_is_stats_rowcol(const parasail_result_t *const result);
int parasail_result_is_table(const parasail_result_t *const result);
int parasail_result_is_rowcol(const parasail_result_t *const result);
int parasail_result_is_trace(const parasail_result_t *const result);
int parasail_result_get_score(const parasail_result_t *const result);
int parasail_result_get_end_query(const parasail_result_t *const result);
int parasail_result_get_end_ref(const parasail_result_t *const result);
int parasail_result_get_matches(const parasail_result_t *const result);
int parasail_result_get_similar(const parasail_result_t *const result);
int parasail_result_get_length(const parasail_result_t *const result);
int *parasail_result_get_trace_table(const parasail_result_t *const result);
int *parasail_result_get_trace_ins_table(const parasail_result_t *const result);
int *parasail_result_get_trace_del_table(const parasail_result_t *const result);
int *parasail_result_get_score_table(const parasail_result_t *const result);
int *parasail_result_get_matches_table(const parasail_result_t *const result);
int *parasail_result_get_similar_table(const parasail_result_t *const result);
int *parasail_result_get_length_table(const parasail_result_t *const result);
int *parasail_result_get_score_row(const parasail_result_t *const result);
int *parasail_result_get_matches_row(const parasail_result_t *const result);
int *parasail_result_get_similar_row(const parasail_result_t *const result);
int *parasail_result_get_length_row(const parasail_result_t *const result);
int *parasail_result_get_score_col(const parasail_result_t *const result);
int *parasail_result_get_matches_col(const parasail_result_t *const result);
int *parasail_result_get_similar_col(const parasail_result_t *const result);
int *parasail_result_get_length_col(const parasail_result_t *const result);

/* alignment strings of a trace result and their printed form
 * (parasail.h:299-373, src/traceback.c:121-285) */
typedef struct parasail_traceback_ {
    char *query;
    char *comp;
    char *ref;
} parasail_traceback_t;
parasail_traceback_t *parasail_result_get_traceback(parasail_result_t *result, const char *seqA,
                                                    int lena, const char *seqB, int lenb,
                                                    const parasail_matrix_t *matrix, char match,
                                                    char pos, char neg);
parasail_traceback_t *parasail_result_get_traceback_extra(
    parasail_result_t *result, const char *seqA, int lena, const char *seqB, int lenb,
    const parasail_matrix_t *matrix, char match, char pos, char neg, int case_sensitive,
    const char *alphabet_aliases);
void parasail_traceback_free(parasail_traceback_t *traceback);
void parasail_traceback_generic(const char *seqA, int lena, const char *seqB, int lenb,
                                const char *nameA, const char *nameB,
                                const parasail_matrix_t *matrix, parasail_result_t *result,
                                char match, char pos, char neg, int width, int name_width,
                                int use_stats);
void parasail_traceback_generic_extra(const char *seqA, int lena, const char *seqB, int lenb,
                                      const char *nameA, const char *nameB,
                                      const parasail_matrix_t *matrix, parasail_result_t *result,
                                      char match, char pos, char neg, int width, int name_width,
                                      int use_stats, int int_width, FILE *stream);
void parasail_traceback_generic_extra2(const char *seqA, int lena, const char *seqB, int lenb,
                                       const char *nameA, const char *nameB,
                                       const parasail_matrix_t *matrix, parasail_result_t *result,
                                       char match, char pos, char neg, int width, int name_width,
                                       int use_stats, int int_width, FILE *stream,
                                       int case_sensitive, const char *alphabet_aliases);

/* the SSW library's interface emulated on the hip functions
 * (parasail.h:451-476, src/pssw.c:16-192): best local alignment, its
 * begin from a pass over the reversed prefixes, its CIGAR from a
 * traceback of the aligned stretch; 8-bit first, then 16 */
typedef struct parasail_result_ssw {
    uint16_t score1;
    int32_t ref_begin1;
    int32_t ref_end1;
    int32_t read_begin1;
    int32_t read_end1;
    uint32_t *cigar;
    int32_t cigarLen;
} parasail_result_ssw_t;
parasail_result_ssw_t *parasail_ssw(const char *const s1, const int s1Len, const char *const s2,
                                    const int s2Len, const int open, const int gap,
                                    const parasail_matrix_t *matrix);
parasail_result_ssw_t *parasail_ssw_profile(const parasail_profile_t *const profile,
                                            const char *const s2, const int s2Len,
                                            const int open, const int gap);
parasail_profile_t *parasail_ssw_init(const char *const s1, const int s1Len,
                                      const parasail_matrix_t *matrix, const int8_t score_size);
void parasail_result_ssw_free(parasail_result_ssw_t *result);

/* striped or scan table -> row-major int table (src/striped_unwind.c:
 * 5-26, :56-111; parasail/memory.h:137-141): `array` is the result's trace
 * table (or any table of the result's layout) */
int *parasail_striped_unwind(int lena, int lenb, parasail_result_t *result, void *array);
#endif /* _PARASAIL_H_ */

/* ---- the hip family (drop-ins) ----
 *
 * For every ALG in {sw, nw, sg, sg_qb, sg_qe, sg_qx, sg_db, sg_de, sg_dx,
 * sg_qb_de, sg_qe_db, sg_qb_db, sg_qe_de} (src/sg_helper.h:41-87 maps the
 * sg_* names to the four end-gap flags), KIND in {"", _stats, _trace},
 * STRATEGY in {striped, scan} and WIDTH in {8, 16, 32, sat}:
 *
 *   parasail_function_t  parasail_<ALG><KIND>_<STRATEGY>_hip_64_<WIDTH>
 *       replaces parasail_<ALG><KIND>_<STRATEGY>_<isa>_<bits>_<WIDTH>,
 *       e.g. src/sw_striped_avx2_256_16.c:140-422
 *   parasail_pfunction_t parasail_<ALG><KIND>_<STRATEGY>_profile_hip_64_<WIDTH>
 *       replaces ..._profile_avx2_256_<WIDTH> (parasail.h:206-209)
 *
 * Widths: every result is computed exactly; a width's window decides only
 * whether the result is reported as saturated (PARASAIL_FLAG_SATURATED with
 * the reference's saturated outputs) and which BITS_* flag it carries.
 * "sat" reports the narrowest width whose window holds the alignment
 * (src/satcheck.c:41520-41537).  Unavailable device: NULL and errno=ENOSYS
 * (util/isastubs.py:23-31).
 */
#define PARASAIL_HIP_ALGS(X)                                                  \
    X(sw) X(nw) X(sg) X(sg_qb) X(sg_qe) X(sg_qx) X(sg_db) X(sg_de) X(sg_dx)     \
    X(sg_qb_de) X(sg_qe_db) X(sg_qb_db) X(sg_qe_de)

#define PARASAIL_HIP_DECL_W(alg, kind, strat, w)                              \
    parasail_result_t *parasail_##alg##kind##_##strat##_hip_64_##w(             \
        const char *const s1, const int s1Len, const char *const s2,           \
        const int s2Len, const int open, const int gap,                        \
        const parasail_matrix_t *matrix);                                      \
    parasail_result_t *parasail_##alg##kind##_##strat##_profile_hip_64_##w(     \
        const parasail_profile_t *const profile, const char *const s2,         \
        const int s2Len, const int open, const int gap);

#define PARASAIL_HIP_DECL_S(alg, kind, strat)                                 \
    PARASAIL_HIP_DECL_W(alg, kind, strat, 8)                                  \
    PARASAIL_HIP_DECL_W(alg, kind, strat, 16)                                 \
    PARASAIL_HIP_DECL_W(alg, kind, strat, 32)                                 \
    PARASAIL_HIP_DECL_W(alg, kind, strat, sat)

#define PARASAIL_HIP_DECL_ALG(alg)                                            \
    PARASAIL_HIP_DECL_S(alg, , striped) PARASAIL_HIP_DECL_S(alg, , scan)       \
    PARASAIL_HIP_DECL_S(alg, _stats, striped)                                 \
    PARASAIL_HIP_DECL_S(alg, _stats, scan)                                    \
    PARASAIL_HIP_DECL_S(alg, _trace, striped)                                 \
    PARASAIL_HIP_DECL_S(alg, _trace, scan)

PARASAIL_HIP_ALGS(PARASAIL_HIP_DECL_ALG)

/* The _table and _rowcol forms, replacing parasail_<ALG>[_stats]_{table,
 * rowcol}_<STRATEGY>[_profile]_<isa>_<bits>_<WIDTH> (built from each kernel
 * with -DPARASAIL_TABLE / -DPARASAIL_ROWCOL, CMakeLists.txt:800-861): the
 * result also carries every cell's H (and, for _stats, the matches,
 * similar and length counts of that H) as s1Len x s2Len row-major tables
 * (parasail_result_new_table1/table3, src/memory.c:171-229), or only the
 * DP's last row (s2Len entries) and last column (s1Len entries)
 * (parasail_result_new_rowcol1/rowcol3, src/memory.c:188-275); read them
 * with parasail_result_get_score_table, _get_score_row, ... */
#define PARASAIL_HIP_DECL_TS(alg, kind, tab, strat)                             \
    PARASAIL_HIP_DECL_W(alg, kind##_##tab, strat, 8)                            \
    PARASAIL_HIP_DECL_W(alg, kind##_##tab, strat, 16)                           \
    PARASAIL_HIP_DECL_W(alg, kind##_##tab, strat, 32)                           \
    PARASAIL_HIP_DECL_W(alg, kind##_##tab, strat, sat)
#define PARASAIL_HIP_DECL_TALG(alg)                                             \
    PARASAIL_HIP_DECL_TS(alg, , table, striped)                                 \
    PARASAIL_HIP_DECL_TS(alg, , table, scan)                                    \
    PARASAIL_HIP_DECL_TS(alg, , rowcol, striped)                                \
    PARASAIL_HIP_DECL_TS(alg, , rowcol, scan)                                   \
    PARASAIL_HIP_DECL_TS(alg, _stats, table, striped)                           \
    PARASAIL_HIP_DECL_TS(alg, _stats, table, scan)                              \
    PARASAIL_HIP_DECL_TS(alg, _stats, rowcol, striped)                          \
    PARASAIL_HIP_DECL_TS(alg, _stats, rowcol, scan)

PARASAIL_HIP_ALGS(PARASAIL_HIP_DECL_TALG)

/* profile creators, replacing parasail_profile_create[_stats]_avx_256_<W>
 * (src/memory_avx2.c:81-132, :240-253, :255-318).  The profile keeps the
 * mapped query and substitution table resident in device memory. */
parasail_profile_t *parasail_profile_create_hip_64_8(const char *const s1, const int s1Len,
                                                     const parasail_matrix_t *matrix);
parasail_profile_t *parasail_profile_create_hip_64_16(const char *const s1, const int s1Len,
                                                      const parasail_matrix_t *matrix);
parasail_profile_t *parasail_profile_create_hip_64_32(const char *const s1, const int s1Len,
                                                      const parasail_matrix_t *matrix);
parasail_profile_t *parasail_profile_create_hip_64_sat(const char *const s1, const int s1Len,
                                                       const parasail_matrix_t *matrix);
parasail_profile_t *parasail_profile_create_stats_hip_64_8(const char *const s1, const int s1Len,
                                                           const parasail_matrix_t *matrix);
parasail_profile_t *parasail_profile_create_stats_hip_64_16(const char *const s1, const int s1Len,
                                                            const parasail_matrix_t *matrix);
parasail_profile_t *parasail_profile_create_stats_hip_64_32(const char *const s1, const int s1Len,
                                                            const parasail_matrix_t *matrix);
parasail_profile_t *parasail_profile_create_stats_hip_64_sat(const char *const s1, const int s1Len,
                                                             const parasail_matrix_t *matrix);

/* banded global alignment, replacing parasail_nw_banded (parasail.h:293-297,
 * src/nw_banded.c:29-189) */
parasail_result_t *parasail_nw_banded_hip(const char *const s1, const int s1Len,
                                          const char *const s2, const int s2Len,
                                          const int open, const int gap, const int k,
                                          const parasail_matrix_t *matrix);

/* ---- batched entry points (no reference counterpart; the reference's
 * callers loop over pairs with OpenMP, tests/test_openmp.c:383-441) ----
 *
 * Aligns npairs (s1[k], s2[k]) in one device pass.  alg_name is any alg
 * above ("sw", "sg_qx", ...); kind 0 score, 1 stats, 2 trace.  Sequences
 * are concatenated with per-pair offsets and lengths.  Results go to flat
 * arrays (any may be NULL): score, end_query, end_ref, matches, similar,
 * length, and flags (the PARASAIL_FLAG_* a per-pair call would set).  For
 * kind 2, cigars[k] receives a parasail_cigar_t (free with
 * parasail_cigar_free) when cigars is non-NULL.  Returns 0, or -1 with a
 * message in parasail_hip_last_error().
 */
typedef struct parasail_hip_batch_result {
    int *score, *end_query, *end_ref;
    int *matches, *similar, *length;
    int *flag;
    parasail_cigar_t **cigars;
} parasail_hip_batch_result_t;

int parasail_hip_align_batch(const char *alg_name, int kind, int width, const char *s1,
                             const int64_t *s1_off, const int *s1_len, const char *s2,
                             const int64_t *s2_off, const int *s2_len, int npairs, int open,
                             int gap, const parasail_matrix_t *matrix,
                             parasail_hip_batch_result_t *out);

/* No counterpart in the reference.  For callers that cannot allocate
 * memory -- code under HIP graph capture, serving loops that must not call
 * hipMalloc: the variant of parasail_hip_align_batch at width 32 (exact
 * int32 results, kind 0 score or 1 stats) whose inputs and outputs are in
 * device memory.  s1/s2 are the residues as given (mapped on the device
 * through matrix->mapper) with device offsets and lengths; every array of
 * out is a device array of npairs ints (any may be NULL; cigars must be
 * NULL).  The work is queued on `stream` (a hipStream_t; NULL: the null
 * stream) and the call returns without waiting.  It allocates nothing:
 * device scratch comes from a static 64 MiB array per GPU, which limits
 * max_len, the longest s2 (a bound the caller gives; a pair whose s2 is
 * longer reports PARASAIL_FLAG_INVALID), to 8,248,832 residues for scores
 * and 2,062,208 for stats; the call fails with a message past that.  The matrix is square, of at most 128 letters.  Calls
 * that overlap on one GPU share the array and race.  0, or -1. */
int parasail_hip_align_batch_noalloc(const char *alg_name, int kind, const char *s1,
                                     const int64_t *s1_off, const int *s1_len, const char *s2,
                                     const int64_t *s2_off, const int *s2_len, int npairs,
                                     int max_len, int open, int gap,
                                     const parasail_matrix_t *matrix,
                                     parasail_hip_batch_result_t *out, void *stream);

/* Banded global alignment (parasail_nw_banded semantics, band k) of npairs
 * pairs in one device pass; score and end positions into out (the other
 * arrays of out are zeroed).  0, or -1 with a message. */
int parasail_hip_nw_banded_batch(const char *s1, const int64_t *s1_off, const int *s1_len,
                                 const char *s2, const int64_t *s2_off, const int *s2_len,
                                 int npairs, int open, int gap, int k,
                                 const parasail_matrix_t *matrix,
                                 parasail_hip_batch_result_t *out);
/* Tuning of the banded entry points for long pairs (no counterpart in the
 * reference): a pair whose target is long enough is cut into segments of
 * about `columns` target columns that run at the same time and are joined
 * exactly (a pair whose join fails is aligned again whole).  0 turns
 * segments off, a negative value keeps the setting; returns the previous
 * value (default 8192). */
int parasail_hip_nw_banded_set_segment(int columns);
/* of the last banded call: pairs cut into segments, segments, and pairs
 * aligned again whole because their segments did not join */
void parasail_hip_nw_banded_last_stats(int *pairs, int *segments, int *redone);

/* A batch kept resident in device memory (HBM): the residues are mapped and
 * uploaded once, and every parasail_hip_batch_align call over it launches
 * kernels only.  Score-only (kind 0) results, as parasail_hip_align_batch
 * with width "sat" returns them.  NULL / -1 on failure. */
typedef struct parasail_hip_batch parasail_hip_batch_t;

parasail_hip_batch_t *parasail_hip_batch_create(const char *s1, const int64_t *s1_off,
                                                const int *s1_len, const char *s2,
                                                const int64_t *s2_off, const int *s2_len,
                                                int npairs, const parasail_matrix_t *matrix);
int parasail_hip_batch_align(parasail_hip_batch_t *batch, const char *alg_name, int open,
                             int gap, parasail_hip_batch_result_t *out);
/* Several algorithms (at most 4, e.g. "sw", "nw", "sg") over one resident
 * batch in one call, results into out[0..nalg-1] as parasail_hip_batch_align
 * gives them: every pass is launched before the first is converted, so the
 * host's work on one pass's results overlaps the next passes' kernels.  No
 * reference counterpart (a batch interface; the reference has none).
 * parasail_hip_last_pass_ms reports pass k's device time of the last such
 * call, all of its kernels and its main kernel alone.  0, or -1. */
int parasail_hip_batch_align_multi(parasail_hip_batch_t *batch, int nalg,
                                   const char *const *alg_names, int open, int gap,
                                   parasail_hip_batch_result_t *out);
int parasail_hip_last_pass_ms(int k, double *kernel_ms, double *main_kernel_ms);
void parasail_hip_batch_free(parasail_hip_batch_t *batch);
/* A parasail_hip_batch_align_multi call that repeats the previous call's
 * algorithms and gaps on the wavefront tier (whose plans that call built)
 * is captured into a HIP graph -- every pass's kernels, result packing and
 * download -- and later such calls replay it with one launch.  on = 0
 * turns this off for the batch (returns the previous setting);
 * graph_stats gives the graphs captured and the calls that replayed one.
 * Neither has a counterpart in the reference. */
int parasail_hip_batch_set_graph(parasail_hip_batch_t *batch, int on);
int parasail_hip_batch_graph_stats(parasail_hip_batch_t *batch, int *captures, int *replays);
/* Measured tuning of the score-only batch planner (no counterpart in the
 * reference).  The planner picks, per pair, the rows per lane R of the
 * packed-int16 wavefront kernel by the cost of a step at each R; untuned
 * that cost comes from an instruction count.  parasail_hip_autotune times
 * the kernel on the device at every R (SW, and NW for NW/SG) over synthetic
 * pairs scored with matrix, open and gap, and installs the measured costs
 * for every later plan; _table reads them back (ns per wave-step; returns
 * how many, 0 if untuned); _reset returns to the instruction count.  Plans
 * already built keep their shapes.  0, or -1 with a message. */
int parasail_hip_autotune(const parasail_matrix_t *matrix, int open, int gap);
int parasail_hip_autotune_table(int *rows, double *sw_cost, double *nw_cost, int cap);
void parasail_hip_autotune_reset(void);

/* ---- database search and serving (no reference counterpart; the
 * reference's callers run parasail_sw_striped_profile_* over every target
 * of a database in an OpenMP loop, tests/test_openmp.c:353-369) ----
 *
 * A target database kept resident in device memory (HBM): residues are
 * mapped with the matrix's mapper and uploaded once.  A search is SW of one
 * query against every target (int16, with an exact int32 re-run of any
 * target whose score leaves the int16 window); it returns the k best hits,
 * score descending then target index ascending, with the end positions
 * parasail_sw would report.  The query is s1 with a square matrix, or a
 * PSSM (parasail_matrix_pssm_create; s1 may be NULL) whose alphabet is the
 * database's. */
typedef struct parasail_hip_db parasail_hip_db_t;
typedef struct parasail_hip_hit {
    int score, end_query, end_ref, target;
} parasail_hip_hit_t;

parasail_hip_db_t *parasail_hip_db_create(const char *seqs, const int64_t *off, const int *len,
                                          int n, const parasail_matrix_t *matrix);
void parasail_hip_db_free(parasail_hip_db_t *db);
/* number of hits written (min(k, targets)), or -1 */
int parasail_hip_db_search(parasail_hip_db_t *db, const char *s1, int s1Len,
                           const parasail_matrix_t *matrix, int open, int gap, int k,
                           parasail_hip_hit_t *hits);

/* Serving loop over a database: requests are queued by submit and run by a
 * background thread with up to four in flight (one HIP stream each), each
 * admitted when a slot and mem_budget bytes of device memory (0: 32 GiB)
 * allow; in-flight requests share the GPU chunk by chunk.  arrival is the
 * request's arrival time on parasail_hip_now()'s clock (0: now); its latency
 * is measured from there to the moment its hits are back in host memory.
 * poll returns 1 with the next finished request (hits[] needs room for k),
 * 0 if none is finished (wait = 0), -1 if that request failed. */
typedef struct parasail_hip_server parasail_hip_server_t;

parasail_hip_server_t *parasail_hip_server_create(parasail_hip_db_t *db, size_t mem_budget);
int parasail_hip_server_submit(parasail_hip_server_t *srv, int64_t id, const char *s1, int s1Len,
                               const parasail_matrix_t *matrix, int open, int gap, int k,
                               double arrival);
int parasail_hip_server_poll(parasail_hip_server_t *srv, int64_t *id, parasail_hip_hit_t *hits,
                             int *nhits, double *latency_ms, int wait);
void parasail_hip_server_free(parasail_hip_server_t *srv);
/* requests completed, requests that waited at admission because their
 * device memory did not fit the budget beside those in flight (a request
 * runs alone when nothing else is in flight), and the most requests in
 * flight at once */
int parasail_hip_server_stats(parasail_hip_server_t *srv, int64_t *completed, int64_t *deferred,
                              int *max_active);
/* monotonic clock (seconds) of the serving loop */
double parasail_hip_now(void);

/* FASTA/FASTQ ingest (parasail/io.h:33-59; src/parser.c:48-259): the
 * records of a file (plain or gzip), and their residues packed back to back
 * with a '$' after each; the residues of record i start after the i-th '$'
 * and run to the next, which gives the offsets and lengths
 * parasail_hip_batch_create takes.  A record may have an empty sequence, as
 * in the reference; a batch rejects empty sequences, so leave those records
 * out of it.  Same struct layouts. */
typedef struct parasail_string {
    size_t l;
    char *s;
} parasail_string_t;
typedef struct parasail_sequence {
    parasail_string_t name;
    parasail_string_t comment;
    parasail_string_t seq;
    parasail_string_t qual;
} parasail_sequence_t;
typedef struct parasail_sequences {
    parasail_sequence_t *seqs;
    size_t l;
    size_t characters;
    size_t shortest;
    size_t longest;
    float mean;
    float stddev;
} parasail_sequences_t;
/* replaces parasail_sequences_from_file (parasail/io.h:55) */
parasail_sequences_t *parasail_sequences_from_file(const char *fname);
/* replaces parasail_sequences_pack (parasail/io.h:57) */
char *parasail_sequences_pack(const parasail_sequences_t *sequences, size_t *size);
/* replaces parasail_sequences_free (parasail/io.h:59) */
void parasail_sequences_free(parasail_sequences_t *sequences);

/* The rest of parasail/io.h:18-103 (src/io.c:53-800): a file mapped into
 * memory, its FASTA/FASTQ statistics and its residues packed with a '$'
 * after each record, for files and for text already in memory. */
typedef struct parasail_file {
    int fd;
    off_t size;
    const char *buf;
} parasail_file_t;
typedef struct parasail_file_stat {
    unsigned long sequences;
    unsigned long characters;
    unsigned long shortest;
    unsigned long longest;
    float mean;
    float stddev;
} parasail_file_stat_t;
parasail_file_t *parasail_open(const char *fname);
void parasail_close(parasail_file_t *file);
int parasail_is_fasta(const parasail_file_t *pf);
int parasail_is_fastq(const parasail_file_t *pf);
parasail_file_stat_t *parasail_stat(const parasail_file_t *pf);
parasail_file_stat_t *parasail_stat_fasta(const parasail_file_t *pf);
parasail_file_stat_t *parasail_stat_fastq(const parasail_file_t *pf);
char *parasail_read(const parasail_file_t *pf, long *size);
char *parasail_pack(const parasail_file_t *pf, long *size);
char *parasail_pack_fasta(const parasail_file_t *pf, long *size);
char *parasail_pack_fastq(const parasail_file_t *pf, long *size);
int parasail_is_fasta_buffer(const char *buf, off_t size);
int parasail_is_fastq_buffer(const char *buf, off_t size);
parasail_file_stat_t *parasail_stat_buffer(const char *buffer, off_t size);
parasail_file_stat_t *parasail_stat_fasta_buffer(const char *buffer, off_t size);
parasail_file_stat_t *parasail_stat_fastq_buffer(const char *buffer, off_t size);
char *parasail_pack_buffer(const char *buffer, off_t size, long *packed_size);
char *parasail_pack_fasta_buffer(const char *buffer, off_t size, long *packed_size);
char *parasail_pack_fastq_buffer(const char *buffer, off_t size, long *packed_size);

/* substitution matrix or PSSM from a text file (src/io.c:921-1178) */
parasail_matrix_t *parasail_matrix_from_file(const char *filename);
parasail_matrix_t *parasail_matrix_from_file_case_sensitive(const char *filename);

/* device time (ms, HIP events) of the last call's alignment kernels */
double parasail_hip_last_kernel_ms(void);
/* device time (ms, HIP events on its stream) of the last call's dominant
 * kernel: the wavefront kernel for score-only int16 batches, otherwise the
 * same as parasail_hip_last_kernel_ms */
double parasail_hip_last_main_kernel_ms(void);
const char *parasail_hip_last_error(void);
/* 1 if a HIP device is usable, 0 otherwise (never falls back to the CPU) */
int parasail_hip_available(void);
/* compute units and peak engine clock (kHz) of the device in use */
int parasail_hip_device_info(int *compute_units, int *clock_khz);

#ifdef __cplusplus
}
#endif
#endif /* PARASAIL_HIP_H */
