/*
 * memory.c -- parasail's object model for the hip family: results,
 * matrices, profiles and their getters, with the reference's layouts,
 * allocators and NULL/stderr error convention (src/memory.c,
 * parasail/memory.h:27-94).  Results are malloc'ed exactly as the reference
 * allocates them, so either library's parasail_result_free releases them.
 */
#define _POSIX_C_SOURCE 200112L
#include <ctype.h>
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "internal.h"

void parasail_version(int *major, int *minor, int *patch)
{
    *major = PARASAIL_VERSION_MAJOR;
    *minor = PARASAIL_VERSION_MINOR;
    *patch = PARASAIL_VERSION_PATCH;
}

/* src/time.c:28-75 (monotonic clock) */
double parasail_time(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + (double)ts.tv_nsec * 1e-9;
}

void *phost_memalign(size_t alignment, size_t size)
{
    void *p = NULL;
    if (size == 0) size = 1;
    if (posix_memalign(&p, alignment, size) != 0) {
        fprintf(stderr, "%s: posix_memalign failed\n", __func__);
        return NULL;
    }
    return p;
}

/* parasail_result_new, src/memory.c:140-154 */
parasail_result_t *phost_result_new(void)
{
    parasail_result_t *r = (parasail_result_t *)malloc(sizeof(parasail_result_t));
    if (!r) {
        fprintf(stderr, "%s: failed to malloc %zu bytes\n", __func__, sizeof(parasail_result_t));
        return NULL;
    }
    r->score = 0;
    r->end_query = 0;
    r->end_ref = 0;
    r->flag = 0;
    r->extra = NULL;
    return r;
}

/* parasail_result_new_stats, src/memory.c:156-169 */
parasail_result_t *phost_result_new_stats(void)
{
    parasail_result_t *r = phost_result_new();
    if (!r) return NULL;
    r->stats = (parasail_result_extra_stats_t *)calloc(1, sizeof(parasail_result_extra_stats_t));
    if (!r->stats) { free(r); return NULL; }
    return r;
}

/* parasail_result_new_trace, src/memory.c:231-249: table of a*b elements of
 * `size` bytes, aligned */
parasail_result_t *phost_result_new_trace(int a, int b, size_t alignment, size_t size)
{
    if (a <= 0 || b <= 0) {
        fprintf(stderr, "%s: inputs must be > 0\n", __func__);
        return NULL;
    }
    parasail_result_t *r = phost_result_new();
    if (!r) return NULL;
    r->trace = (parasail_result_extra_trace_t *)malloc(sizeof(parasail_result_extra_trace_t));
    if (!r->trace) { free(r); return NULL; }
    r->trace->trace_table = phost_memalign(alignment, size * (size_t)a * (size_t)b);
    if (!r->trace->trace_table) { free(r->trace); free(r); return NULL; }
    r->trace->trace_ins_table = NULL;
    r->trace->trace_del_table = NULL;
    return r;
}

/* parasail_result_new_table1 / rowcol1 / table3 / rowcol3,
 * src/memory.c:171-275: a (query) by b (target) tables, zeroed; rows hold
 * b entries, columns a */
parasail_result_t *phost_result_new_tables(int a, int b, int stats, int rowcol)
{
    if (a <= 0 || b <= 0) {
        fprintf(stderr, "%s: inputs must be > 0\n", __func__);
        return NULL;
    }
    parasail_result_t *r = stats ? phost_result_new_stats() : phost_result_new();
    if (!r) return NULL;
    int ok = 1;
    const size_t ab = (size_t)a * (size_t)b;
    if (!stats && !rowcol) {
        r->tables = (parasail_result_extra_tables_t *)calloc(1, sizeof(*r->tables));
        ok = r->tables && (r->tables->score_table = (int *)calloc(ab, sizeof(int)));
        if (r->tables) r->flag |= PARASAIL_FLAG_TABLE;
    } else if (!stats) {
        r->rowcols = (parasail_result_extra_rowcols_t *)calloc(1, sizeof(*r->rowcols));
        if (r->rowcols) {
            r->flag |= PARASAIL_FLAG_ROWCOL;
            r->rowcols->score_row = (int *)calloc((size_t)b, sizeof(int));
            r->rowcols->score_col = (int *)calloc((size_t)a, sizeof(int));
        }
        ok = r->rowcols && r->rowcols->score_row && r->rowcols->score_col;
    } else if (!rowcol) {
        parasail_result_extra_stats_tables_t *t =
            (parasail_result_extra_stats_tables_t *)calloc(1, sizeof(*t));
        r->stats->tables = t;
        r->flag |= PARASAIL_FLAG_STATS;
        if (t) {
            r->flag |= PARASAIL_FLAG_TABLE;
            t->score_table = (int *)calloc(ab, sizeof(int));
            t->matches_table = (int *)calloc(ab, sizeof(int));
            t->similar_table = (int *)calloc(ab, sizeof(int));
            t->length_table = (int *)calloc(ab, sizeof(int));
        }
        ok = t && t->score_table && t->matches_table && t->similar_table && t->length_table;
    } else {
        parasail_result_extra_stats_rowcols_t *c =
            (parasail_result_extra_stats_rowcols_t *)calloc(1, sizeof(*c));
        r->stats->rowcols = c;
        r->flag |= PARASAIL_FLAG_STATS;
        if (c) {
            r->flag |= PARASAIL_FLAG_ROWCOL;
            c->score_row = (int *)calloc((size_t)b, sizeof(int));
            c->matches_row = (int *)calloc((size_t)b, sizeof(int));
            c->similar_row = (int *)calloc((size_t)b, sizeof(int));
            c->length_row = (int *)calloc((size_t)b, sizeof(int));
            c->score_col = (int *)calloc((size_t)a, sizeof(int));
            c->matches_col = (int *)calloc((size_t)a, sizeof(int));
            c->similar_col = (int *)calloc((size_t)a, sizeof(int));
            c->length_col = (int *)calloc((size_t)a, sizeof(int));
        }
        ok = c && c->score_row && c->matches_row && c->similar_row && c->length_row &&
             c->score_col && c->matches_col && c->similar_col && c->length_col;
    }
    if (!ok) {
        parasail_result_free(r);
        return NULL;
    }
    return r;
}

/* parasail_result_free, src/memory.c:277-327 */
void parasail_result_free(parasail_result_t *result)
{
    if (!result) {
        fprintf(stderr, "%s: attempted free of NULL result pointer\n", __func__);
        return;
    }
    if (result->flag & PARASAIL_FLAG_STATS) {
        if (result->stats) {
            if ((result->flag & PARASAIL_FLAG_TABLE) && result->stats->tables) {
                free(result->stats->tables->score_table);
                free(result->stats->tables->matches_table);
                free(result->stats->tables->similar_table);
                free(result->stats->tables->length_table);
                free(result->stats->tables);
            }
            if ((result->flag & PARASAIL_FLAG_ROWCOL) && result->stats->rowcols) {
                parasail_result_extra_stats_rowcols_t *rc = result->stats->rowcols;
                free(rc->score_row); free(rc->matches_row); free(rc->similar_row);
                free(rc->length_row); free(rc->score_col); free(rc->matches_col);
                free(rc->similar_col); free(rc->length_col);
                free(rc);
            }
            free(result->stats);
        }
    } else {
        if ((result->flag & PARASAIL_FLAG_TABLE) && result->tables) {
            free(result->tables->score_table);
            free(result->tables);
        }
        if ((result->flag & PARASAIL_FLAG_ROWCOL) && result->rowcols) {
            free(result->rowcols->score_row);
            free(result->rowcols->score_col);
            free(result->rowcols);
        }
    }
    if ((result->flag & PARASAIL_FLAG_TRACE) && result->trace) {
        free(result->trace->trace_table);
        if (result->trace->trace_ins_table) free(result->trace->trace_ins_table);
        if (result->trace->trace_del_table) free(result->trace->trace_del_table);
        free(result->trace);
    }
    free(result);
}

/* ---------------------------------------------------------------- getters */

#define CHECK_RESULT(r)                                                  \
    do {                                                                 \
        if (!(r)) {                                                      \
            fprintf(stderr, "%s: missing %s\n", __func__, "result");     \
            return 0;                                                    \
        }                                                                \
    } while (0)

int parasail_result_is_nw(const parasail_result_t *const r) { CHECK_RESULT(r); return r->flag & PARASAIL_FLAG_NW; }
int parasail_result_is_sg(const parasail_result_t *const r) { CHECK_RESULT(r); return r->flag & PARASAIL_FLAG_SG; }
int parasail_result_is_sw(const parasail_result_t *const r) { CHECK_RESULT(r); return r->flag & PARASAIL_FLAG_SW; }
int parasail_result_is_saturated(const parasail_result_t *const r) { CHECK_RESULT(r); return r->flag & PARASAIL_FLAG_SATURATED; }
int parasail_result_is_banded(const parasail_result_t *const r) { CHECK_RESULT(r); return r->flag & PARASAIL_FLAG_BANDED; }
int parasail_result_is_scan(const parasail_result_t *const r) { CHECK_RESULT(r); return r->flag & PARASAIL_FLAG_SCAN; }
int parasail_result_is_striped(const parasail_result_t *const r) { CHECK_RESULT(r); return r->flag & PARASAIL_FLAG_STRIPED; }
int parasail_result_is_diag(const parasail_result_t *const r) { CHECK_RESULT(r); return r->flag & PARASAIL_FLAG_DIAG; }
int parasail_result_is_blocked(const parasail_result_t *const r) { CHECK_RESULT(r); return r->flag & PARASAIL_FLAG_BLOCKED; }
int parasail_result_is_stats(const parasail_result_t *const r) { CHECK_RESULT(r); return r->flag & PARASAIL_FLAG_STATS; }
int parasail_result_is_stats_table(const parasail_result_t *const r) { CHECK_RESULT(r); return (r->flag & PARASAIL_FLAG_STATS) && (r->flag & PARASAIL_FLAG_TABLE); }
int parasail_result_is_stats_rowcol(const parasail_result_t *const r) { CHECK_RESULT(r); return (r->flag & PARASAIL_FLAG_STATS) && (r->flag & PARASAIL_FLAG_ROWCOL); }
int parasail_result_is_table(const parasail_result_t *const r) { CHECK_RESULT(r); return r->flag & PARASAIL_FLAG_TABLE; }
int parasail_result_is_rowcol(const parasail_result_t *const r) { CHECK_RESULT(r); return r->flag & PARASAIL_FLAG_ROWCOL; }
int parasail_result_is_trace(const parasail_result_t *const r) { CHECK_RESULT(r); return r->flag & PARASAIL_FLAG_TRACE; }
int parasail_result_get_score(const parasail_result_t *const r) { CHECK_RESULT(r); return r->score; }
int parasail_result_get_end_query(const parasail_result_t *const r) { CHECK_RESULT(r); return r->end_query; }
int parasail_result_get_end_ref(const parasail_result_t *const r) { CHECK_RESULT(r); return r->end_ref; }

/* src/memory.c:877-926: stats getters assert a stats result */
int parasail_result_get_matches(const parasail_result_t *const r)
{
    CHECK_RESULT(r);
    if (!(r->flag & PARASAIL_FLAG_STATS)) {
        fprintf(stderr, "%s: assert(%s) failed\n", __func__, "parasail_result_is_stats(result)");
        return 0;
    }
    return r->stats->matches;
}
int parasail_result_get_similar(const parasail_result_t *const r)
{
    CHECK_RESULT(r);
    if (!(r->flag & PARASAIL_FLAG_STATS)) {
        fprintf(stderr, "%s: assert(%s) failed\n", __func__, "parasail_result_is_stats(result)");
        return 0;
    }
    return r->stats->similar;
}
int parasail_result_get_length(const parasail_result_t *const r)
{
    CHECK_RESULT(r);
    if (!(r->flag & PARASAIL_FLAG_STATS)) {
        fprintf(stderr, "%s: assert(%s) failed\n", __func__, "parasail_result_is_stats(result)");
        return 0;
    }
    return r->stats->length;
}
/* table and row/column getters, src/memory.c:906-1027 */
#define NEED(cond, text)                                                            \
    do {                                                                            \
        if (!(cond)) {                                                              \
            fprintf(stderr, "%s: assert(%s) failed\n", __func__, text);            \
            return NULL;                                                            \
        }                                                                           \
    } while (0)
#define STATS_TABLE(r) (((r)->flag & PARASAIL_FLAG_STATS) && ((r)->flag & PARASAIL_FLAG_TABLE))
#define STATS_ROWCOL(r) (((r)->flag & PARASAIL_FLAG_STATS) && ((r)->flag & PARASAIL_FLAG_ROWCOL))
#define PLAIN_TABLE(r) (!((r)->flag & PARASAIL_FLAG_STATS) && ((r)->flag & PARASAIL_FLAG_TABLE))
#define PLAIN_ROWCOL(r) (!((r)->flag & PARASAIL_FLAG_STATS) && ((r)->flag & PARASAIL_FLAG_ROWCOL))

int *parasail_result_get_score_table(const parasail_result_t *const r)
{
    NEED(r, "result");
    NEED(STATS_TABLE(r) || PLAIN_TABLE(r), "parasail_result_is_table(result)");
    return STATS_TABLE(r) ? r->stats->tables->score_table : r->tables->score_table;
}
#define STATS_TABLE_GETTER(field)                                                   \
    int *parasail_result_get_##field##_table(const parasail_result_t *const r)      \
    {                                                                               \
        NEED(r, "result");                                                          \
        NEED(STATS_TABLE(r), "parasail_result_is_stats_table(result)");             \
        return r->stats->tables->field##_table;                                     \
    }
STATS_TABLE_GETTER(matches)
STATS_TABLE_GETTER(similar)
STATS_TABLE_GETTER(length)
int *parasail_result_get_score_row(const parasail_result_t *const r)
{
    NEED(r, "result");
    NEED(STATS_ROWCOL(r) || PLAIN_ROWCOL(r), "parasail_result_is_rowcol(result)");
    return STATS_ROWCOL(r) ? r->stats->rowcols->score_row : r->rowcols->score_row;
}
int *parasail_result_get_score_col(const parasail_result_t *const r)
{
    NEED(r, "result");
    NEED(STATS_ROWCOL(r) || PLAIN_ROWCOL(r), "parasail_result_is_rowcol(result)");
    return STATS_ROWCOL(r) ? r->stats->rowcols->score_col : r->rowcols->score_col;
}
#define STATS_ROWCOL_GETTER(field, rc)                                              \
    int *parasail_result_get_##field##_##rc(const parasail_result_t *const r)       \
    {                                                                               \
        NEED(r, "result");                                                          \
        NEED(STATS_ROWCOL(r), "parasail_result_is_stats_rowcol(result)");           \
        return r->stats->rowcols->field##_##rc;                                     \
    }
STATS_ROWCOL_GETTER(matches, row)
STATS_ROWCOL_GETTER(similar, row)
STATS_ROWCOL_GETTER(length, row)
STATS_ROWCOL_GETTER(matches, col)
STATS_ROWCOL_GETTER(similar, col)
STATS_ROWCOL_GETTER(length, col)
int *parasail_result_get_trace_ins_table(const parasail_result_t *const r)
{
    NEED(r, "result");
    NEED(r->flag & PARASAIL_FLAG_TRACE, "parasail_result_is_trace(result)");
    return (int *)r->trace->trace_ins_table;
}
int *parasail_result_get_trace_del_table(const parasail_result_t *const r)
{
    NEED(r, "result");
    NEED(r->flag & PARASAIL_FLAG_TRACE, "parasail_result_is_trace(result)");
    return (int *)r->trace->trace_del_table;
}

int *parasail_result_get_trace_table(const parasail_result_t *const r)
{
    CHECK_RESULT(r);
    if (!(r->flag & PARASAIL_FLAG_TRACE)) {
        fprintf(stderr, "%s: assert(%s) failed\n", __func__, "parasail_result_is_trace(result)");
        return 0;
    }
    return (int *)r->trace->trace_table;
}

/* ---------------------------------------------------------------- matrices */

static int *make_mapper(const char *alphabet, size_t size, int case_sensitive)
{
    int *mapper = (int *)malloc(sizeof(int) * 256);
    if (!mapper) return NULL;
    for (int c = 0; c < 256; ++c) mapper[c] = (int)size;
    for (size_t i = 0; i < size; ++i) {
        unsigned char a = (unsigned char)alphabet[i];
        if (case_sensitive) {
            mapper[a] = (int)i;
        } else {
            mapper[toupper(a)] = (int)i;
            mapper[tolower(a)] = (int)i;
        }
    }
    return mapper;
}

static char *alphabet_with_star(const char *alphabet, size_t size)
{
    char *a = (char *)calloc(size + 2, 1);
    if (!a) return NULL;
    memcpy(a, alphabet, size);
    a[size] = '*';
    return a;
}

/* parasail_matrix_create_internal, src/memory.c:336-407 */
static parasail_matrix_t *matrix_create(const char *alphabet, int match, int mismatch, int cs)
{
    if (!alphabet) { fprintf(stderr, "%s: missing %s\n", __func__, "alphabet"); return NULL; }
    size_t size = strlen(alphabet);
    if (size >= INT_MAX || size + 1 > 256) {
        fprintf(stderr, "%s: alphabet is too large\n", __func__);
        return NULL;
    }
    size_t size1 = size + 1;
    int *matrix = (int *)calloc(size1 * size1, sizeof(int));
    int *mapper = make_mapper(alphabet, size, cs);
    char *alpha = alphabet_with_star(alphabet, size);
    parasail_matrix_t *m = (parasail_matrix_t *)malloc(sizeof(parasail_matrix_t));
    if (!matrix || !mapper || !alpha || !m) {
        free(matrix); free(mapper); free(alpha); free(m);
        return NULL;
    }
    for (size_t i = 0; i < size; ++i)
        for (size_t j = 0; j < size; ++j) matrix[i * size1 + j] = (i == j) ? match : mismatch;
    m->name = "";
    m->matrix = matrix;
    m->mapper = mapper;
    m->size = (int)size1;
    m->max = match > mismatch ? match : mismatch;
    m->min = match > mismatch ? mismatch : match;
    m->user_matrix = matrix;
    m->type = PARASAIL_MATRIX_TYPE_SQUARE;
    m->length = m->size;
    m->alphabet = alpha;
    m->query = NULL;
    return m;
}

parasail_matrix_t *parasail_matrix_create(const char *alphabet, const int match, const int mismatch)
{
    return matrix_create(alphabet, match, mismatch, 0);
}

parasail_matrix_t *parasail_matrix_create_case_sensitive(const char *alphabet, const int match,
                                                         const int mismatch)
{
    return matrix_create(alphabet, match, mismatch, 1);
}

/* parasail_matrix_pssm_create_internal, src/memory.c:421-498 */
static parasail_matrix_t *pssm_create(const char *alphabet, const int *values, int length, int cs)
{
    if (!alphabet) { fprintf(stderr, "%s: missing %s\n", __func__, "alphabet"); return NULL; }
    if (!values) { fprintf(stderr, "%s: missing %s\n", __func__, "values"); return NULL; }
    if (length <= 0) { fprintf(stderr, "%s: length must be > 0\n", __func__); return NULL; }
    size_t size = strlen(alphabet);
    if (size >= INT_MAX || size + 1 > 256) {
        fprintf(stderr, "%s: alphabet is too large\n", __func__);
        return NULL;
    }
    size_t size1 = size + 1;
    int mn = INT_MAX, mx = INT_MIN;
    for (size_t i = 0; i < size * (size_t)length; ++i) {
        if (values[i] < mn) mn = values[i];
        if (values[i] > mx) mx = values[i];
    }
    int *matrix = (int *)calloc(size1 * (size_t)length, sizeof(int));
    int *mapper = make_mapper(alphabet, size, cs);
    char *alpha = alphabet_with_star(alphabet, size);
    parasail_matrix_t *m = (parasail_matrix_t *)malloc(sizeof(parasail_matrix_t));
    if (!matrix || !mapper || !alpha || !m) {
        free(matrix); free(mapper); free(alpha); free(m);
        return NULL;
    }
    size_t v = 0, c = 0;
    for (int r = 0; r < length; ++r) {
        for (size_t k = 0; k < size; ++k) matrix[c++] = values[v++];
        matrix[c++] = mn;
    }
    m->name = "";
    m->matrix = matrix;
    m->mapper = mapper;
    m->size = (int)size1;
    m->max = mx;
    m->min = mn;
    m->user_matrix = matrix;
    m->type = PARASAIL_MATRIX_TYPE_PSSM;
    m->length = length;
    m->alphabet = alpha;
    m->query = NULL;
    return m;
}

parasail_matrix_t *parasail_matrix_pssm_create(const char *alphabet, const int *values,
                                               const int length)
{
    return pssm_create(alphabet, values, length, 0);
}

parasail_matrix_t *parasail_matrix_pssm_create_case_sensitive(const char *alphabet,
                                                              const int *values, const int length)
{
    return pssm_create(alphabet, values, length, 1);
}

/* parasail_matrix_copy, src/memory.c:506-565 */
parasail_matrix_t *parasail_matrix_copy(const parasail_matrix_t *matrix)
{
    if (!matrix) { fprintf(stderr, "%s: missing %s\n", __func__, "matrix"); return NULL; }
    size_t msize = matrix->type == PARASAIL_MATRIX_TYPE_SQUARE
                       ? (size_t)matrix->size * matrix->size
                       : (size_t)matrix->size * matrix->length;
    parasail_matrix_t *m = (parasail_matrix_t *)malloc(sizeof(parasail_matrix_t));
    int *mapper = (int *)malloc(sizeof(int) * 256);
    int *vals = (int *)malloc(sizeof(int) * (msize ? msize : 1));
    char *alpha = (char *)calloc((size_t)matrix->size + 1, 1);
    if (!m || !mapper || !vals || !alpha) {
        free(m); free(mapper); free(vals); free(alpha);
        return NULL;
    }
    *m = *matrix;
    memcpy(mapper, matrix->mapper, sizeof(int) * 256);
    memcpy(vals, matrix->matrix, sizeof(int) * msize);
    memcpy(alpha, matrix->alphabet, (size_t)matrix->size + 1);
    m->mapper = mapper;
    m->matrix = vals;
    m->user_matrix = vals;
    m->alphabet = alpha;
    m->query = NULL;
    if (matrix->query) {
        size_t ql = strlen(matrix->query);
        char *q = (char *)malloc(ql + 1);
        if (!q) { parasail_matrix_free(m); return NULL; }
        memcpy(q, matrix->query, ql + 1);
        m->query = q;
    }
    return m;
}

/* parasail_matrix_convert_square_to_pssm, src/memory.c:567-621 */
parasail_matrix_t *parasail_matrix_convert_square_to_pssm(const parasail_matrix_t *matrix,
                                                          const char *s1, int s1Len)
{
    if (!matrix) { fprintf(stderr, "%s: missing %s\n", __func__, "matrix"); return NULL; }
    if (matrix->type != PARASAIL_MATRIX_TYPE_SQUARE) {
        fprintf(stderr, "%s: attempted to convert non-square matrix to pssm\n", __func__);
        return NULL;
    }
    if (!s1 || s1Len <= 0) { fprintf(stderr, "%s: missing %s\n", __func__, "s1"); return NULL; }
    parasail_matrix_t *m = (parasail_matrix_t *)malloc(sizeof(parasail_matrix_t));
    int *mapper = (int *)malloc(sizeof(int) * 256);
    int *vals = (int *)calloc((size_t)matrix->size * s1Len, sizeof(int));
    char *alpha = (char *)calloc((size_t)matrix->size + 1, 1);
    char *query = (char *)calloc((size_t)s1Len + 1, 1);
    if (!m || !mapper || !vals || !alpha || !query) {
        free(m); free(mapper); free(vals); free(alpha); free(query);
        return NULL;
    }
    memcpy(mapper, matrix->mapper, sizeof(int) * 256);
    memcpy(alpha, matrix->alphabet, (size_t)matrix->size + 1);
    memcpy(query, s1, (size_t)s1Len);
    for (int i = 0; i < s1Len; ++i)
        memcpy(&vals[(size_t)matrix->size * i],
               &matrix->matrix[(size_t)matrix->size * matrix->mapper[(unsigned char)s1[i]]],
               sizeof(int) * (size_t)matrix->size);
    m->name = matrix->name;
    m->matrix = vals;
    m->mapper = mapper;
    m->size = matrix->size;
    m->max = matrix->max;
    m->min = matrix->min;
    m->user_matrix = vals;
    m->type = PARASAIL_MATRIX_TYPE_PSSM;
    m->length = s1Len;
    m->alphabet = alpha;
    m->query = query;
    return m;
}

/* src/memory.c:623-640 */
void parasail_matrix_set_value(parasail_matrix_t *matrix, int row, int col, int value)
{
    if (!matrix) { fprintf(stderr, "%s: missing %s\n", __func__, "matrix"); return; }
    if (!matrix->user_matrix) {
        fprintf(stderr, "%s: attempted to set value of built-in matrix '%s'\n", __func__,
                matrix->name);
        return;
    }
    matrix->user_matrix[row * matrix->size + col] = value;
    if (value > matrix->max) matrix->max = value;
    if (value < matrix->min) matrix->min = value;
}

/* src/memory.c:642-658: built-in matrices (user_matrix == NULL) are refused */
void parasail_matrix_free(parasail_matrix_t *matrix)
{
    if (!matrix) { fprintf(stderr, "%s: missing %s\n", __func__, "matrix"); return; }
    if (matrix->user_matrix) {
        free((void *)matrix->matrix);
        free((void *)matrix->mapper);
        free((void *)matrix->alphabet);
        if (matrix->query) free((void *)matrix->query);
        free(matrix);
    } else {
        fprintf(stderr, "%s: attempted to free built-in matrix '%s'\n", __func__, matrix->name);
    }
}

/* ---------------------------------------------------------------- profiles */

/* parasail_profile_new, src/memory.c:660-696: borrows s1 and matrix */
parasail_profile_t *phost_profile_new(const char *s1, int s1Len, const parasail_matrix_t *matrix)
{
    if (!matrix) { fprintf(stderr, "%s: missing %s\n", __func__, "matrix"); return NULL; }
    if (matrix->type == PARASAIL_MATRIX_TYPE_SQUARE && !s1) {
        fprintf(stderr, "%s: missing %s\n", __func__, "s1");
        return NULL;
    }
    parasail_profile_t *p = (parasail_profile_t *)calloc(1, sizeof(parasail_profile_t));
    if (!p) return NULL;
    p->s1 = s1;
    p->s1Len = matrix->type == PARASAIL_MATRIX_TYPE_PSSM ? matrix->length : s1Len;
    p->matrix = matrix;
    p->free = NULL;
    p->stop = INT_MAX;
    return p;
}

/* parasail_profile_free, src/memory.c:698-746 */
void parasail_profile_free(parasail_profile_t *profile)
{
    if (!profile) {
        fprintf(stderr, "%s: attempted free of NULL profile pointer\n", __func__);
        return;
    }
    struct parasail_profile_data *d[4] = {&profile->profile8, &profile->profile16,
                                          &profile->profile32, &profile->profile64};
    for (int k = 0; k < 4; ++k) {
        if (d[k]->score) profile->free(d[k]->score);
        if (d[k]->matches) profile->free(d[k]->matches);
        if (d[k]->similar) profile->free(d[k]->similar);
    }
    free(profile);
}

/* ---------------------------------------------------------------- unwind */

/* parasail_striped_unwind (parasail/memory.h:137-141, src/striped_unwind.c:
 * 5-26, :56-111): `array` holds a striped or scan table of lenb columns of
 * segLen*lanes entries, entry i of column j being query row
 * i / lanes + (i % lanes) * segLen; lanes and element width come from the
 * result's LANES_* and BITS_* flags.  Returns the row-major int table,
 * malloc'ed; NULL for a result that is neither striped nor scan. */
int *parasail_striped_unwind(int lena, int lenb, parasail_result_t *result, void *array)
{
    if (!result || !array || lena <= 0 || lenb <= 0) return NULL;
    if (!(result->flag & (PARASAIL_FLAG_STRIPED | PARASAIL_FLAG_SCAN))) return NULL;
    int lanes = phost_flag_lanes(result->flag);
    int bits = phost_flag_bits(result->flag);
    int *out = (int *)malloc(sizeof(int) * (size_t)lena * lenb);
    if (!out) return NULL;
    long segLen = (lena + lanes - 1) / lanes;
    for (long j = 0; j < lenb; ++j)
        for (long i = 0; i < lena; ++i) {
            long loc = j * segLen * lanes + (i % segLen) * lanes + i / segLen;
            out[i * lenb + j] = phost_trace_at(array, bits, loc);
        }
    return out;
}

int phost_flag_lanes(int flag)
{
    if (flag & PARASAIL_FLAG_LANES_64) return 64;
    if (flag & PARASAIL_FLAG_LANES_32) return 32;
    if (flag & PARASAIL_FLAG_LANES_16) return 16;
    if (flag & PARASAIL_FLAG_LANES_8) return 8;
    if (flag & PARASAIL_FLAG_LANES_4) return 4;
    if (flag & PARASAIL_FLAG_LANES_2) return 2;
    return 1;
}

int phost_flag_bits(int flag)
{
    if (flag & PARASAIL_FLAG_BITS_8) return 8;
    if (flag & PARASAIL_FLAG_BITS_16) return 16;
    if (flag & PARASAIL_FLAG_BITS_32) return 32;
    if (flag & PARASAIL_FLAG_BITS_64) return 64;
    return 32;
}

int phost_trace_at(const void *t, int bits, long loc)
{
    switch (bits) {
    case 8: return ((const int8_t *)t)[loc];
    case 16: return ((const int16_t *)t)[loc];
    case 32: return ((const int32_t *)t)[loc];
    default: return (int)((const int64_t *)t)[loc];
    }
}
