/*
 * align.c -- the hip family's per-call entry points.
 *
 * Each parasail_<alg>[_stats|_trace]_<strategy>[_profile]_hip_64_<width>
 * validates its arguments the way the reference does (PARASAIL_CHECK_*,
 * parasail/memory.h:47-94), maps the residues through matrix->mapper, runs
 * one pair through the device (phip_align) and returns an ordinary
 * parasail_result_t.
 *
 * Widths.  The device computes every pair exactly (packed int16 when the
 * pair provably fits, int32 otherwise).  A width only decides how the
 * result is reported: a width whose window the exact DP values leave is
 * reported saturated with the reference's saturated outputs (SW: score
 * 2^w-1 and ends 0, src/sw_striped_avx2_256_16.c:374-382; NW/SG: score 0,
 * ends 0, src/nw_striped_avx2_256_16.c:380-385), and "sat" reports the
 * narrowest width that holds the alignment (src/satcheck.c:41520-41537).
 * The windows are those of the reference's kernel of the same algorithm,
 * kind and strategy (phost_saturates below).
 */
#include <errno.h>
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

static int width_max(int w) { return w == 8 ? INT8_MAX : w == 16 ? INT16_MAX : INT32_MAX; }
static int width_min(int w) { return w == 8 ? INT8_MIN : w == 16 ? INT16_MIN : INT32_MIN; }

/* Whether the reference's w-bit kernel of this algorithm, kind and
 * strategy reports saturation, decided on the exact DP values:
 *   SW striped score and _stats keep H biased by INT_MIN(w) and stop past
 *   maxp = INT_MAX(w) - (max+1) on that scale, i.e. a score above
 *   2^w - max - 2 (sw_striped_avx2_256_16.c:205-210,350-362); _stats also
 *   when a count of a cell's H clamps at INT_MAX(w), a true 2^w - 1
 *   (sw_stats_striped_avx2_256_16.c:440-442,612-614);
 *   SW striped _trace is unbiased, maxp = INT_MAX(w) - (max+1)
 *   (sw_trace_striped_avx2_256_16.c:177,403-416), and SW scan tests H (and
 *   the _stats counts) against POS_LIMIT, the same number
 *   (sw_scan_avx2_256_16.c:218-219,334-390, sw_stats_scan...:537-541);
 *   NW/SG test min < NEG_LIMIT or max > POS_LIMIT
 *   (sg_striped_avx2_256_16.c:209-210) over H, E and F (striped score and
 *   _trace, scan _trace; sg_striped...:317-320), over H alone for the min
 *   and H with its counts for the max (_stats, striped and scan,
 *   sg_stats_striped...:441-445), and over H alone for nw scan score
 *   (nw_scan_avx2_256_16.c:345-346).  sg scan score also folds in its
 *   prefix-scan gap vector (sg_scan_avx2_256_16.c:298-304), whose length
 *   depends on the ISA's lane count; here it takes the H/E/F window. */
int phost_saturates(int alg, int kind, int scan, int w, int open,
                    const parasail_matrix_t *matrix, const phip_out_t *o)
{
    if (w >= 32) return 0;
    const long pos = (long)width_max(w) - matrix->max - 1;
    const long neg =
        (-open < matrix->min ? (long)width_min(w) + open : (long)width_min(w) - matrix->min) + 1;
    if (alg == PHIP_SW) {
        if (!scan && kind != PHIP_TRACE) {
            if (o->score > (1L << w) - matrix->max - 2) return 1;
            return kind == PHIP_STATS && o->max_len >= (1L << w) - 1;
        }
        if (o->score > pos) return 1;
        return kind == PHIP_STATS && o->max_len > pos;
    }
    if (kind == PHIP_STATS) return o->min_h < neg || o->max_val > pos || o->max_len > pos;
    if (scan && kind == PHIP_SCORE && alg == PHIP_NW) return o->min_h < neg || o->max_val > pos;
    return o->min_val < neg || o->max_val > pos;
}

/* The values a saturated result reports: SW striped score 2^w - 1
 * (INT_MAX(w) less the bias, sw_striped_avx2_256_16.c:378-382,412), SW
 * striped _stats 2^(w-1) for the score and every count (0 less the bias,
 * sw_stats_striped_avx2_256_16.c:616-623,661-667), everything else 0
 * (sw_trace_striped...:419-423, sg_striped...:433-439); ends are 0. */
int phost_saturated_value(int alg, int kind, int scan, int w)
{
    if (alg != PHIP_SW || scan || kind == PHIP_TRACE) return 0;
    return kind == PHIP_STATS ? (int)(1L << (w - 1)) : (int)((1L << w) - 1);
}

/* The width whose saturation a call must test (phip_batch_t
 * want_sat_check): the narrowest one its result may report. */
int phost_check_width(int alg, int kind, int scan, int width)
{
    if (width == 32) return 0;
    /* SW's windows on the score need no extrema, its _stats ones the counts */
    if (alg == PHIP_SW && kind != PHIP_STATS) return 0;
    (void)scan;
    return width == 16 ? 16 : 8;
}

static int bits_flag(int w)
{
    return w == 8 ? PARASAIL_FLAG_BITS_8 : w == 16 ? PARASAIL_FLAG_BITS_16 : PARASAIL_FLAG_BITS_32;
}

#define CHECK_NULL(fn, v)                                        \
    do {                                                         \
        if (!(v)) {                                              \
            fprintf(stderr, "%s: missing %s\n", fn, #v);         \
            return NULL;                                         \
        }                                                        \
    } while (0)
#define CHECK_GT0(fn, v)                                         \
    do {                                                         \
        if ((v) <= 0) {                                          \
            fprintf(stderr, "%s: %s must be > 0\n", fn, #v);     \
            return NULL;                                         \
        }                                                        \
    } while (0)
#define CHECK_GE0(fn, v)                                         \
    do {                                                         \
        if ((v) < 0) {                                           \
            fprintf(stderr, "%s: %s must be >= 0\n", fn, #v);    \
            return NULL;                                         \
        }                                                        \
    } while (0)

static int sg_result_flags(int sg_flags)
{
    int f = 0;
    if (sg_flags & PHIP_S1_BEG) f |= PARASAIL_FLAG_SG_S1_BEG;
    if (sg_flags & PHIP_S1_END) f |= PARASAIL_FLAG_SG_S1_END;
    if (sg_flags & PHIP_S2_BEG) f |= PARASAIL_FLAG_SG_S2_BEG;
    if (sg_flags & PHIP_S2_END) f |= PARASAIL_FLAG_SG_S2_END;
    return f;
}

/* tab: 0, or 1 for the _table and 2 for the _rowcol form of the function
 * (src/memory.c:171-275 results, src/sw.c:159-170 values) */
/* qcodes/dprof: a profile's query codes (host) and their device copy with
 * the matrix, or NULL to encode s1 here */
static parasail_result_t *align_ex(const char *fname, int alg, int sg_flags, int kind, int width,
                                   int scan, int tab, const char *s1, int s1Len, const char *s2,
                                   int s2Len, int open, int gap, const parasail_matrix_t *matrix,
                                   const uint8_t *qcodes, const phip_dprofile_t *dprof)
{
    CHECK_NULL(fname, s2);
    CHECK_GT0(fname, s2Len);
    CHECK_GE0(fname, open);
    CHECK_GE0(fname, gap);
    CHECK_NULL(fname, matrix);
    const int pssm = matrix->type == PARASAIL_MATRIX_TYPE_PSSM;
    if (!pssm) {
        CHECK_NULL(fname, s1);
        CHECK_GT0(fname, s1Len);
    }
    if (kind == PHIP_STATS && !s1) {
        fprintf(stderr, "%s: stats functions with pssm require a consensus sequence\n", fname);
        return NULL;
    }
    if (matrix->size > 256 || matrix->size <= 0) {
        fprintf(stderr, "%s: matrix size %d unsupported\n", fname, matrix->size);
        return NULL;
    }
    if (phip_init() != 0) {
        errno = ENOSYS;
        return NULL;
    }
    const int m = pssm ? matrix->length : s1Len;
    const int mq = s1 ? (pssm && s1Len > 0 ? s1Len : m) : 0;
    if (pssm && s1 && mq < m) {
        fprintf(stderr, "%s: consensus sequence shorter than the pssm\n", fname);
        return NULL;
    }
    uint8_t *q = NULL, *t = (uint8_t *)malloc((size_t)s2Len);
    if (s1) q = (uint8_t *)malloc((size_t)m);
    if (!t || (s1 && !q)) { free(q); free(t); return NULL; }
    if (s1 && qcodes) memcpy(q, qcodes, (size_t)m);
    for (int i = 0; s1 && !qcodes && i < m; ++i)
        q[i] = (uint8_t)matrix->mapper[(unsigned char)s1[i]];
    for (int j = 0; j < s2Len; ++j) t[j] = (uint8_t)matrix->mapper[(unsigned char)s2[j]];

    phip_batch_t b;
    memset(&b, 0, sizeof(b));
    b.alg = alg;
    b.kind = kind;
    b.sg_flags = sg_flags;
    b.open = open;
    b.gap = gap;
    /* the narrowest width whose saturation the result must report */
    b.want_sat_check = phost_check_width(alg, kind, scan, width);
    b.want_min_h = kind == PHIP_STATS || (scan && kind == PHIP_SCORE && alg == PHIP_NW);
    b.matrix.values = matrix->matrix;
    b.matrix.size = matrix->size;
    b.matrix.length = pssm ? matrix->length : matrix->size;
    b.matrix.is_pssm = pssm;
    b.matrix.max = matrix->max;
    b.matrix.min = matrix->min;
    b.npairs = 1;
    int64_t qo = 0, to = 0;
    int32_t ql = m, tl = s2Len;
    b.q = q;
    b.qoff = &qo;
    b.qlen = &ql;
    b.t = t;
    b.toff = &to;
    b.tlen = &tl;
    b.dprof = s1 ? dprof : NULL;
    phip_out_t o;
    int8_t *tr = NULL;
    parasail_result_t *rt = NULL; /* the table result, made before the call */
    int *rcbuf = NULL;            /* _rowcol: last row then last column, per plane */
    int rc;
    if (tab) {
        const int nplanes = kind == PHIP_STATS ? 4 : 1;
        rt = phost_result_new_tables(m, s2Len, kind == PHIP_STATS, tab == 2);
        if (tab == 2) rcbuf = (int *)malloc(sizeof(int) * (size_t)nplanes * (m + s2Len));
        if (!rt || (tab == 2 && !rcbuf)) {
            if (rt) parasail_result_free(rt);
            free(q); free(t); free(rcbuf);
            return NULL;
        }
        int *dst[4] = {NULL, NULL, NULL, NULL};
        for (int k = 0; k < nplanes; ++k) {
            if (tab == 2) dst[k] = rcbuf + (size_t)k * (m + s2Len);
            else if (kind != PHIP_STATS) dst[k] = rt->tables->score_table;
            else dst[k] = k == 0 ? rt->stats->tables->score_table
                          : k == 1 ? rt->stats->tables->matches_table
                          : k == 2 ? rt->stats->tables->similar_table
                                   : rt->stats->tables->length_table;
        }
        rc = phip_align_tables(&b, tab, &o, dst);
    } else {
        if (kind == PHIP_TRACE) {
            tr = (int8_t *)malloc((size_t)m * s2Len);
            if (!tr) { free(q); free(t); return NULL; }
        }
        rc = phip_align_one(&b, &o, kind == PHIP_TRACE ? &tr : NULL);
    }
    free(q);
    free(t);
    if (rc != 0) {
        fprintf(stderr, "%s: %s\n", fname, phip_last_error());
        free(tr);
        free(rcbuf);
        if (rt) parasail_result_free(rt);
        return NULL;
    }
    if (tab == 2) {
        /* rows hold s2Len entries, columns m (src/memory.c:188-204, 251-275) */
        const int n = s2Len;
        int **rows[4], **cols[4];
        if (kind == PHIP_STATS) {
            parasail_result_extra_stats_rowcols_t *c = rt->stats->rowcols;
            rows[0] = &c->score_row; rows[1] = &c->matches_row;
            rows[2] = &c->similar_row; rows[3] = &c->length_row;
            cols[0] = &c->score_col; cols[1] = &c->matches_col;
            cols[2] = &c->similar_col; cols[3] = &c->length_col;
        } else {
            rows[0] = &rt->rowcols->score_row;
            cols[0] = &rt->rowcols->score_col;
        }
        for (int k = 0; k < (kind == PHIP_STATS ? 4 : 1); ++k) {
            memcpy(*rows[k], rcbuf + (size_t)k * (m + n), sizeof(int) * (size_t)n);
            memcpy(*cols[k], rcbuf + (size_t)k * (m + n) + n, sizeof(int) * (size_t)m);
        }
        free(rcbuf);
    }

    int w = width;
    int saturated = 0;
    if (w == 0) {
        w = !phost_saturates(alg, kind, scan, 8, open, matrix, &o) ? 8
            : !phost_saturates(alg, kind, scan, 16, open, matrix, &o) ? 16 : 32;
    } else {
        saturated = phost_saturates(alg, kind, scan, w, open, matrix, &o);
    }

    parasail_result_t *r = NULL;
    if (rt) {
        r = rt;
    } else if (kind == PHIP_STATS) {
        r = phost_result_new_stats();
    } else if (kind == PHIP_TRACE) {
        const int lanes = 64, seglen = (m + lanes - 1) / lanes;
        const size_t esz = w == 8 ? 1 : w == 16 ? 2 : 4;
        r = phost_result_new_trace(seglen * lanes, s2Len, 64, esz);
        if (r) {
            /* striped layout of cigar.c:157 with 64 segments */
            memset(r->trace->trace_table, 0, esz * (size_t)seglen * lanes * s2Len);
            for (long j = 0; j < s2Len; ++j)
                for (long i = 0; i < m; ++i) {
                    long loc = j * seglen * lanes + (i % seglen) * lanes + i / seglen;
                    int v = tr[i * s2Len + j];
                    if (esz == 1) ((int8_t *)r->trace->trace_table)[loc] = (int8_t)v;
                    else if (esz == 2) ((int16_t *)r->trace->trace_table)[loc] = (int16_t)v;
                    else ((int32_t *)r->trace->trace_table)[loc] = v;
                }
        }
    } else {
        r = phost_result_new();
    }
    free(tr);
    if (!r) return NULL;
    r->flag |= alg == PHIP_SW ? PARASAIL_FLAG_SW : alg == PHIP_NW ? PARASAIL_FLAG_NW : PARASAIL_FLAG_SG;
    if (alg == PHIP_SG) r->flag |= sg_result_flags(sg_flags);
    r->flag |= scan ? PARASAIL_FLAG_SCAN : PARASAIL_FLAG_STRIPED;
    r->flag |= bits_flag(w) | PARASAIL_FLAG_LANES_64;
    if (kind == PHIP_STATS) r->flag |= PARASAIL_FLAG_STATS;
    if (kind == PHIP_TRACE) r->flag |= PARASAIL_FLAG_TRACE;
    if (saturated) {
        const int sv = phost_saturated_value(alg, kind, scan, w);
        r->flag |= PARASAIL_FLAG_SATURATED;
        r->score = sv;
        r->end_query = 0;
        r->end_ref = 0;
        if (kind == PHIP_STATS) {
            r->stats->matches = sv;
            r->stats->similar = sv;
            r->stats->length = sv;
        }
    } else {
        r->score = o.score;
        r->end_query = o.end_query;
        r->end_ref = o.end_ref;
        if (kind == PHIP_STATS) {
            r->stats->matches = o.matches;
            r->stats->similar = o.similar;
            r->stats->length = o.length;
        }
    }
    return r;
}

parasail_result_t *phost_align(const char *fname, int alg, int sg_flags, int kind, int width,
                               int scan, const char *s1, int s1Len, const char *s2, int s2Len,
                               int open, int gap, const parasail_matrix_t *matrix,
                               const parasail_profile_t *profile)
{
    (void)profile;
    return align_ex(fname, alg, sg_flags, kind, width, scan, 0, s1, s1Len, s2, s2Len, open, gap,
                    matrix, NULL, NULL);
}

/* The hip profile's data (the `score` member of the width's
 * parasail_profile_data, freed by the profile's `free`): the query's
 * residue codes and, for a square matrix, their device copy with the
 * matrix values, uploaded when the profile is made. */
typedef struct hip_pdata {
    uint8_t *codes;
    phip_dprofile_t *dev;
} hip_pdata_t;

/* the data for a call at width w (0: _sat), else any the profile holds */
static const hip_pdata_t *profile_data(const parasail_profile_t *p, int w)
{
    const void *pick = w == 8 ? p->profile8.score : w == 32 ? p->profile32.score
                                                            : p->profile16.score;
    if (!pick) pick = p->profile16.score ? p->profile16.score
                    : p->profile8.score  ? p->profile8.score
                                         : p->profile32.score;
    return (const hip_pdata_t *)pick;
}

static parasail_result_t *phost_align_profile(const char *fname, int alg, int sg_flags, int kind,
                                              int width, int scan, int tab,
                                              const parasail_profile_t *profile, const char *s2,
                                              int s2Len, int open, int gap)
{
    CHECK_NULL(fname, profile);
    CHECK_NULL(fname, profile->matrix);
    CHECK_GT0(fname, profile->s1Len);
    const hip_pdata_t *d = profile_data(profile, width);
    return align_ex(fname, alg, sg_flags, kind, width, scan, tab, profile->s1, profile->s1Len, s2,
                    s2Len, open, gap, profile->matrix, d ? d->codes : NULL, d ? d->dev : NULL);
}

/* ---- the generated entry points ---- */

#define ALG_sw PHIP_SW
#define ALG_nw PHIP_NW
#define ALG_sg PHIP_SG
#define ALG_sg_qb PHIP_SG
#define ALG_sg_qe PHIP_SG
#define ALG_sg_qx PHIP_SG
#define ALG_sg_db PHIP_SG
#define ALG_sg_de PHIP_SG
#define ALG_sg_dx PHIP_SG
#define ALG_sg_qb_de PHIP_SG
#define ALG_sg_qe_db PHIP_SG
#define ALG_sg_qb_db PHIP_SG
#define ALG_sg_qe_de PHIP_SG
/* src/sg_helper.h:53-63: (s1_beg, s1_end, s2_beg, s2_end) per name */
#define SGF_sw 0
#define SGF_nw 0
#define SGF_sg (PHIP_S1_BEG | PHIP_S1_END | PHIP_S2_BEG | PHIP_S2_END)
#define SGF_sg_qb (PHIP_S1_BEG)
#define SGF_sg_qe (PHIP_S1_END)
#define SGF_sg_qx (PHIP_S1_BEG | PHIP_S1_END)
#define SGF_sg_db (PHIP_S2_BEG)
#define SGF_sg_de (PHIP_S2_END)
#define SGF_sg_dx (PHIP_S2_BEG | PHIP_S2_END)
#define SGF_sg_qb_de (PHIP_S1_BEG | PHIP_S2_END)
#define SGF_sg_qe_db (PHIP_S1_END | PHIP_S2_BEG)
#define SGF_sg_qb_db (PHIP_S1_BEG | PHIP_S2_BEG)
#define SGF_sg_qe_de (PHIP_S1_END | PHIP_S2_END)

#define KIND_ PHIP_SCORE
#define KIND__stats PHIP_STATS
#define KIND__trace PHIP_TRACE
#define SCAN_striped 0
#define SCAN_scan 1
#define WID_8 8
#define WID_16 16
#define WID_32 32
#define WID_sat 0

#define DEF_W(alg, kind, strat, w)                                                          \
    PHOST_EXPORT parasail_result_t *parasail_##alg##kind##_##strat##_hip_64_##w(            \
        const char *const s1, const int s1Len, const char *const s2, const int s2Len,        \
        const int open, const int gap, const parasail_matrix_t *matrix)                      \
    {                                                                                        \
        return phost_align(__func__, ALG_##alg, SGF_##alg, KIND_##kind, WID_##w,             \
                           SCAN_##strat, s1, s1Len, s2, s2Len, open, gap, matrix, NULL);     \
    }                                                                                        \
    PHOST_EXPORT parasail_result_t *parasail_##alg##kind##_##strat##_profile_hip_64_##w(    \
        const parasail_profile_t *const profile, const char *const s2, const int s2Len,     \
        const int open, const int gap)                                                       \
    {                                                                                        \
        return phost_align_profile(__func__, ALG_##alg, SGF_##alg, KIND_##kind, WID_##w,     \
                                   SCAN_##strat, 0, profile, s2, s2Len, open, gap);          \
    }

#define DEF_S(alg, kind, strat) \
    DEF_W(alg, kind, strat, 8) DEF_W(alg, kind, strat, 16) DEF_W(alg, kind, strat, 32) DEF_W(alg, kind, strat, sat)

#define DEF_ALG(alg)                                                                        \
    DEF_S(alg, , striped) DEF_S(alg, , scan) DEF_S(alg, _stats, striped)                    \
    DEF_S(alg, _stats, scan) DEF_S(alg, _trace, striped) DEF_S(alg, _trace, scan)

PARASAIL_HIP_ALGS(DEF_ALG)

/* the _table and _rowcol forms (KIND "" or _stats, TAB table or rowcol) */
#define TAB_table 1
#define TAB_rowcol 2
#define DEF_TW(alg, kind, tab, strat, w)                                                    \
    PHOST_EXPORT parasail_result_t *parasail_##alg##kind##_##tab##_##strat##_hip_64_##w(    \
        const char *const s1, const int s1Len, const char *const s2, const int s2Len,        \
        const int open, const int gap, const parasail_matrix_t *matrix)                      \
    {                                                                                        \
        return align_ex(__func__, ALG_##alg, SGF_##alg, KIND_##kind, WID_##w, SCAN_##strat,  \
                        TAB_##tab, s1, s1Len, s2, s2Len, open, gap, matrix, NULL, NULL);     \
    }                                                                                        \
    PHOST_EXPORT parasail_result_t                                                           \
        *parasail_##alg##kind##_##tab##_##strat##_profile_hip_64_##w(                        \
            const parasail_profile_t *const profile, const char *const s2, const int s2Len,  \
            const int open, const int gap)                                                   \
    {                                                                                        \
        return phost_align_profile(__func__, ALG_##alg, SGF_##alg, KIND_##kind, WID_##w,     \
                                   SCAN_##strat, TAB_##tab, profile, s2, s2Len, open, gap);  \
    }
#define DEF_TS(alg, kind, tab, strat)                                                       \
    DEF_TW(alg, kind, tab, strat, 8) DEF_TW(alg, kind, tab, strat, 16)                       \
    DEF_TW(alg, kind, tab, strat, 32) DEF_TW(alg, kind, tab, strat, sat)
#define DEF_TALG(alg)                                                                       \
    DEF_TS(alg, , table, striped) DEF_TS(alg, , table, scan)                                 \
    DEF_TS(alg, , rowcol, striped) DEF_TS(alg, , rowcol, scan)                               \
    DEF_TS(alg, _stats, table, striped) DEF_TS(alg, _stats, table, scan)                     \
    DEF_TS(alg, _stats, rowcol, striped) DEF_TS(alg, _stats, rowcol, scan)

PARASAIL_HIP_ALGS(DEF_TALG)

/* ---- profiles ---- */

static void profile_data_free(void *p)
{
    hip_pdata_t *d = (hip_pdata_t *)p;
    phip_profile_release(d->dev);
    free(d->codes);
    free(d);
}

/* The hip profile keeps the query's residue codes and, for a square
 * matrix, a device copy of the codes and the matrix values that every
 * per-call alignment with the profile reads in place of an upload
 * (src/memory_avx2.c:81-132 builds striped score vectors instead, which
 * the wavefront kernels do not need; src/memory.c:698-746 frees them
 * through the profile's `free`). */
static parasail_profile_t *profile_create(const char *fname, const char *s1, int s1Len,
                                          const parasail_matrix_t *matrix, int stats, int w)
{
    CHECK_NULL(fname, matrix);
    if (matrix->type == PARASAIL_MATRIX_TYPE_SQUARE) {
        CHECK_NULL(fname, s1);
        CHECK_GT0(fname, s1Len);
    }
    if (stats && !s1) {
        fprintf(stderr, "%s: stats functions with pssm require a consensus sequence\n", fname);
        return NULL;
    }
    if (phip_init() != 0) {
        errno = ENOSYS;
        return NULL;
    }
    parasail_profile_t *p = phost_profile_new(s1, s1Len, matrix);
    if (!p) return NULL;
    const int m = p->s1Len;
    uint8_t *codes = (uint8_t *)malloc((size_t)m + 1);
    hip_pdata_t *pd = (hip_pdata_t *)calloc(1, sizeof(hip_pdata_t));
    if (!codes || !pd) { free(codes); free(pd); free(p); return NULL; }
    for (int i = 0; i < m; ++i) codes[i] = s1 ? (uint8_t)matrix->mapper[(unsigned char)s1[i]] : 0;
    pd->codes = codes;
    if (s1 && matrix->type == PARASAIL_MATRIX_TYPE_SQUARE) {
        phip_matrix_t mx;
        memset(&mx, 0, sizeof(mx));
        mx.values = matrix->matrix;
        mx.size = matrix->size;
        mx.length = matrix->size;
        pd->dev = phip_profile_upload(codes, m, &mx);
        if (!pd->dev) {
            fprintf(stderr, "%s: %s\n", fname, phip_last_error());
            free(codes); free(pd); free(p);
            return NULL;
        }
    }
    p->free = profile_data_free;
    /* one width's slot; _sat's profile16 (src/memory_avx2.c:240-253 fills
     * both 8 and 16 there, which share one copy here) */
    struct parasail_profile_data *d = w == 8 ? &p->profile8 : w == 32 ? &p->profile32 : &p->profile16;
    d->score = pd;
    return p;
}

#define DEF_PC(w, W)                                                                         \
    PHOST_EXPORT parasail_profile_t *parasail_profile_create_hip_64_##w(                     \
        const char *const s1, const int s1Len, const parasail_matrix_t *matrix)               \
    {                                                                                         \
        return profile_create(__func__, s1, s1Len, matrix, 0, W);                            \
    }                                                                                         \
    PHOST_EXPORT parasail_profile_t *parasail_profile_create_stats_hip_64_##w(               \
        const char *const s1, const int s1Len, const parasail_matrix_t *matrix)               \
    {                                                                                         \
        return profile_create(__func__, s1, s1Len, matrix, 1, W);                            \
    }
DEF_PC(8, 8)
DEF_PC(16, 16)
DEF_PC(32, 32)
DEF_PC(sat, 0)

/* ---- banded NW ---- */

PHOST_EXPORT parasail_result_t *parasail_nw_banded_hip(const char *const s1, const int s1Len,
                                                       const char *const s2, const int s2Len,
                                                       const int open, const int gap, const int k,
                                                       const parasail_matrix_t *matrix)
{
    const char *fname = __func__;
    CHECK_NULL(fname, s1);
    CHECK_GT0(fname, s1Len);
    CHECK_NULL(fname, s2);
    CHECK_GT0(fname, s2Len);
    CHECK_GE0(fname, open);
    CHECK_GE0(fname, gap);
    CHECK_GT0(fname, k);
    CHECK_NULL(fname, matrix);
    if (phip_init() != 0) {
        errno = ENOSYS;
        return NULL;
    }
    uint8_t *q = (uint8_t *)malloc((size_t)s1Len), *t = (uint8_t *)malloc((size_t)s2Len);
    if (!q || !t) { free(q); free(t); return NULL; }
    for (int i = 0; i < s1Len; ++i) q[i] = (uint8_t)matrix->mapper[(unsigned char)s1[i]];
    for (int j = 0; j < s2Len; ++j) t[j] = (uint8_t)matrix->mapper[(unsigned char)s2[j]];
    phip_matrix_t mx;
    mx.values = matrix->matrix;
    mx.size = matrix->size;
    mx.length = matrix->type == PARASAIL_MATRIX_TYPE_PSSM ? matrix->length : matrix->size;
    mx.is_pssm = matrix->type == PARASAIL_MATRIX_TYPE_PSSM;
    mx.max = matrix->max;
    mx.min = matrix->min;
    phip_out_t o;
    int rc = phip_nw_banded(q, s1Len, t, s2Len, open, gap, k, &mx, &o);
    free(q);
    free(t);
    if (rc != 0) {
        fprintf(stderr, "%s: %s\n", fname, phip_last_error());
        return NULL;
    }
    parasail_result_t *r = phost_result_new();
    if (!r) return NULL;
    /* src/nw_banded.c:75-77 */
    r->flag |= PARASAIL_FLAG_NW | PARASAIL_FLAG_BANDED | PARASAIL_FLAG_BITS_32 | PARASAIL_FLAG_LANES_64;
    r->score = o.score;
    r->end_query = s1Len - 1;
    r->end_ref = s2Len - 1;
    return r;
}

/* ---- misc ---- */

PHOST_EXPORT double parasail_hip_last_kernel_ms(void) { return phip_last_kernel_ms(); }
PHOST_EXPORT double parasail_hip_last_main_kernel_ms(void) { return phip_last_main_kernel_ms(); }
PHOST_EXPORT const char *parasail_hip_last_error(void) { return phip_last_error(); }
PHOST_EXPORT int parasail_hip_available(void) { return phip_init() == 0; }
PHOST_EXPORT int parasail_hip_device_info(int *compute_units, int *clock_khz)
{
    return phip_device_info(compute_units, clock_khz);
}
