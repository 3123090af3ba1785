/*
 * traceback.c -- alignment strings and the printed alignment of a trace
 * result (parasail_result_get_traceback[_extra], parasail_traceback_generic
 * [_extra[2]]), and the SSW-library emulation on the hip functions
 * (parasail_ssw, parasail_ssw_profile, parasail_ssw_init).
 *
 * The strings follow src/traceback_template.c:13-214: from (end_query,
 * end_ref) the trace codes are followed back -- a diagonal move prints both
 * residues and a comparison character (match, pos for a positive
 * substitution, neg otherwise; equality case-insensitive unless asked, and
 * the alias pairs count as equal), a move along s2 prints '-' against the
 * target residue, a move along s1 the query residue against '-'.  Semi-
 * global results print their end gaps first; once the walk leaves the
 * table, NW and SG print the rest of the other sequence against gaps, SW
 * prints nothing more.  A PSSM result without a query prints '*' for it.
 * The printed form and its statistics follow :216-430 (blocks of `width`
 * columns, names cut or padded to name_width, positions in int_width
 * columns).  These are output formats, so the text is the reference's.
 */
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

/* the walk's strings, built back to front, and the gap counts the printed
 * form's positions need */
typedef struct {
    char *q, *c, *d;
    size_t n;
    long c_ins, c_del;
} walk_t;

static char query_at(const char *seqA, long i) { return seqA ? seqA[i] : '*'; }

static char compare_char(const char *seqA, long i, const char *seqB, long j,
                         const parasail_matrix_t *matrix, char match, char pos, char neg,
                         int case_sensitive, const char *aliases, size_t naliases)
{
    const char a = query_at(seqA, i), b = seqB[j];
    int eq = case_sensitive ? a == b : toupper((unsigned char)a) == toupper((unsigned char)b);
    /* every adjacent pair of the alias string, as traceback.c:41-50 steps
     * through it one character at a time (its last step reads past the
     * string; that one is left out) */
    for (size_t k = 0; k + 1 < naliases; ++k) {
        if (aliases[k] == a) eq |= aliases[k + 1] == b;
        else if (aliases[k + 1] == a) eq |= aliases[k] == b;
    }
    if (eq) return match;
    const int col = matrix->mapper[(unsigned char)b];
    const int sub = matrix->type == PARASAIL_MATRIX_TYPE_SQUARE
                        ? matrix->matrix[matrix->mapper[(unsigned char)a] * matrix->size + col]
                        : matrix->matrix[(long)i * matrix->size + col];
    return sub > 0 ? pos : neg;
}

static void put(walk_t *w, char q, char c, char d)
{
    w->q[w->n] = q;
    w->c[w->n] = c;
    w->d[w->n] = d;
    ++w->n;
}

static void reverse(char *s, size_t n)
{
    for (size_t a = 0, z = n ? n - 1 : 0; a < z; ++a, --z) {
        const char t = s[a];
        s[a] = s[z];
        s[z] = t;
    }
}

/* 0, or -1 when the trace codes are inconsistent */
static int walk(parasail_result_t *result, const char *seqA, int lena, const char *seqB, int lenb,
                const parasail_matrix_t *matrix, char match, char pos, char neg,
                int case_sensitive, const char *aliases_in, walk_t *w)
{
    char aliases[256];
    size_t naliases = 0;
    if (aliases_in) {
        naliases = strlen(aliases_in);
        if (naliases % 2 || naliases >= 256) {
            fprintf(stderr, "traceback: alphabet_aliases needs an even length below 256\n");
            return -1;
        }
        for (size_t k = 0; k < naliases; ++k)
            aliases[k] = case_sensitive ? aliases_in[k] : (char)toupper((unsigned char)aliases_in[k]);
    }
    const size_t cap = (size_t)lena + lenb + 1;
    w->q = (char *)malloc(cap);
    w->c = (char *)malloc(cap);
    w->d = (char *)malloc(cap);
    w->n = 0;
    w->c_ins = w->c_del = 0;
    if (!w->q || !w->c || !w->d) return -1;
    const int striped = (result->flag & (PARASAIL_FLAG_STRIPED | PARASAIL_FLAG_SCAN)) != 0;
    const int bits = striped ? phost_flag_bits(result->flag) : 8;
    const long lanes = phost_flag_lanes(result->flag);
    const long seglen = (lena + lanes - 1) / lanes;
    const int sw = (result->flag & PARASAIL_FLAG_SW) != 0;
    const void *tab = result->trace->trace_table;
    long i = result->end_query, j = result->end_ref;
    if (result->flag & PARASAIL_FLAG_SG) {
        if (result->end_query + 1 == lena) {
            for (long k = lenb - 1; k > j; --k, ++w->c_ins) put(w, '-', ' ', seqB[k]);
        } else if (result->end_ref + 1 == lenb) {
            for (long k = lena - 1; k > i; --k, ++w->c_del) put(w, query_at(seqA, k), ' ', '-');
        } else {
            return -1;
        }
    }
    int where = PARASAIL_DIAG;
    while (i >= 0 || j >= 0) {
        if (i < 0 || j < 0) {
            /* off the table: NW and SG print the rest against gaps */
            if (!sw) {
                for (; j >= 0; --j, ++w->c_ins) put(w, '-', ' ', seqB[j]);
                for (; i >= 0; --i, ++w->c_del) put(w, query_at(seqA, i), ' ', '-');
            }
            break;
        }
        const long loc = striped ? j * seglen * lanes + (i % seglen) * lanes + i / seglen
                                 : i * (long)lenb + j;
        const int code = phost_trace_at(tab, bits, loc);
        if (where == PARASAIL_DIAG) {
            if (code & PARASAIL_DIAG) {
                put(w, query_at(seqA, i),
                    compare_char(seqA, i, seqB, j, matrix, match, pos, neg, case_sensitive,
                                 aliases, naliases),
                    seqB[j]);
                --i;
                --j;
            } else if (code & PARASAIL_INS) {
                where = PARASAIL_INS;
            } else if (code & PARASAIL_DEL) {
                where = PARASAIL_DEL;
            } else {
                break;
            }
        } else if (where == PARASAIL_INS) {
            put(w, '-', ' ', seqB[j]);
            ++w->c_ins;
            --j;
            if (code & PARASAIL_DIAG_E) where = PARASAIL_DIAG;
            else if (!(code & PARASAIL_INS_E)) return -1;
        } else {
            put(w, query_at(seqA, i), ' ', '-');
            ++w->c_del;
            --i;
            if (code & PARASAIL_DIAG_F) where = PARASAIL_DIAG;
            else if (!(code & PARASAIL_DEL_F)) return -1;
        }
    }
    reverse(w->q, w->n);
    reverse(w->c, w->n);
    reverse(w->d, w->n);
    w->q[w->n] = w->c[w->n] = w->d[w->n] = 0;
    return 0;
}

static void walk_free(walk_t *w)
{
    free(w->q);
    free(w->c);
    free(w->d);
}

PHOST_EXPORT parasail_traceback_t *parasail_result_get_traceback_extra(
    parasail_result_t *result, const char *seqA, int lena, const char *seqB, int lenb,
    const parasail_matrix_t *matrix, char match, char pos, char neg, int case_sensitive,
    const char *alphabet_aliases)
{
    if (!result || !(result->flag & PARASAIL_FLAG_TRACE) || !result->trace || !seqB ||
        !matrix || lena <= 0 || lenb <= 0) {
        fprintf(stderr, "%s: assert(parasail_result_is_trace(result)) failed\n", __func__);
        return NULL;
    }
    walk_t w;
    memset(&w, 0, sizeof(w));
    parasail_traceback_t *tb = NULL;
    if (walk(result, seqA, lena, seqB, lenb, matrix, match, pos, neg, case_sensitive,
             alphabet_aliases, &w) == 0 &&
        (tb = (parasail_traceback_t *)malloc(sizeof(*tb)))) {
        tb->query = w.q;
        tb->comp = w.c;
        tb->ref = w.d;
        return tb;
    }
    fprintf(stderr, "%s: internal error\n", __func__);
    walk_free(&w);
    return NULL;
}

PHOST_EXPORT parasail_traceback_t *parasail_result_get_traceback(
    parasail_result_t *result, const char *seqA, int lena, const char *seqB, int lenb,
    const parasail_matrix_t *matrix, char match, char pos, char neg)
{
    return parasail_result_get_traceback_extra(result, seqA, lena, seqB, lenb, matrix, match, pos,
                                               neg, 0, NULL);
}

PHOST_EXPORT void parasail_traceback_free(parasail_traceback_t *traceback)
{
    if (!traceback) return;
    free(traceback->query);
    free(traceback->comp);
    free(traceback->ref);
    free(traceback);
}

/* one labelled line of a printed block: name, the position before the
 * block, the block's characters, the position after it */
static void print_line(FILE *f, const char *name, int name_width, int int_width, long *pindex,
                       const char *s, long at, long cnt)
{
    const long nl = name ? (long)strlen(name) : 0;
    fprintf(f, "%-*.*s", name_width, (int)(nl < name_width ? nl : name_width), name ? name : "");
    fprintf(f, " %*lld ", int_width, (long long)*pindex + 1);
    for (long k = 0; k < cnt; ++k) {
        if (s[at + k] != '-') ++*pindex;
        fputc(s[at + k], f);
    }
    fprintf(f, " %*lld\n", int_width, (long long)*pindex);
}

PHOST_EXPORT void parasail_traceback_generic_extra2(
    const char *seqA, int lena, const char *seqB, int lenb, const char *nameA, const char *nameB,
    const parasail_matrix_t *matrix, parasail_result_t *result, char match, char pos, char neg,
    int width, int name_width, int use_stats, int int_width, FILE *stream, int case_sensitive,
    const char *alphabet_aliases)
{
    if (!result || !(result->flag & PARASAIL_FLAG_TRACE) || !stream || width <= 0) {
        fprintf(stderr, "%s: assert(parasail_result_is_trace(result)) failed\n", __func__);
        return;
    }
    walk_t w;
    memset(&w, 0, sizeof(w));
    if (walk(result, seqA, lena, seqB, lenb, matrix, match, pos, neg, case_sensitive,
             alphabet_aliases, &w) != 0) {
        fprintf(stderr, "%s: internal error\n", __func__);
        walk_free(&w);
        return;
    }
    /* positions need room for lena + lenb */
    char tmp[32];
    const int need = snprintf(tmp, sizeof(tmp), "%llu", (unsigned long long)lena + (unsigned long long)lenb);
    if (int_width < need) int_width = need;
    const long len = (long)w.n;
    long qp = 0, dp = 0, mch = 0, sim = 0, gaps = 0;
    if (result->flag & PARASAIL_FLAG_SW) {
        qp = result->end_query + 1 - len + w.c_ins;
        dp = result->end_ref + 1 - len + w.c_del;
    }
    for (long at = 0; at < len; at += width) {
        const long cnt = len - at < width ? len - at : width;
        fputc('\n', stream);
        print_line(stream, nameB, name_width, int_width, &dp, w.d, at, cnt);
        fprintf(stream, "%*s", name_width + 1 + int_width + 1, "");
        for (long k = 0; k < cnt; ++k) {
            const char ch = w.c[at + k];
            if (ch == match) { ++mch; ++sim; }
            else if (ch == pos) ++sim;
            else if (ch == neg) {}
            else if (ch == ' ') ++gaps;
            fputc(ch, stream);
        }
        fputc('\n', stream);
        print_line(stream, nameA, name_width, int_width, &qp, w.q, at, cnt);
    }
    if (use_stats) {
        fprintf(stream, "\n");
        fprintf(stream, "Length: %lld\n", (long long)len);
        fprintf(stream, "Identity:   %*lld/%lld (%4.1f%%)\n", int_width, (long long)mch,
                (long long)len, 100.0 * mch / len);
        fprintf(stream, "Similarity: %*lld/%lld (%4.1f%%)\n", int_width, (long long)sim,
                (long long)len, 100.0 * sim / len);
        fprintf(stream, "Gaps:       %*lld/%lld (%4.1f%%)\n", int_width, (long long)gaps,
                (long long)len, 100.0 * gaps / len);
        fprintf(stream, "Score: %d\n", result->score);
    }
    walk_free(&w);
}

PHOST_EXPORT void parasail_traceback_generic_extra(
    const char *seqA, int lena, const char *seqB, int lenb, const char *nameA, const char *nameB,
    const parasail_matrix_t *matrix, parasail_result_t *result, char match, char pos, char neg,
    int width, int name_width, int use_stats, int int_width, FILE *stream)
{
    parasail_traceback_generic_extra2(seqA, lena, seqB, lenb, nameA, nameB, matrix, result, match,
                                      pos, neg, width, name_width, use_stats, int_width, stream, 0,
                                      NULL);
}

PHOST_EXPORT void parasail_traceback_generic(const char *seqA, int lena, const char *seqB,
                                             int lenb, const char *nameA, const char *nameB,
                                             const parasail_matrix_t *matrix,
                                             parasail_result_t *result, char match, char pos,
                                             char neg, int width, int name_width, int use_stats)
{
    parasail_traceback_generic_extra2(seqA, lena, seqB, lenb, nameA, nameB, matrix, result, match,
                                      pos, neg, width, name_width, use_stats, 7, stdout, 0, NULL);
}

/* ---------------------------------------------------------------- SSW */

/* parasail_ssw_init (src/pssw.c:16-43): score_size 0 keeps the 8-bit
 * profile, 1 the 16-bit one, 2 both */
PHOST_EXPORT parasail_profile_t *parasail_ssw_init(const char *const s1, const int s1Len,
                                                   const parasail_matrix_t *matrix,
                                                   const int8_t score_size)
{
    parasail_profile_t *p8 = NULL, *p16 = NULL;
    if (score_size == 0 || score_size == 2) p8 = parasail_profile_create_hip_64_8(s1, s1Len, matrix);
    if (score_size == 1 || score_size == 2)
        p16 = parasail_profile_create_hip_64_16(s1, s1Len, matrix);
    if (p8 && p16) {
        p8->profile16 = p16->profile16;
        free(p16);
        return p8;
    }
    return p8 ? p8 : p16;
}

/* SW at 8 bits while that width holds, then 16 (has8 turns off for the
 * rest of the call once 8 bits saturate); NULL when 16 bits saturate too */
typedef parasail_result_t *(*pair_fn)(const char *, int, const char *, int, int, int,
                                      const parasail_matrix_t *);

static parasail_result_t *try_widths(int *has8, int has16, pair_fn f8, pair_fn f16,
                                     const char *s1, int l1, const char *s2, int l2, int open,
                                     int gap, const parasail_matrix_t *matrix)
{
    parasail_result_t *r = NULL;
    if (*has8) {
        r = f8(s1, l1, s2, l2, open, gap, matrix);
        if (r && (r->flag & PARASAIL_FLAG_SATURATED)) {
            *has8 = 0;
            parasail_result_free(r);
            r = NULL;
        }
    }
    if (!r && has16) {
        r = f16(s1, l1, s2, l2, open, gap, matrix);
        if (r && (r->flag & PARASAIL_FLAG_SATURATED)) {
            parasail_result_free(r);
            r = NULL;
        }
    }
    return r;
}

static char *reversed(const char *s, int n)
{
    char *r = (char *)malloc((size_t)n + 1);
    if (!r) return NULL;
    for (int k = 0; k < n; ++k) r[k] = s[n - 1 - k];
    r[n] = 0;
    return r;
}

/* parasail_ssw_profile (src/pssw.c:56-183): the end of the best local
 * alignment from a forward pass, its start from a pass over the reversed
 * prefixes, and its CIGAR from a traceback of the aligned stretch alone */
PHOST_EXPORT parasail_result_ssw_t *parasail_ssw_profile(const parasail_profile_t *const profile,
                                                         const char *const s2, const int s2Len,
                                                         const int open, const int gap)
{
    if (!profile || !profile->s1 || !s2) {
        fprintf(stderr, "%s: missing arguments\n", __func__);
        return NULL;
    }
    const char *s1 = profile->s1;
    const parasail_matrix_t *matrix = profile->matrix;
    int has8 = profile->profile8.score != NULL, has16 = profile->profile16.score != NULL;
    parasail_result_t *fwd = NULL, *rev = NULL, *fin = NULL;
    if (has8) {
        fwd = parasail_sw_striped_profile_hip_64_8(profile, s2, s2Len, open, gap);
        if (fwd && (fwd->flag & PARASAIL_FLAG_SATURATED)) {
            has8 = 0;
            parasail_result_free(fwd);
            fwd = NULL;
        }
    }
    if (!fwd && has16) {
        fwd = parasail_sw_striped_profile_hip_64_16(profile, s2, s2Len, open, gap);
        if (fwd && (fwd->flag & PARASAIL_FLAG_SATURATED)) {
            parasail_result_free(fwd);
            fwd = NULL;
        }
    }
    if (!fwd) return NULL;
    int l1 = fwd->end_query + 1, l2 = fwd->end_ref + 1;
    char *r1 = reversed(s1, l1), *r2 = reversed(s2, l2);
    if (r1 && r2)
        rev = try_widths(&has8, has16, parasail_sw_striped_hip_64_8, parasail_sw_striped_hip_64_16,
                         r1, l1, r2, l2, open, gap, matrix);
    free(r1);
    free(r2);
    if (!rev) {
        parasail_result_free(fwd);
        return NULL;
    }
    const int off1 = fwd->end_query - rev->end_query, off2 = fwd->end_ref - rev->end_ref;
    l1 = rev->end_query + 1;
    l2 = rev->end_ref + 1;
    fin = try_widths(&has8, has16, parasail_sw_trace_striped_hip_64_8,
                     parasail_sw_trace_striped_hip_64_16, s1 + off1, l1, s2 + off2, l2, open, gap,
                     matrix);
    parasail_result_ssw_t *out = NULL;
    parasail_cigar_t *cig = fin ? parasail_result_get_cigar(fin, s1 + off1, l1, s2 + off2, l2, matrix)
                                : NULL;
    if (cig && (out = (parasail_result_ssw_t *)malloc(sizeof(*out)))) {
        out->score1 = (uint16_t)fwd->score;
        out->ref_begin1 = off2;
        out->ref_end1 = fwd->end_ref;
        out->read_begin1 = off1;
        out->read_end1 = fwd->end_query;
        out->cigar = cig->seq;
        out->cigarLen = cig->len;
        free(cig);
    } else if (cig) {
        parasail_cigar_free(cig);
    }
    if (fin) parasail_result_free(fin);
    parasail_result_free(rev);
    parasail_result_free(fwd);
    return out;
}

/* parasail_ssw (src/pssw.c:45-54) */
PHOST_EXPORT parasail_result_ssw_t *parasail_ssw(const char *const s1, const int s1Len,
                                                 const char *const s2, const int s2Len,
                                                 const int open, const int gap,
                                                 const parasail_matrix_t *matrix)
{
    parasail_profile_t *p = parasail_ssw_init(s1, s1Len, matrix, 2);
    if (!p) return NULL;
    parasail_result_ssw_t *r = parasail_ssw_profile(p, s2, s2Len, open, gap);
    parasail_profile_free(p);
    return r;
}

PHOST_EXPORT void parasail_result_ssw_free(parasail_result_ssw_t *result)
{
    if (!result) return;
    free(result->cigar);
    free(result);
}
