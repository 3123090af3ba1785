/*
 * cigar.c -- SAM CIGAR from a trace table (src/cigar.c, the walk of
 * src/cigar_template.c:58-263).
 *
 * The walk starts at (end_query, end_ref) in state DIAG and follows the
 * trace codes of parasail.h:31-41 backwards: DIAG emits '='/'X' and moves
 * diagonally; INS (target consumed) and DEL (query consumed) states extend or
 * close by the E/F bits of the cell they leave.  I and D are emitted swapped,
 * as the reference does (cigar_template.c:38-56).  Semi-global results are
 * prefixed with the free end gap (cigar_template.c:121-142).  The table is
 * read in whichever layout the result's flags name: row-major int8
 * (LOC_NOVEC) or striped with the BITS and LANES flags (LOC_STRIPED,
 * cigar.c:156-157).
 */
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

static const char kOps[] = "MIDNSHP=XB";

uint32_t parasail_cigar_encode(uint32_t length, char op_letter)
{
    uint32_t code = 0;
    const char *p = strchr(kOps, op_letter);
    if (p && op_letter) code = (uint32_t)(p - kOps);
    return (length << 4) | code;
}

char parasail_cigar_decode_op(uint32_t cigar_int)
{
    return (cigar_int & 0xfU) > 9 ? 'M' : kOps[cigar_int & 0xfU];
}

uint32_t parasail_cigar_decode_len(uint32_t cigar_int) { return cigar_int >> 4; }

/* parasail_cigar_encode_string, src/cigar.c:63-103 */
parasail_cigar_t *parasail_cigar_encode_string(const char *cigar)
{
    if (!cigar) return NULL;
    size_t n = strlen(cigar);
    parasail_cigar_t *c = (parasail_cigar_t *)malloc(sizeof(parasail_cigar_t));
    if (!c) return NULL;
    c->seq = (uint32_t *)malloc(sizeof(uint32_t) * (n + 1));
    c->len = 0;
    c->beg_query = 0;
    c->beg_ref = 0;
    size_t off = 0;
    while (off < n) {
        unsigned int len = 0;
        char op = 'M';
        int used = 0;
        if (sscanf(cigar + off, "%u%c%n", &len, &op, &used) != 2) {
            fprintf(stderr, "invalid CIGAR string\n");
            parasail_cigar_free(c);
            return NULL;
        }
        off += (size_t)used;
        c->seq[c->len++] = parasail_cigar_encode(len, op);
    }
    return c;
}

/* parasail_cigar_decode, src/cigar.c:113-145 */
char *parasail_cigar_decode(parasail_cigar_t *cigar)
{
    if (!cigar) return NULL;
    size_t cap = (size_t)cigar->len * 12 + 1, used = 0;
    char *s = (char *)malloc(cap);
    if (!s) return NULL;
    s[0] = '\0';
    for (int i = 0; i < cigar->len; ++i) {
        int w = snprintf(s + used, cap - used, "%u%c", parasail_cigar_decode_len(cigar->seq[i]),
                         parasail_cigar_decode_op(cigar->seq[i]));
        if (w < 0 || (size_t)w >= cap - used) {
            fprintf(stderr, "invalid CIGAR\n");
            free(s);
            return NULL;
        }
        used += (size_t)w;
    }
    return s;
}

void parasail_cigar_free(parasail_cigar_t *cigar)
{
    if (!cigar) return;
    free(cigar->seq);
    free(cigar);
}

typedef struct {
    const void *tab;
    int striped, bits;
    long seglen, segwidth, lenb;
} tview_t;

static inline int code_at(const tview_t *v, long i, long j)
{
    if (!v->striped) return ((const int8_t *)v->tab)[i * v->lenb + j];
    long loc = j * v->seglen * v->segwidth + (i % v->seglen) * v->segwidth + i / v->seglen;
    return phost_trace_at(v->tab, v->bits, loc);
}

typedef struct {
    uint32_t *seq;
    int len, cap;
    uint32_t mat, mis, del, ins;
} run_t;

static int push(run_t *r, uint32_t cnt, char op)
{
    if (r->len >= r->cap) {
        int ncap = r->cap * 2 + 8;
        uint32_t *s = (uint32_t *)realloc(r->seq, sizeof(uint32_t) * (size_t)ncap);
        if (!s) return -1;
        r->seq = s;
        r->cap = ncap;
    }
    r->seq[r->len++] = parasail_cigar_encode(cnt, op);
    return 0;
}

static void flush(run_t *r)
{
    if (r->mat) push(r, r->mat, '=');
    else if (r->mis) push(r, r->mis, 'X');
    else if (r->del) push(r, r->del, 'I');
    else if (r->ins) push(r, r->ins, 'D');
    r->mat = r->mis = r->del = r->ins = 0;
}

static parasail_cigar_t *walk(const tview_t *v, const char *seqA, int lena, const char *seqB,
                              int lenb, int is_sg, int end_query, int end_ref, int case_sensitive,
                              const char *aliases_in)
{
    char aliases[256];
    size_t nal = 0;
    if (aliases_in) {
        nal = strlen(aliases_in);
        if (nal % 2 != 0 || nal >= 256) {
            fprintf(stderr, "%s: assert(%s) failed\n", __func__, "aliases_size % 2 == 0");
            return NULL;
        }
        for (size_t a = 0; a < nal; ++a)
            aliases[a] = case_sensitive ? aliases_in[a] : (char)toupper((unsigned char)aliases_in[a]);
    }
    run_t r;
    memset(&r, 0, sizeof(r));
    r.cap = lena + lenb + 1;
    r.seq = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)r.cap);
    if (!r.seq) return NULL;
    long i = end_query, j = end_ref;
    int where = PARASAIL_DIAG;
    if (is_sg) {
        if (end_query + 1 == lena) {
            r.ins += (uint32_t)(lenb - 1 - j > 0 ? lenb - 1 - j : 0);
        } else if (end_ref + 1 == lenb) {
            r.del += (uint32_t)(lena - 1 - i > 0 ? lena - 1 - i : 0);
        } else {
            free(r.seq);
            return NULL;
        }
    }
    while (i >= 0 || j >= 0) {
        if (i < 0) {
            if (r.ins == 0) flush(&r);
            r.ins += (uint32_t)(j + 1);
            j = -1;
            break;
        }
        if (j < 0) {
            if (r.del == 0) flush(&r);
            r.del += (uint32_t)(i + 1);
            i = -1;
            break;
        }
        const int code = code_at(v, i, j);
        if (where == PARASAIL_DIAG) {
            if (code & PARASAIL_DIAG) {
                char a = case_sensitive ? seqA[i] : (char)toupper((unsigned char)seqA[i]);
                char b = case_sensitive ? seqB[j] : (char)toupper((unsigned char)seqB[j]);
                int same = a == b;
                for (size_t k = 0; aliases_in && k < nal; k += 2) {
                    if (aliases[k] == a) same |= aliases[k + 1] == b;
                    else if (aliases[k + 1] == a) same |= aliases[k] == b;
                }
                if (same) {
                    if (r.mat == 0) flush(&r);
                    r.mat += 1;
                } else {
                    if (r.mis == 0) flush(&r);
                    r.mis += 1;
                }
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
            if (r.ins == 0) flush(&r);
            r.ins += 1;
            --j;
            if (code & PARASAIL_DIAG_E) where = PARASAIL_DIAG;
            else if (!(code & PARASAIL_INS_E)) { free(r.seq); return NULL; }
        } else {
            if (r.del == 0) flush(&r);
            r.del += 1;
            --i;
            if (code & PARASAIL_DIAG_F) where = PARASAIL_DIAG;
            else if (!(code & PARASAIL_DEL_F)) { free(r.seq); return NULL; }
        }
    }
    flush(&r);
    parasail_cigar_t *c = (parasail_cigar_t *)malloc(sizeof(parasail_cigar_t));
    if (!c) { free(r.seq); return NULL; }
    for (int a = 0, z = r.len - 1; a < z; ++a, --z) {
        uint32_t t = r.seq[a];
        r.seq[a] = r.seq[z];
        r.seq[z] = t;
    }
    c->seq = r.seq;
    c->len = r.len;
    c->beg_query = (int)(i + 1);
    c->beg_ref = (int)(j + 1);
    return c;
}

/* parasail_result_get_cigar_extra, src/cigar.c:199-231 */
parasail_cigar_t *parasail_result_get_cigar_extra(parasail_result_t *result, const char *seqA,
                                                  int lena, const char *seqB, int lenb,
                                                  const parasail_matrix_t *matrix,
                                                  int case_sensitive, const char *alphabet_aliases)
{
    (void)matrix;
    if (!result || !(result->flag & PARASAIL_FLAG_TRACE)) {
        fprintf(stderr, "%s: assert(%s) failed\n", __func__, "parasail_result_is_trace(result)");
        return NULL;
    }
    tview_t v;
    v.tab = result->trace->trace_table;
    v.striped = (result->flag & (PARASAIL_FLAG_STRIPED | PARASAIL_FLAG_SCAN)) != 0;
    v.bits = v.striped ? phost_flag_bits(result->flag) : 8;
    v.segwidth = phost_flag_lanes(result->flag);
    v.seglen = (lena + v.segwidth - 1) / v.segwidth;
    v.lenb = lenb;
    return walk(&v, seqA, lena, seqB, lenb, (result->flag & PARASAIL_FLAG_SG) != 0,
                result->end_query, result->end_ref, case_sensitive, alphabet_aliases);
}

parasail_cigar_t *parasail_result_get_cigar(parasail_result_t *result, const char *seqA, int lena,
                                            const char *seqB, int lenb,
                                            const parasail_matrix_t *matrix)
{
    return parasail_result_get_cigar_extra(result, seqA, lena, seqB, lenb, matrix, 0, NULL);
}

parasail_cigar_t *phost_cigar_rowmajor(const char *seqA, int lena, const char *seqB, int lenb,
                                       int is_sg, int end_query, int end_ref, const int8_t *tr)
{
    tview_t v;
    v.tab = tr;
    v.striped = 0;
    v.bits = 8;
    v.segwidth = 1;
    v.seglen = lena;
    v.lenb = lenb;
    return walk(&v, seqA, lena, seqB, lenb, is_sg, end_query, end_ref, 0, NULL);
}
