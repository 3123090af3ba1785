/*
 * batch.c -- parasail_hip_align_batch: many pairs in one device pass.
 *
 * The reference has no batched entry point; its callers loop over pairs
 * with OpenMP (tests/test_openmp.c:383-441, apps/parasail_aligner.cpp:
 * 1606-1625).  Here the loop moves onto the device: one call maps the
 * residues, ships the batch to HBM once and aligns every pair there.
 * Results are the ones the per-pair hip functions return, in flat arrays.
 */
#include <ctype.h>
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

static int parse_alg(const char *name, int *alg, int *flags)
{
    static const struct { const char *n; int alg, flags; } tab[] = {
        {"sw", PHIP_SW, 0},
        {"nw", PHIP_NW, 0},
        {"sg", PHIP_SG, PHIP_S1_BEG | PHIP_S1_END | PHIP_S2_BEG | PHIP_S2_END},
        {"sg_qb", PHIP_SG, PHIP_S1_BEG},
        {"sg_qe", PHIP_SG, PHIP_S1_END},
        {"sg_qx", PHIP_SG, PHIP_S1_BEG | PHIP_S1_END},
        {"sg_db", PHIP_SG, PHIP_S2_BEG},
        {"sg_de", PHIP_SG, PHIP_S2_END},
        {"sg_dx", PHIP_SG, PHIP_S2_BEG | PHIP_S2_END},
        {"sg_qb_de", PHIP_SG, PHIP_S1_BEG | PHIP_S2_END},
        {"sg_qe_db", PHIP_SG, PHIP_S1_END | PHIP_S2_BEG},
        {"sg_qb_db", PHIP_SG, PHIP_S1_BEG | PHIP_S2_BEG},
        {"sg_qe_de", PHIP_SG, PHIP_S1_END | PHIP_S2_END},
    };
    for (size_t i = 0; i < sizeof(tab) / sizeof(tab[0]); ++i)
        if (strcmp(name, tab[i].n) == 0) {
            *alg = tab[i].alg;
            *flags = tab[i].flags;
            return 0;
        }
    return -1;
}

static char g_err[256];

/* the device results of [lo, hi) into the caller's arrays and CIGARs, with
 * the flags and saturation the reference's functions report */
typedef struct {
    const phip_out_t *res;
    parasail_hip_batch_result_t *out;
    const parasail_matrix_t *matrix;
    const phip_cigars_t *cg;
    int alg, flags, kind, width, open, dev_cigar;
} convert_ctx_t;

static void convert_range(void *ctx, int lo, int hi)
{
    const convert_ctx_t *x = (const convert_ctx_t *)ctx;
    const int alg = x->alg, flags = x->flags, kind = x->kind, width = x->width, open = x->open;
    const parasail_matrix_t *matrix = x->matrix;
    parasail_hip_batch_result_t *out = x->out;
    for (int k = lo; k < hi; ++k) {
        const phip_out_t *o = &x->res[k];
        int w = width, sat = 0;
        if (w == 0) {
            /* the narrowest width whose window holds (satcheck.c tries 8,
             * then 16, then 32), from the exact score (SW) or the extrema
             * the device returned (NW/SG, _stats) */
            w = !phost_saturates(alg, kind, 0, 8, open, matrix, o) ? 8
                : !phost_saturates(alg, kind, 0, 16, open, matrix, o) ? 16 : 32;
        } else {
            sat = phost_saturates(alg, kind, 0, w, open, matrix, o);
        }
        int f = alg == PHIP_SW ? PARASAIL_FLAG_SW : alg == PHIP_NW ? PARASAIL_FLAG_NW : PARASAIL_FLAG_SG;
        if (alg == PHIP_SG) {
            if (flags & PHIP_S1_BEG) f |= PARASAIL_FLAG_SG_S1_BEG;
            if (flags & PHIP_S1_END) f |= PARASAIL_FLAG_SG_S1_END;
            if (flags & PHIP_S2_BEG) f |= PARASAIL_FLAG_SG_S2_BEG;
            if (flags & PHIP_S2_END) f |= PARASAIL_FLAG_SG_S2_END;
        }
        f |= PARASAIL_FLAG_STRIPED | PARASAIL_FLAG_LANES_64;
        f |= w == 8 ? PARASAIL_FLAG_BITS_8 : w == 16 ? PARASAIL_FLAG_BITS_16 : PARASAIL_FLAG_BITS_32;
        if (kind == PHIP_STATS) f |= PARASAIL_FLAG_STATS;
        if (kind == PHIP_TRACE) f |= PARASAIL_FLAG_TRACE;
        if (sat) f |= PARASAIL_FLAG_SATURATED;
        const int sv = sat ? phost_saturated_value(alg, kind, 0, w) : 0;
        if (out->score) out->score[k] = sat ? sv : o->score;
        if (out->end_query) out->end_query[k] = sat ? 0 : o->end_query;
        if (out->end_ref) out->end_ref[k] = sat ? 0 : o->end_ref;
        if (out->matches) out->matches[k] = sat ? (kind == PHIP_STATS ? sv : 0) : o->matches;
        if (out->similar) out->similar[k] = sat ? (kind == PHIP_STATS ? sv : 0) : o->similar;
        if (out->length) out->length[k] = sat ? (kind == PHIP_STATS ? sv : 0) : o->length;
        if (out->flag) out->flag[k] = f;
        if (x->dev_cigar) {
            const phip_cigars_t *cg = x->cg;
            parasail_cigar_t *c = NULL;
            if (!sat && cg->nops[k] >= 0) {
                c = (parasail_cigar_t *)malloc(sizeof(parasail_cigar_t));
                uint32_t *seq = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(cg->nops[k] + 1));
                if (!c || !seq) {
                    free(c);
                    free(seq);
                    c = NULL;
                } else {
                    memcpy(seq, cg->ops + cg->off[k], sizeof(uint32_t) * (size_t)cg->nops[k]);
                    c->seq = seq;
                    c->len = cg->nops[k];
                    c->beg_query = cg->beg[2 * k];
                    c->beg_ref = cg->beg[2 * k + 1];
                }
            }
            out->cigars[k] = c;
        }
    }
}

PHOST_EXPORT int parasail_hip_align_batch(const char *alg_name, int kind, int width,
                                          const char *s1, const int64_t *s1_off,
                                          const int *s1_len, const char *s2,
                                          const int64_t *s2_off, const int *s2_len, int npairs,
                                          int open, int gap, const parasail_matrix_t *matrix,
                                          parasail_hip_batch_result_t *out)
{
    int alg, flags;
    if (!alg_name || parse_alg(alg_name, &alg, &flags) != 0) {
        fprintf(stderr, "%s: unknown algorithm %s\n", __func__, alg_name ? alg_name : "(null)");
        return -1;
    }
    if (!matrix || !s2 || !s2_off || !s2_len || npairs < 0 || open < 0 || gap < 0 || !out ||
        kind < 0 || kind > 2 || (width != 0 && width != 8 && width != 16 && width != 32)) {
        fprintf(stderr, "%s: invalid arguments\n", __func__);
        return -1;
    }
    const int pssm = matrix->type == PARASAIL_MATRIX_TYPE_PSSM;
    if (!pssm && (!s1 || !s1_off || !s1_len)) {
        fprintf(stderr, "%s: missing s1\n", __func__);
        return -1;
    }
    if (kind == PHIP_STATS && !s1) {
        fprintf(stderr, "%s: stats functions with pssm require a consensus sequence\n", __func__);
        return -1;
    }
    if (kind == PHIP_TRACE && out->cigars && !s1) {
        fprintf(stderr, "%s: CIGAR strings require the query sequence\n", __func__);
        return -1;
    }
    if (phip_init() != 0) {
        errno = ENOSYS;
        return -1;
    }
    if (npairs == 0) return 0;
    int64_t qtot = 0, ttot = 0;
    int32_t *ql = (int32_t *)malloc(sizeof(int32_t) * (size_t)npairs);
    int32_t *tl = (int32_t *)malloc(sizeof(int32_t) * (size_t)npairs);
    int64_t *qo = (int64_t *)malloc(sizeof(int64_t) * (size_t)npairs);
    int64_t *to = (int64_t *)malloc(sizeof(int64_t) * (size_t)npairs);
    /* the results: in the pinned staging buffer with the codes when it can
     * be had (the download then runs at DMA speed), else ordinary memory */
    phip_out_t *res = NULL;
    int res_owned = 0;
    if (!ql || !tl || !qo || !to) goto oom;
    for (int k = 0; k < npairs; ++k) {
        ql[k] = pssm ? matrix->length : s1_len[k];
        tl[k] = s2_len[k];
        if (ql[k] <= 0 || tl[k] <= 0) {
            fprintf(stderr, "%s: pair %d has an empty sequence\n", __func__, k);
            free(ql); free(tl); free(qo); free(to);
            return -1;
        }
        qo[k] = qtot;
        to[k] = ttot;
        qtot += s1 ? ql[k] : 0;
        ttot += tl[k];
    }
    {
        /* CIGARs: traceback and walk on the device, from the residues as
         * given, upper-cased (cigar_template.c compares them) */
        const int dev_cigar = kind == PHIP_TRACE && out->cigars;
        /* the codes (and upper-cased residues) go to the runtime's pinned
         * staging buffer when it can be had, else to ordinary memory */
        const size_t sq = ((size_t)(s1 ? qtot + 1 : 0) + 255) & ~(size_t)255;
        const size_t st = ((size_t)ttot + 1 + 255) & ~(size_t)255;
        const size_t scodes = (sq + st) * (dev_cigar ? 2 : 1);
        uint8_t *stage =
            (uint8_t *)phip_stage_acquire(scodes + sizeof(phip_out_t) * (size_t)npairs);
        uint8_t *q = NULL, *t = NULL, *qraw = NULL, *traw = NULL;
        if (stage) {
            q = s1 ? stage : NULL;
            t = stage + sq;
            if (dev_cigar) {
                qraw = stage + sq + st;
                traw = qraw + sq;
            }
            res = (phip_out_t *)(stage + scodes);
        } else {
            q = s1 ? (uint8_t *)malloc((size_t)qtot + 1) : NULL;
            t = (uint8_t *)malloc((size_t)ttot + 1);
            if (dev_cigar) {
                qraw = (uint8_t *)malloc((size_t)qtot + 1);
                traw = (uint8_t *)malloc((size_t)ttot + 1);
            }
            res = (phip_out_t *)malloc(sizeof(phip_out_t) * (size_t)npairs);
            res_owned = 1;
        }
        phip_cigars_t cg;
        memset(&cg, 0, sizeof(cg));
        if ((s1 && !q) || !t || (dev_cigar && (!qraw || !traw)) || !res) {
            if (stage) phip_stage_release(stage);
            else { free(q); free(t); free(qraw); free(traw); free(res); }
            res = NULL;
            goto oom;
        }
        if (dev_cigar) {
            int64_t otot = 0;
            for (int k = 0; k < npairs; ++k) otot += (int64_t)ql[k] + tl[k] + 2;
            cg.ops = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(otot + 1));
            cg.off = (int64_t *)malloc(sizeof(int64_t) * (size_t)npairs);
            cg.nops = (int32_t *)malloc(sizeof(int32_t) * (size_t)npairs);
            cg.beg = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)npairs);
            if (!cg.ops || !cg.off || !cg.nops || !cg.beg) {
                free(cg.ops); free(cg.off); free(cg.nops); free(cg.beg);
                if (stage) phip_stage_release(stage);
                else { free(q); free(t); free(qraw); free(traw); free(res); }
                res = NULL;
                goto oom;
            }
        }
        {
            /* codes, and the upper-cased residues for device CIGARs */
            phost_enc_t e = {s2, s2_off, tl, to, matrix->mapper, t, traw};
            phost_encode(&e, npairs);
            if (s1) {
                phost_enc_t f = {s1, s1_off, ql, qo, matrix->mapper, q, qraw};
                phost_encode(&f, npairs);
            }
        }
        phip_batch_t b;
        memset(&b, 0, sizeof(b));
        b.alg = alg;
        b.kind = kind;
        b.sg_flags = flags;
        b.open = open;
        b.gap = gap;
        /* the extrema the flags need */
        b.want_sat_check = phost_check_width(alg, kind, 0, width);
        b.want_min_h = kind == PHIP_STATS;
        b.matrix.values = matrix->matrix;
        b.matrix.size = matrix->size;
        b.matrix.length = pssm ? matrix->length : matrix->size;
        b.matrix.is_pssm = pssm;
        b.matrix.max = matrix->max;
        b.matrix.min = matrix->min;
        b.npairs = npairs;
        b.q = q;
        b.qoff = qo;
        b.qlen = ql;
        b.t = t;
        b.toff = to;
        b.tlen = tl;
        int rc = dev_cigar ? phip_align_cigars(&b, qraw, traw, res, &cg) : phip_align(&b, res, NULL);
        if (!stage) { free(q); free(t); free(qraw); free(traw); }
        if (rc != 0) {
            snprintf(g_err, sizeof(g_err), "%s", phip_last_error());
            fprintf(stderr, "%s: %s\n", __func__, g_err);
            free(cg.ops); free(cg.off); free(cg.nops); free(cg.beg);
            if (stage) phip_stage_release(stage);
            if (res_owned) free(res);
            free(ql); free(tl); free(qo); free(to);
            return -1;
        }
        {
            convert_ctx_t cc = {res, out, matrix, &cg, alg, flags, kind, width, open, dev_cigar};
            /* the CIGAR objects are two mallocs each: several threads share them */
            phost_parallel_for(npairs, dev_cigar ? 25000 : 50000, convert_range, &cc);
        }
        /* the results lived in the staging buffer until converted */
        if (stage) phip_stage_release(stage);
        if (res_owned) free(res);
        free(cg.ops); free(cg.off); free(cg.nops); free(cg.beg);
    }
    free(ql); free(tl); free(qo); free(to);
    return 0;
oom:
    free(ql); free(tl); free(qo); free(to);
    fprintf(stderr, "%s: out of host memory\n", __func__);
    return -1;
}

/* ---- resident batches ---- */

struct parasail_hip_batch {
    void *dev;
    int npairs;
    phip_matrix_t mx;
    const parasail_matrix_t *matrix;
};

PHOST_EXPORT parasail_hip_batch_t *parasail_hip_batch_create(const char *s1, const int64_t *s1_off,
                                                             const int *s1_len, const char *s2,
                                                             const int64_t *s2_off,
                                                             const int *s2_len, int npairs,
                                                             const parasail_matrix_t *matrix)
{
    if (!matrix || !s1 || !s1_off || !s1_len || !s2 || !s2_off || !s2_len || npairs <= 0) {
        fprintf(stderr, "%s: invalid arguments\n", __func__);
        return NULL;
    }
    if (matrix->type == PARASAIL_MATRIX_TYPE_PSSM) {
        fprintf(stderr, "%s: resident batches take a square matrix\n", __func__);
        return NULL;
    }
    if (phip_init() != 0) {
        errno = ENOSYS;
        return NULL;
    }
    parasail_hip_batch_t *hb = (parasail_hip_batch_t *)calloc(1, sizeof(*hb));
    int32_t *ql = (int32_t *)malloc(sizeof(int32_t) * (size_t)npairs);
    int32_t *tl = (int32_t *)malloc(sizeof(int32_t) * (size_t)npairs);
    int64_t *qo = (int64_t *)malloc(sizeof(int64_t) * (size_t)npairs);
    int64_t *to = (int64_t *)malloc(sizeof(int64_t) * (size_t)npairs);
    uint8_t *q = NULL, *t = NULL;
    int64_t qtot = 0, ttot = 0;
    if (!hb || !ql || !tl || !qo || !to) goto fail;
    for (int k = 0; k < npairs; ++k) {
        if (s1_len[k] <= 0 || s2_len[k] <= 0) {
            fprintf(stderr, "%s: pair %d has an empty sequence\n", __func__, k);
            goto fail;
        }
        ql[k] = s1_len[k];
        tl[k] = s2_len[k];
        qo[k] = qtot;
        to[k] = ttot;
        qtot += ql[k];
        ttot += tl[k];
    }
    q = (uint8_t *)malloc((size_t)qtot);
    t = (uint8_t *)malloc((size_t)ttot);
    if (!q || !t) goto fail;
    {
        phost_enc_t e = {s1, s1_off, ql, qo, matrix->mapper, q, NULL};
        phost_encode(&e, npairs);
        phost_enc_t f = {s2, s2_off, tl, to, matrix->mapper, t, NULL};
        phost_encode(&f, npairs);
    }
    {
        phip_batch_t b;
        memset(&b, 0, sizeof(b));
        b.kind = PHIP_SCORE;
        b.matrix.values = matrix->matrix;
        b.matrix.size = matrix->size;
        b.matrix.length = matrix->size;
        b.matrix.max = matrix->max;
        b.matrix.min = matrix->min;
        b.npairs = npairs;
        b.q = q;
        b.qoff = qo;
        b.qlen = ql;
        b.t = t;
        b.toff = to;
        b.tlen = tl;
        hb->dev = phip_resident_create(&b);
    }
    if (!hb->dev) {
        fprintf(stderr, "%s: %s\n", __func__, phip_last_error());
        goto fail;
    }
    hb->npairs = npairs;
    hb->matrix = matrix;
    free(q); free(t); free(ql); free(tl); free(qo); free(to);
    return hb;
fail:
    free(hb); free(q); free(t); free(ql); free(tl); free(qo); free(to);
    return NULL;
}

/* the device's 16-byte results into the caller's arrays, with the width
 * flag of each pair */
typedef struct {
    const phip_out4_t *res;
    parasail_hip_batch_result_t *out;
    const parasail_matrix_t *matrix;
    int alg, open, f0;
} unpack_ctx_t;

static void unpack_range(void *ctx, int lo, int hi)
{
    const unpack_ctx_t *c = (const unpack_ctx_t *)ctx;
    int *score = c->out->score, *eq = c->out->end_query, *er = c->out->end_ref, *fl = c->out->flag;
    for (int k = lo; k < hi; ++k) {
        const phip_out4_t *o = &c->res[k];
        int w = (o->status & PHIP_ST_W32) ? 32 : 16;
        if (c->alg == PHIP_SW) {
            phip_out_t full;
            memset(&full, 0, sizeof(full));
            full.score = o->score; /* the SW width test reads the score only */
            w = !phost_saturates(c->alg, PHIP_SCORE, 0, 8, c->open, c->matrix, &full) ? 8
                : !phost_saturates(c->alg, PHIP_SCORE, 0, 16, c->open, c->matrix, &full) ? 16 : 32;
        }
        if (score) score[k] = o->score;
        if (eq) eq[k] = o->end_query;
        if (er) er[k] = o->end_ref;
        if (fl)
            fl[k] = c->f0 | (w == 8 ? PARASAIL_FLAG_BITS_8
                             : w == 16 ? PARASAIL_FLAG_BITS_16 : PARASAIL_FLAG_BITS_32);
    }
}

PHOST_EXPORT int parasail_hip_batch_align(parasail_hip_batch_t *hb, const char *alg_name, int open,
                                          int gap, parasail_hip_batch_result_t *out)
{
    int alg, flags;
    if (!hb || !alg_name || parse_alg(alg_name, &alg, &flags) != 0 || open < 0 || gap < 0) {
        fprintf(stderr, "%s: invalid arguments\n", __func__);
        return -1;
    }
    /* score, ends and status of each pair, 16 bytes a pair, in the device
     * batch's pinned buffer */
    const phip_out4_t *res = NULL;
    if (phip_resident_align4(hb->dev, alg, flags, open, gap, &res) != 0) {
        fprintf(stderr, "%s: %s\n", __func__, phip_last_error());
        return -1;
    }
    if (!out) return 0;
    int f0 = alg == PHIP_SW ? PARASAIL_FLAG_SW : alg == PHIP_NW ? PARASAIL_FLAG_NW : PARASAIL_FLAG_SG;
    if (alg == PHIP_SG) {
        if (flags & PHIP_S1_BEG) f0 |= PARASAIL_FLAG_SG_S1_BEG;
        if (flags & PHIP_S1_END) f0 |= PARASAIL_FLAG_SG_S1_END;
        if (flags & PHIP_S2_BEG) f0 |= PARASAIL_FLAG_SG_S2_BEG;
        if (flags & PHIP_S2_END) f0 |= PARASAIL_FLAG_SG_S2_END;
    }
    f0 |= PARASAIL_FLAG_STRIPED | PARASAIL_FLAG_LANES_64;
    unpack_ctx_t c = {res, out, hb->matrix, alg, open, f0};
    phost_parallel_for(hb->npairs, 50000, unpack_range, &c);
    return 0;
}

/* per pass of the last parasail_hip_batch_align_multi: all kernels, main kernel */
static double g_pass_ms[4][2];
static int g_npass_ms = 0;

PHOST_EXPORT int parasail_hip_batch_align_multi(parasail_hip_batch_t *hb, int nalg,
                                                const char *const *alg_names, int open, int gap,
                                                parasail_hip_batch_result_t *out)
{
    int alg[4], flags[4];
    if (!hb || nalg < 1 || nalg > 4 || !alg_names || !out || open < 0 || gap < 0) {
        fprintf(stderr, "%s: invalid arguments\n", __func__);
        return -1;
    }
    for (int k = 0; k < nalg; ++k)
        if (!alg_names[k] || parse_alg(alg_names[k], &alg[k], &flags[k]) != 0) {
            fprintf(stderr, "%s: invalid arguments\n", __func__);
            return -1;
        }
    if (phip_resident_launch4(hb->dev, nalg, alg, flags, open, gap) != 0) {
        fprintf(stderr, "%s: %s\n", __func__, phip_last_error());
        return -1;
    }
    int rc = 0;
    g_npass_ms = 0;
    for (int k = 0; k < nalg && rc == 0; ++k) {
        const phip_out4_t *res = NULL;
        if (phip_resident_wait4(hb->dev, k, &res, &g_pass_ms[k][0], &g_pass_ms[k][1]) != 0) {
            fprintf(stderr, "%s: %s\n", __func__, phip_last_error());
            rc = -1;
            break;
        }
        g_npass_ms = k + 1;
        int f0 = alg[k] == PHIP_SW ? PARASAIL_FLAG_SW : alg[k] == PHIP_NW ? PARASAIL_FLAG_NW : PARASAIL_FLAG_SG;
        if (alg[k] == PHIP_SG) {
            if (flags[k] & PHIP_S1_BEG) f0 |= PARASAIL_FLAG_SG_S1_BEG;
            if (flags[k] & PHIP_S1_END) f0 |= PARASAIL_FLAG_SG_S1_END;
            if (flags[k] & PHIP_S2_BEG) f0 |= PARASAIL_FLAG_SG_S2_BEG;
            if (flags[k] & PHIP_S2_END) f0 |= PARASAIL_FLAG_SG_S2_END;
        }
        f0 |= PARASAIL_FLAG_STRIPED | PARASAIL_FLAG_LANES_64;
        unpack_ctx_t c = {res, &out[k], hb->matrix, alg[k], open, f0};
        phost_parallel_for(hb->npairs, 50000, unpack_range, &c);
    }
    phip_resident_end4(hb->dev);
    return rc;
}

PHOST_EXPORT int parasail_hip_last_pass_ms(int k, double *kernel_ms, double *main_kernel_ms)
{
    if (k < 0 || k >= g_npass_ms) return -1;
    if (kernel_ms) *kernel_ms = g_pass_ms[k][0];
    if (main_kernel_ms) *main_kernel_ms = g_pass_ms[k][1];
    return 0;
}

PHOST_EXPORT void parasail_hip_batch_free(parasail_hip_batch_t *hb)
{
    if (!hb) return;
    phip_resident_free(hb->dev);
    free(hb);
}

PHOST_EXPORT int parasail_hip_autotune(const parasail_matrix_t *matrix, int open, int gap)
{
    if (!matrix || matrix->type == PARASAIL_MATRIX_TYPE_PSSM || open < 0 || gap < 0) {
        fprintf(stderr, "%s: invalid arguments\n", __func__);
        return -1;
    }
    phip_matrix_t mx;
    mx.values = matrix->matrix;
    mx.size = matrix->size;
    mx.length = matrix->size;
    mx.is_pssm = 0;
    mx.max = matrix->max;
    mx.min = matrix->min;
    if (phip_wave_autotune(&mx, open, gap) != 0) {
        fprintf(stderr, "%s: %s\n", __func__, phip_last_error());
        return -1;
    }
    return 0;
}

PHOST_EXPORT int parasail_hip_autotune_table(int *rows, double *sw_cost, double *nw_cost, int cap)
{
    return phip_wave_autotune_table(rows, sw_cost, nw_cost, cap);
}

PHOST_EXPORT void parasail_hip_autotune_reset(void) { phip_wave_autotune_reset(); }

PHOST_EXPORT int parasail_hip_batch_set_graph(parasail_hip_batch_t *hb, int on)
{
    return hb ? phip_resident_set_graph(hb->dev, on) : -1;
}

PHOST_EXPORT int parasail_hip_batch_graph_stats(parasail_hip_batch_t *hb, int *captures, int *replays)
{
    return hb ? phip_resident_graph_stats(hb->dev, captures, replays) : -1;
}

/* ---- banded global alignment of many pairs ---- */

PHOST_EXPORT int parasail_hip_nw_banded_batch(const char *s1, const int64_t *s1_off,
                                              const int *s1_len, const char *s2,
                                              const int64_t *s2_off, const int *s2_len,
                                              int npairs, int open, int gap, int k,
                                              const parasail_matrix_t *matrix,
                                              parasail_hip_batch_result_t *out)
{
    if (!s1 || !s1_off || !s1_len || !s2 || !s2_off || !s2_len || npairs < 0 || open < 0 ||
        gap < 0 || k <= 0 || !matrix || !out) {
        fprintf(stderr, "%s: invalid arguments\n", __func__);
        return -1;
    }
    if (phip_init() != 0) {
        errno = ENOSYS;
        return -1;
    }
    if (npairs == 0) return 0;
    const int pssm = matrix->type == PARASAIL_MATRIX_TYPE_PSSM;
    int64_t qtot = 0, ttot = 0;
    int32_t *ql = (int32_t *)malloc(sizeof(int32_t) * (size_t)npairs);
    int32_t *tl = (int32_t *)malloc(sizeof(int32_t) * (size_t)npairs);
    int64_t *qo = (int64_t *)malloc(sizeof(int64_t) * (size_t)npairs);
    int64_t *to = (int64_t *)malloc(sizeof(int64_t) * (size_t)npairs);
    phip_out_t *res = (phip_out_t *)malloc(sizeof(phip_out_t) * (size_t)npairs);
    uint8_t *q = NULL, *t = NULL, *stage = NULL;
    int rc = -1;
    if (!ql || !tl || !qo || !to || !res) goto done;
    for (int p = 0; p < npairs; ++p) {
        ql[p] = s1_len[p];
        tl[p] = s2_len[p];
        if (ql[p] <= 0 || tl[p] <= 0) {
            fprintf(stderr, "%s: pair %d has an empty sequence\n", __func__, p);
            goto done;
        }
        if (pssm && ql[p] != matrix->length) {
            fprintf(stderr, "%s: pair %d: query length differs from the PSSM's\n", __func__, p);
            goto done;
        }
        qo[p] = qtot;
        to[p] = ttot;
        qtot += ql[p];
        ttot += tl[p];
    }
    {
        const size_t sq = ((size_t)qtot + 1 + 255) & ~(size_t)255;
        stage = (uint8_t *)phip_stage_acquire(sq + (size_t)ttot + 1);
        if (stage) {
            q = stage;
            t = stage + sq;
        } else {
            q = (uint8_t *)malloc((size_t)qtot + 1);
            t = (uint8_t *)malloc((size_t)ttot + 1);
        }
    }
    if (!q || !t) goto done;
    {
        phost_enc_t e = {s1, s1_off, ql, qo, matrix->mapper, q, NULL};
        phost_encode(&e, npairs);
        phost_enc_t f = {s2, s2_off, tl, to, matrix->mapper, t, NULL};
        phost_encode(&f, npairs);
    }
    {
        phip_batch_t bt;
        memset(&bt, 0, sizeof(bt));
        bt.alg = PHIP_NW;
        bt.open = open;
        bt.gap = gap;
        bt.matrix.values = matrix->matrix;
        bt.matrix.size = matrix->size;
        bt.matrix.length = pssm ? matrix->length : matrix->size;
        bt.matrix.is_pssm = pssm;
        bt.matrix.max = matrix->max;
        bt.matrix.min = matrix->min;
        bt.npairs = npairs;
        bt.q = q;
        bt.qoff = qo;
        bt.qlen = ql;
        bt.t = t;
        bt.toff = to;
        bt.tlen = tl;
        if (phip_nw_banded_batch(&bt, k, res) != 0) {
            fprintf(stderr, "%s: %s\n", __func__, phip_last_error());
            goto done;
        }
    }
    for (int p = 0; p < npairs; ++p) {
        if (out->score) out->score[p] = res[p].score;
        if (out->end_query) out->end_query[p] = res[p].end_query;
        if (out->end_ref) out->end_ref[p] = res[p].end_ref;
        if (out->matches) out->matches[p] = 0;
        if (out->similar) out->similar[p] = 0;
        if (out->length) out->length[p] = 0;
        /* the flags parasail_nw_banded_hip sets (src/nw_banded.c:75-77 with
         * the hip family's width and lanes) */
        if (out->flag)
            out->flag[p] = PARASAIL_FLAG_NW | PARASAIL_FLAG_BANDED | PARASAIL_FLAG_BITS_32 |
                           PARASAIL_FLAG_LANES_64;
    }
    rc = 0;
done:
    if (stage) phip_stage_release(stage);
    else { free(q); free(t); }
    free(ql); free(tl); free(qo); free(to); free(res);
    return rc;
}

PHOST_EXPORT int parasail_hip_nw_banded_set_segment(int columns)
{
    return phip_nw_banded_segment_cols(columns);
}

PHOST_EXPORT void parasail_hip_nw_banded_last_stats(int *pairs, int *segments, int *redone)
{
    phip_nw_banded_last_stats(pairs, segments, redone);
}
