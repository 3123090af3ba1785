/*
 * serve.c -- HBM-resident target database, database search and the serving
 * loop, as C entry points over the device code in search.hip.
 *
 * The reference has no database or serving API: its callers search a
 * database by running parasail_sw_striped_profile_* over every target in an
 * OpenMP loop (tests/test_openmp.c:353-369, apps/parasail_aligner.cpp:
 * 1714-1731) and keep the best hits themselves.  Here the database is mapped
 * and uploaded once; a search is SW (int16 with an exact int32 re-run of any
 * target whose score leaves the int16 window) of one query -- residues with a
 * square matrix, or a PSSM (parasail_matrix_pssm_create) -- against every
 * target, returning the k best hits (score desc, then target index asc).
 */
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

struct parasail_hip_db {
    phip_db_t *db;
    int size;
    int mapper[256];
};

struct parasail_hip_server {
    phip_server_t *srv;
    parasail_hip_db_t *db;
};

PHOST_EXPORT parasail_hip_db_t *parasail_hip_db_create(const char *seqs, const int64_t *off,
                                                       const int *len, int n,
                                                       const parasail_matrix_t *matrix)
{
    if (!seqs || !off || !len || n <= 0 || !matrix) {
        fprintf(stderr, "%s: invalid arguments\n", __func__);
        return NULL;
    }
    if (matrix->type != PARASAIL_MATRIX_TYPE_SQUARE || matrix->size > 255) {
        fprintf(stderr, "%s: the database alphabet must come from a square matrix of < 256 symbols\n",
                __func__);
        return NULL;
    }
    if (phip_init() != 0) {
        errno = ENOSYS;
        return NULL;
    }
    int64_t tot = 0;
    for (int i = 0; i < n; ++i) {
        if (len[i] <= 0) {
            fprintf(stderr, "%s: target %d is empty\n", __func__, i);
            return NULL;
        }
        tot += len[i];
    }
    uint8_t *codes = (uint8_t *)malloc((size_t)tot);
    int64_t *doff = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    int32_t *dlen = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    parasail_hip_db_t *h = (parasail_hip_db_t *)calloc(1, sizeof(parasail_hip_db_t));
    if (!codes || !doff || !dlen || !h) {
        free(codes); free(doff); free(dlen); free(h);
        fprintf(stderr, "%s: out of host memory\n", __func__);
        return NULL;
    }
    int64_t p = 0;
    for (int i = 0; i < n; ++i) {
        const unsigned char *src = (const unsigned char *)seqs + off[i];
        for (int j = 0; j < len[i]; ++j) codes[p + j] = (uint8_t)matrix->mapper[src[j]];
        doff[i] = p;
        dlen[i] = len[i];
        p += len[i];
    }
    h->size = matrix->size;
    memcpy(h->mapper, matrix->mapper, sizeof(h->mapper));
    h->db = phip_db_create(codes, doff, dlen, n, matrix->size);
    free(codes); free(doff); free(dlen);
    if (!h->db) {
        fprintf(stderr, "%s: %s\n", __func__, phip_last_error());
        free(h);
        return NULL;
    }
    return h;
}

PHOST_EXPORT void parasail_hip_db_free(parasail_hip_db_t *db)
{
    if (!db) return;
    phip_db_free(db->db);
    free(db);
}

/* the query of one request in device terms; *codes is malloc'd (or NULL) */
static int make_query(const parasail_hip_db_t *db, const char *s1, int s1Len,
                      const parasail_matrix_t *matrix, int open, int gap, phip_query_t *q,
                      uint8_t **codes, const char *fn)
{
    *codes = NULL;
    if (!matrix || open < 0 || gap < 0) {
        fprintf(stderr, "%s: invalid arguments\n", fn);
        return -1;
    }
    if (matrix->size != db->size || memcmp(matrix->mapper, db->mapper, sizeof(db->mapper)) != 0) {
        fprintf(stderr, "%s: the matrix alphabet differs from the database's\n", fn);
        return -1;
    }
    const int pssm = matrix->type == PARASAIL_MATRIX_TYPE_PSSM;
    const int m = pssm ? matrix->length : s1Len;
    if (m <= 0 || (!pssm && !s1)) {
        fprintf(stderr, "%s: missing query\n", fn);
        return -1;
    }
    memset(q, 0, sizeof(*q));
    if (!pssm) {
        *codes = (uint8_t *)malloc((size_t)m);
        if (!*codes) return -1;
        for (int i = 0; i < m; ++i) (*codes)[i] = (uint8_t)matrix->mapper[(unsigned char)s1[i]];
    }
    q->codes = *codes;
    q->m = m;
    q->mx.values = matrix->matrix;
    q->mx.size = matrix->size;
    q->mx.length = pssm ? matrix->length : matrix->size;
    q->mx.is_pssm = pssm;
    q->mx.max = matrix->max;
    q->mx.min = matrix->min;
    q->open = open;
    q->gap = gap;
    return 0;
}

PHOST_EXPORT int parasail_hip_db_search(parasail_hip_db_t *db, const char *s1, int s1Len,
                                        const parasail_matrix_t *matrix, int open, int gap, int k,
                                        parasail_hip_hit_t *hits)
{
    if (!db || !hits || k <= 0) {
        fprintf(stderr, "%s: invalid arguments\n", __func__);
        return -1;
    }
    phip_query_t q;
    uint8_t *codes;
    if (make_query(db, s1, s1Len, matrix, open, gap, &q, &codes, __func__) != 0) return -1;
    int nh = 0;
    const int rc = phip_db_search(db->db, &q, k, (phip_hit_t *)hits, &nh);
    free(codes);
    if (rc != 0) {
        fprintf(stderr, "%s: %s\n", __func__, phip_last_error());
        return -1;
    }
    return nh;
}

PHOST_EXPORT parasail_hip_server_t *parasail_hip_server_create(parasail_hip_db_t *db,
                                                               size_t mem_budget)
{
    if (!db) {
        fprintf(stderr, "%s: missing database\n", __func__);
        return NULL;
    }
    parasail_hip_server_t *s = (parasail_hip_server_t *)calloc(1, sizeof(*s));
    if (!s) return NULL;
    s->db = db;
    s->srv = phip_server_create(db->db, mem_budget);
    if (!s->srv) {
        fprintf(stderr, "%s: %s\n", __func__, phip_last_error());
        free(s);
        return NULL;
    }
    return s;
}

PHOST_EXPORT int parasail_hip_server_submit(parasail_hip_server_t *s, int64_t id, const char *s1,
                                            int s1Len, const parasail_matrix_t *matrix, int open,
                                            int gap, int k, double arrival)
{
    if (!s || k <= 0) {
        fprintf(stderr, "%s: invalid arguments\n", __func__);
        return -1;
    }
    phip_query_t q;
    uint8_t *codes;
    if (make_query(s->db, s1, s1Len, matrix, open, gap, &q, &codes, __func__) != 0) return -1;
    const int rc = phip_server_submit(s->srv, id, &q, k, arrival);
    free(codes);
    if (rc != 0) fprintf(stderr, "%s: %s\n", __func__, phip_last_error());
    return rc;
}

PHOST_EXPORT int parasail_hip_server_poll(parasail_hip_server_t *s, int64_t *id,
                                          parasail_hip_hit_t *hits, int *nhits,
                                          double *latency_ms, int wait)
{
    if (!s) return -1;
    return phip_server_poll(s->srv, id, (phip_hit_t *)hits, nhits, latency_ms, wait);
}

PHOST_EXPORT void parasail_hip_server_free(parasail_hip_server_t *s)
{
    if (!s) return;
    phip_server_free(s->srv);
    free(s);
}

PHOST_EXPORT int parasail_hip_server_stats(parasail_hip_server_t *s, int64_t *completed,
                                           int64_t *deferred, int *max_active)
{
    if (!s) return -1;
    return phip_server_stats(s->srv, completed, deferred, max_active);
}

PHOST_EXPORT double parasail_hip_now(void) { return phip_now(); }
