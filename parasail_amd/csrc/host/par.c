/* par.c -- host threads for the batch layer's byte work: turning residues
 * into matrix codes (and upper-cased copies for the CIGAR walk), and the
 * per-pair results into the caller's arrays, cost more than the kernels of
 * a large batch (or a noticeable share of them) when one thread does it. */
#include <ctype.h>
#include <pthread.h>
#include <unistd.h>

#include "internal.h"

#define PHOST_MAX_THREADS 16

typedef struct {
    const phost_enc_t *e;
    int lo, hi;
} enc_part_t;

static void encode_range(const phost_enc_t *e, int lo, int hi)
{
    for (int k = lo; k < hi; ++k) {
        const unsigned char *src = (const unsigned char *)e->src + e->src_off[k];
        const int n = e->len[k];
        if (e->codes) {
            uint8_t *dst = e->codes + e->dst_off[k];
            for (int i = 0; i < n; ++i) dst[i] = (uint8_t)e->mapper[src[i]];
        }
        if (e->upper) {
            uint8_t *up = e->upper + e->dst_off[k];
            for (int i = 0; i < n; ++i) up[i] = (uint8_t)toupper(src[i]);
        }
    }
}

static void *encode_part(void *arg)
{
    const enc_part_t *p = (const enc_part_t *)arg;
    encode_range(p->e, p->lo, p->hi);
    return NULL;
}

void phost_encode(const phost_enc_t *e, int npairs)
{
    int64_t bytes = 0;
    for (int k = 0; k < npairs; ++k) bytes += e->len[k];
    long nt = sysconf(_SC_NPROCESSORS_ONLN);
    if (nt > PHOST_MAX_THREADS) nt = PHOST_MAX_THREADS;
    /* threads only where each gets at least ~1 MB */
    if (nt > bytes / (1 << 20)) nt = (long)(bytes / (1 << 20));
    if (nt > npairs) nt = npairs;
    if (nt <= 1) {
        encode_range(e, 0, npairs);
        return;
    }
    /* split at equal byte counts */
    pthread_t th[PHOST_MAX_THREADS];
    enc_part_t part[PHOST_MAX_THREADS];
    int started[PHOST_MAX_THREADS] = {0};
    int k = 0;
    int64_t acc = 0;
    for (long i = 0; i < nt; ++i) {
        const int64_t goal = bytes * (i + 1) / nt;
        part[i].e = e;
        part[i].lo = k;
        while (k < npairs && (acc < goal || i == nt - 1)) acc += e->len[k++];
        part[i].hi = k;
    }
    for (long i = 1; i < nt; ++i)
        started[i] = pthread_create(&th[i], NULL, encode_part, &part[i]) == 0;
    encode_range(e, part[0].lo, part[0].hi);
    for (long i = 1; i < nt; ++i) {
        if (started[i])
            pthread_join(th[i], NULL);
        else
            encode_range(e, part[i].lo, part[i].hi);
    }
}

typedef struct {
    void (*fn)(void *ctx, int lo, int hi);
    void *ctx;
    int lo, hi;
} for_part_t;

static void *for_part(void *arg)
{
    const for_part_t *p = (const for_part_t *)arg;
    p->fn(p->ctx, p->lo, p->hi);
    return NULL;
}

void phost_parallel_for(int n, int min_per_thread, void (*fn)(void *ctx, int lo, int hi),
                        void *ctx)
{
    long nt = sysconf(_SC_NPROCESSORS_ONLN);
    if (nt > PHOST_MAX_THREADS) nt = PHOST_MAX_THREADS;
    if (min_per_thread < 1) min_per_thread = 1;
    if (nt > n / min_per_thread) nt = n / min_per_thread;
    if (nt <= 1) {
        if (n > 0) fn(ctx, 0, n);
        return;
    }
    pthread_t th[PHOST_MAX_THREADS];
    for_part_t part[PHOST_MAX_THREADS];
    int started[PHOST_MAX_THREADS] = {0};
    for (long i = 0; i < nt; ++i) {
        part[i].fn = fn;
        part[i].ctx = ctx;
        part[i].lo = (int)((int64_t)n * i / nt);
        part[i].hi = (int)((int64_t)n * (i + 1) / nt);
    }
    for (long i = 1; i < nt; ++i)
        started[i] = pthread_create(&th[i], NULL, for_part, &part[i]) == 0;
    fn(ctx, part[0].lo, part[0].hi);
    for (long i = 1; i < nt; ++i) {
        if (started[i])
            pthread_join(th[i], NULL);
        else
            fn(ctx, part[i].lo, part[i].hi);
    }
}
