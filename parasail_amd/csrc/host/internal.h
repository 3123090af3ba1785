/* internal.h -- shared declarations of the C host layer. */
#ifndef PHOST_INTERNAL_H
#define PHOST_INTERNAL_H

#include <stddef.h>
#include <stdint.h>

#include "../../../include/parasail_hip.h"
#include "../phip.h"

#define PHOST_EXPORT __attribute__((visibility("default")))

parasail_result_t *phost_result_new(void);
parasail_result_t *phost_result_new_stats(void);
parasail_result_t *phost_result_new_trace(int a, int b, size_t alignment, size_t size);
/* tables (rowcol 0) or last row and column (rowcol 1) of an a x b DP, of
 * scores alone or (stats) with the three counts; flags TABLE/ROWCOL and
 * STATS set */
parasail_result_t *phost_result_new_tables(int a, int b, int stats, int rowcol);
parasail_profile_t *phost_profile_new(const char *s1, int s1Len, const parasail_matrix_t *matrix);
void *phost_memalign(size_t alignment, size_t size);
int phost_flag_lanes(int flag);
int phost_flag_bits(int flag);
int phost_trace_at(const void *t, int bits, long loc);

/* one alignment through the device (align.c) */
parasail_result_t *phost_align(const char *fname, int alg, int sg_flags, int kind, int width,
                               int scan, const char *s1, int s1Len, const char *s2, int s2Len,
                               int open, int gap, const parasail_matrix_t *matrix,
                               const parasail_profile_t *profile);

/* saturation predicate of one width on exact DP values (align.c) */
/* Codes (through mapper) and/or upper-cased copies of npairs sequences:
 * pair k's len[k] residues at src + src_off[k] go to codes/upper +
 * dst_off[k] (either destination may be NULL); host threads share the
 * pairs when the batch is large. */
typedef struct {
    const char *src;
    const int64_t *src_off;
    const int32_t *len;
    const int64_t *dst_off;
    const int *mapper;
    uint8_t *codes, *upper;
} phost_enc_t;
void phost_encode(const phost_enc_t *e, int npairs);
/* fn(ctx, lo, hi) over [0, n) split across host threads, each taking at
 * least min_per_thread items (one thread, the caller's, below that) */
void phost_parallel_for(int n, int min_per_thread, void (*fn)(void *ctx, int lo, int hi),
                        void *ctx);

/* saturation window of the reference's w-bit kernel of (alg, kind,
 * strategy) on exact values, the values a saturated result reports, and
 * the width whose extrema a call must ask the device for (align.c) */
int phost_saturates(int alg, int kind, int scan, int w, int open,
                    const parasail_matrix_t *matrix, const phip_out_t *o);
int phost_saturated_value(int alg, int kind, int scan, int w);
int phost_check_width(int alg, int kind, int scan, int width);

/* CIGAR walk over a row-major int8 trace (cigar.c) */
parasail_cigar_t *phost_cigar_rowmajor(const char *seqA, int lena, const char *seqB, int lenb,
                                       int is_sg, int end_query, int end_ref, const int8_t *tr);

extern const parasail_matrix_t phost_blosum62;

#endif
