/*
 * phip.h -- the thin C ABI between the C host layer (parasail_amd/csrc/host)
 * and the HIP device layer (parasail_amd/csrc/device).  Plain C types only.
 *
 * One call aligns a batch of pairs that share an algorithm, result kind,
 * gap penalties and a substitution matrix (or one PSSM).  The host layer
 * builds every public parasail_* entry point on top of it.
 */
#ifndef PHIP_H
#define PHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { PHIP_SW = 0, PHIP_NW = 1, PHIP_SG = 2 };
enum { PHIP_SCORE = 0, PHIP_STATS = 1, PHIP_TRACE = 2 };

/* sg flag bits (parasail_sg_flags argument order, src/sg.c:31-35) */
#define PHIP_S1_BEG 1
#define PHIP_S1_END 2
#define PHIP_S2_BEG 4
#define PHIP_S2_END 8

/* per-pair status bits returned in phip_out_t.status */
#define PHIP_ST_SAT8 1   /* exact values leave the 8-bit window  */
#define PHIP_ST_SAT16 2  /* exact values leave the 16-bit window */
#define PHIP_ST_SAT32 4  /* exact values leave the 32-bit window */
#define PHIP_ST_W32 8    /* computed by the int32 tier            */
#define PHIP_ST_ERR 16   /* could not be computed                 */

typedef struct phip_matrix {
    const int *values;  /* row-major, `length` rows of `size` ints     */
    int size;           /* columns (alphabet + 1)                      */
    int length;         /* rows: size for square, query length for PSSM */
    int is_pssm;
    int max, min;
} phip_matrix_t;

typedef struct phip_batch {
    int alg, kind, sg_flags;
    int open, gap;
    int want_sat_check;   /* 0, or the width (8, 16) whose saturation the
                           * caller will test: exact extrema are returned */
    int want_min_h;       /* with want_sat_check: min_h is tested too (NW
                           * scan score), which the packed kernel omits */
    phip_matrix_t matrix;
    int npairs;
    /* mapped residue codes (matrix->mapper), host memory.  For a PSSM the
     * query codes are still used by _stats (matches) and may be NULL
     * otherwise; qlen must then equal matrix.length. */
    const uint8_t *q;
    const int64_t *qoff;
    const int32_t *qlen;
    const uint8_t *t;
    const int64_t *toff;
    const int32_t *tlen;
    /* NULL, or the device copy of pair 0's query codes and of the matrix
     * (phip_profile_upload) that phip_align_one uses instead of uploading
     * q and matrix.values again */
    const struct phip_dprofile *dprof;
} phip_batch_t;

typedef struct phip_out {
    int32_t score, end_query, end_ref;
    int32_t matches, similar, length;
    int32_t status;
    /* exact extrema for the saturation windows (want_sat_check): the
     * smallest H/E/F over the cells and the DP's boundary row and column,
     * the largest H, the smallest H of a cell, and (_stats) the largest
     * length count of a cell's H (a cell's length is never below its
     * matches or similar counts) */
    int32_t min_val, max_val;
    int32_t min_h, max_len;
} phip_out_t;

/* Initialise the device context once (thread-safe).  0 on success; on a
 * host without a usable HIP device returns -1 and sets errno = ENOSYS. */
int phip_init(void);
const char *phip_last_error(void);

/* Align every pair of the batch.  out[npairs]; for kind == PHIP_TRACE,
 * traces[p] (if non-NULL) receives the row-major int8 trace codes of pair p
 * (qlen*tlen bytes, the LOC_NOVEC layout of cigar.c:156).  0 on success. */
int phip_align(const phip_batch_t *b, phip_out_t *out, int8_t **traces);

/* One pair (npairs == 1) with the lowest latency: one upload, one launch,
 * one download when the pair fits the 64-lane kernel, else phip_align. */
int phip_align_one(const phip_batch_t *b, phip_out_t *out, int8_t **traces);

/* A profile's query kept in device memory: its m residue codes and the
 * matrix values (length x size ints, square or PSSM) uploaded once, for
 * every per-call alignment made with the profile.  NULL on failure (the
 * error in phip_last_error).  Released by phip_profile_release. */
typedef struct phip_dprofile phip_dprofile_t;
phip_dprofile_t *phip_profile_upload(const uint8_t *q, int m, const phip_matrix_t *mx);
void phip_profile_release(phip_dprofile_t *p);

/* One pair (npairs == 1, kind score or stats) with its DP tables: tab 1
 * writes every cell's H (and, for stats, the matches, similar and length of
 * that H) row-major, qlen*tlen ints into dst[0] (dst[1..3]); tab 2 writes
 * the last row (tlen ints) followed by the last column (qlen ints).  The
 * PARASAIL_TABLE / PARASAIL_ROWCOL outputs of src/sw.c:159-170.  0 on
 * success. */
int phip_align_tables(const phip_batch_t *b, int tab, phip_out_t *out, int **dst);

/* Traceback of every pair and its CIGAR, both on the device (kind must be
 * PHIP_TRACE).  qraw/traw: the residues as given, upper-cased, at the
 * batch's offsets (the walk compares residues, cigar_template.c:104-118).
 * The caller sizes ops for sum(qlen + tlen + 2) entries, off/nops for
 * npairs and beg for 2*npairs; pair p's nops[p] ops (-1: no CIGAR) start at
 * ops + off[p], its start (beg_query, beg_ref) at beg[2p]. */
typedef struct phip_cigars {
    uint32_t *ops;
    int64_t *off;
    int32_t *nops;
    int32_t *beg;
} phip_cigars_t;
int phip_align_cigars(const phip_batch_t *b, const uint8_t *qraw, const uint8_t *traw,
                      phip_out_t *out, phip_cigars_t *cig);

/* Device-resident batch: residues, lengths and results stay in HBM across
 * calls, and the wavefront plan for each scoring is built once.  Score-only
 * alignments; NULL / -1 on failure. */
/* The process's pinned staging buffer of at least `bytes` (NULL when it
 * cannot be had: use ordinary memory).  The caller holds it until
 * phip_stage_release; other callers wait meanwhile. */
void *phip_stage_acquire(size_t bytes);
void phip_stage_release(void *p);
void *phip_resident_create(const phip_batch_t *b);
/* Score-only results of a resident batch, 16 bytes a pair. */
typedef struct phip_out4 {
    int32_t score, end_query, end_ref, status;
} phip_out4_t;
/* Score-only pass with the results packed on the device and downloaded into
 * the handle's pinned buffer (*res, valid until the next call or free). */
int phip_resident_align4(void *handle, int alg, int sg_flags, int open, int gap,
                         const phip_out4_t **res);
/* Several score passes over a resident batch in one call (at most 4): launch
 * them all (this takes the device lock), collect pass k's packed results with
 * phip_resident_wait4 (its kernel time, and its main kernel's, in ms), then
 * release the lock with phip_resident_end4, also after a failed wait. */
int phip_resident_launch4(void *handle, int npass, const int *alg, const int *sg_flags, int open,
                          int gap);
int phip_resident_wait4(void *handle, int k, const phip_out4_t **res, double *kernel_ms,
                        double *main_ms);
void phip_resident_end4(void *handle);
/* whether a multi-pass call repeating the previous one is captured into a
 * HIP graph and replayed (default on); returns the previous setting */
int phip_resident_set_graph(void *handle, int on);
/* graphs captured, and multi-pass calls that replayed one */
int phip_resident_graph_stats(void *handle, int *captures, int *replays);
/* out == NULL: into the handle's pinned buffer, phip_resident_results() */
int phip_resident_align(void *handle, int alg, int sg_flags, int open, int gap, phip_out_t *out);
const phip_out_t *phip_resident_results(void *handle);
void phip_resident_free(void *handle);

/* banded global alignment (src/nw_banded.c) for one pair */
int phip_nw_banded(const uint8_t *s1, int m, const uint8_t *s2, int n,
                   int open, int gap, int k, const phip_matrix_t *mx,
                   phip_out_t *out);

/* banded global alignment of every pair of b (alg/kind ignored), band k */
int phip_nw_banded_batch(const phip_batch_t *b, int k, phip_out_t *out);
/* columns a segment of a long banded pair owns (0: no segments; < 0: keep);
 * returns the previous value */
int phip_nw_banded_segment_cols(int cols);
/* the last banded batch: segmented pairs, segments, pairs aligned again
 * without segments */
void phip_nw_banded_last_stats(int *pairs, int *segments, int *redone);

int phip_device_info(int *cus, int *clock_khz);

/* measured step costs of the wavefront planner (see wave.hip) */
int phip_wave_autotune(const phip_matrix_t *mx, int open, int gap);
int phip_wave_autotune_table(int *rows, double *sw_cost, double *nw_cost, int cap);
void phip_wave_autotune_reset(void);

/* ---- database search and serving (search.hip) ---- */

typedef struct phip_hit {
    int score, end_query, end_ref, target;
} phip_hit_t;

/* one query: residue codes (NULL for a PSSM) and its scoring */
typedef struct phip_query {
    const uint8_t *codes;
    int m;
    phip_matrix_t mx;   /* square (codes index rows) or PSSM (m rows) */
    int open, gap;
} phip_query_t;

typedef struct phip_db phip_db_t;
typedef struct phip_server phip_server_t;

/* targets: mapped codes with per-target offsets/lengths; size = alphabet
 * columns of the matrices that will search it (< 256) */
phip_db_t *phip_db_create(const uint8_t *codes, const int64_t *off, const int32_t *len, int n,
                          int size);
void phip_db_free(phip_db_t *db);
int64_t phip_db_residues(const phip_db_t *db);
/* SW of the query against every target; the k best (score desc, target
 * index asc) into hits; *nhits = min(k, targets).  0 on success. */
int phip_db_search(phip_db_t *db, const phip_query_t *q, int k, phip_hit_t *hits, int *nhits);

/* serving loop: up to 4 requests in flight (one stream each), admitted as
 * slots and device memory (mem_budget bytes, 0 = 32 GiB) free up */
phip_server_t *phip_server_create(phip_db_t *db, size_t mem_budget);
/* queues a request (inputs are copied); arrival = phip_now() time of its
 * arrival (0 = now), from which its latency is measured */
int phip_server_submit(phip_server_t *s, int64_t id, const phip_query_t *q, int k, double arrival);
/* next finished request: 1 with its hits and latency, 0 if none (wait = 0),
 * -1 if it failed */
int phip_server_poll(phip_server_t *s, int64_t *id, phip_hit_t *hits, int *nhits,
                     double *latency_ms, int wait);
void phip_server_free(phip_server_t *s);
/* requests completed, requests that waited for memory at admission, and
 * the most requests on the device at once */
int phip_server_stats(phip_server_t *s, int64_t *completed, int64_t *deferred, int *max_active);
double phip_now(void);

/* timing of the last phip_align call's dominant kernel (ms, HIP events) */
double phip_last_kernel_ms(void);
/* device time of the last call's dominant kernel (the wavefront kernel when
 * it ran, else the same as phip_last_kernel_ms) */
double phip_last_main_kernel_ms(void);

#ifdef __cplusplus
}
#endif
#endif
