// dp32.h -- int32 wavefront kernels for _stats and _trace (dp32.hip) and
// the device CIGAR walk.
#pragma once

#include <vector>

#include "wave.h"

// Whether a pair of this length and kind goes to the wavefront kernel.
// check: 0, or the width whose saturation the caller tests (the kernel then
// returns the pair's exact extrema).
bool dp32_eligible(const phip_matrix_t &mx, int alg, int kind, int check, int open, int gap, int m,
                   int n);
// Whether every value of a pair provably stays inside the `check`-bit
// windows of the reference's kernels, so no extrema are needed.
bool dp32_window_holds(const phip_matrix_t &mx, int alg, int kind, int check, int open, int gap,
                       int m, int n);
// Whether one pair on its own (a per-call alignment) goes to the kernel with
// 64-lane jobs; score-only pairs go whatever the check (extrema computed).
bool dp32_eligible_one(const phip_matrix_t &mx, int alg, int kind, int check, int open, int gap,
                       int m, int n);
// Trace codes in the wavefront kernel's layout are 4-bit fields: bits 0-1
// the source of H (0 none -- SW at 0 --, 1 INS, 2 DEL, 3 DIAG), bit 2 set
// when E opens (DIAG_E, else INS_E), bit 3 when F opens (DIAG_F, else
// DEL_F); field ((s*T + t)*rpad(R) + r) of a pair's region holds row
// t*R + r at step s (column s - t), two fields a byte, low field first.
__host__ __device__ inline int dp32_rpad(int R) { return (R + 7) & ~7; }
__host__ __device__ inline int dp32_code_of(int f)
{
    return (f & 8 ? 32 : 64) | (f & 4 ? 8 : 16) | ((f & 3) == 3 ? 4 : (f & 3));
}
// the reference's trace code (parasail.h:31-41) of cell (i, j) of a pair
// traced in layout lay (T | R << 8)
__host__ __device__ inline int dp32_trace_code(const uint8_t *tr, int lay, int64_t i, int64_t j)
{
    const int T = lay & 0xff, R = lay >> 8;
    const int64_t t = i / R, r = i - t * R;
    const int64_t f = ((j + t) * T + t) * dp32_rpad(R) + r;
    return dp32_code_of((tr[f >> 1] >> ((f & 1) * 4)) & 15);
}

// Trace layout code of the wavefront kernel for a query of m rows and jobs
// of at least min_lanes lanes (T | R << 8; 0 if the pair is outside its
// range) and the bytes its trace takes for n columns.
int32_t dp32_layout(int m, int min_lanes = 16);
int64_t dp32_trace_bytes(int m, int n, int min_lanes = 16);

// Jobs and waves of one pass, as a host image of the device region the
// caller uploads (bytes long).
struct Dp32Plan {
    std::vector<char> image;
    size_t bytes = 0, o_jobs = 0, o_waves = 0;
    int njobs = 0, nwaves = 0;
};
// status[p] and troff[p] (may be NULL) are indexed by pair.
// fine: rows per lane from a finer set (no traceback)
bool dp32_plan(const HostBatch &hb, const std::vector<int> &pairs, const int32_t *status,
               const int64_t *troff, Dp32Plan &P, int min_lanes = 16, bool fine = false);
int dp32_launch(const DPParams &dp, const Dp32Plan &P, const char *region, hipStream_t st);

// CIGAR of every pair from its trace (row-major when tlay[p] is 0, else
// the wavefront order of layout tlay[p]), ends from out[]; ops[opoff[p]..]
// receives nops[p] ops (-1: none), beg[2p], beg[2p+1] the start.
struct CigarArgs {
    const uint8_t *qraw, *traw;
    const int64_t *qoff, *toff;
    const int32_t *qlen, *tlen;
    const int8_t *trace;
    const int64_t *troff;
    const int32_t *tlay;
    const phip_out_t *out;
    int is_sg, npairs;
    uint32_t *ops;
    const int64_t *opoff;
    int32_t *nops, *beg;
};
int cigar_launch(const CigarArgs &a, hipStream_t st);
// pair p's ops moved from opoff[p] to coff[p] of dst (after cigar_launch)
int cigar_compact_launch(const uint32_t *ops, const int64_t *opoff, const int32_t *nops,
                         const int64_t *coff, int np, uint32_t *dst, hipStream_t st);
