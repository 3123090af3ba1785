// common.h -- shared device-side definitions for the gfx950 DP kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../phip.h"

#define PHIP_NEG_INF32 (-(1 << 29))  // far below any reachable DP value

// Parameters shared by every kernel of one phip_align call.
struct DPParams {
    int alg, kind, sg_flags;
    int open, gap;
    int check;             // track exact extrema for the saturation predicate
    const int *matrix;     // device, row-major length*size
    int msize, is_pssm;
    const uint8_t *q;      // device residue codes
    const int64_t *qoff;   // NULL: one query (offset 0) for every pair
    const int32_t *qlen;   // NULL: every pair's query has qlen1 rows
    int qlen1;
    const uint8_t *t;
    const int64_t *toff;
    const int32_t *tlen;
    int npairs;
    phip_out_t *out;       // device
    int8_t *trace;         // device, row-major per pair (kind == TRACE)
    const int64_t *trace_off;
    const int *mapper;     // k_simple: device, 256 entries; q/t then hold the
                           // residues as given, mapped cell by cell (NULL:
                           // q/t hold codes)
    int tab;               // one-pair launches: 1 DP tables, 2 last row and
    int *tabs[4];          // column, of score (and matches, similar, length)
};

__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }

// Workgroup barrier that orders LDS accesses only: loads from HBM still in
// flight stay in flight (__syncthreads also waits for every one of them).
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

#define HIP_OK(expr)                                                        \
    do {                                                                    \
        hipError_t _e = (expr);                                             \
        if (_e != hipSuccess) {                                             \
            phip_set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr,       \
                           hipGetErrorString(_e));                          \
            return -1;                                                      \
        }                                                                   \
    } while (0)

extern "C" void phip_set_error(const char *fmt, ...);

extern "C" __global__ void k_simple(DPParams p, int first, int count, const int32_t *list,
                                    const int *dcount, int *scratch, int stride, int maxn);
