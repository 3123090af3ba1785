// noalloc.hip -- parasail_hip_align_batch_noalloc: score or _stats of many
// pairs held in device memory, with every byte of device memory beyond the
// caller's inputs and outputs taken from one static array.
//
// It is the device-memory form of parasail_hip_align_batch at width 32 (the
// exact int32 results that call reports without saturation) over the
// general int32 tier, k_simple, whose per-pair row scratch the other entry
// points take from the allocator.  Here it comes from `scratch`, a
// file-scope __device__ array of kScratchBytes per GPU, and the call makes
// no allocation of any kind, so a caller that cannot allocate -- code under
// HIP graph capture, a serving loop that must not call hipMalloc -- can use
// it.  The array's device address is fetched once per device with
// hipGetSymbolAddress; host code never uses the array's name as an address.
//
// Layout of the array for one call:
//   [0, 64 KiB)            substitution matrix values (square, size <= 128)
//   [64 KiB, 65 KiB)       the matrix's mapper (256 ints)
//   [68 KiB, +chunk*44 B)  per-pair results of the current chunk
//   [.., end)              k_simple's row planes: lanes x (max_len + 1) ints
//                          per plane, 2 planes (score) or 8 (_stats)
// Every part is written before it is read within a call, so nothing needs
// clearing; the array keeps the last call's contents.  Calls that overlap
// on one GPU share the array and race: a caller serialises them.
#include <pthread.h>
#include <string.h>

#include <algorithm>

#include "../../../include/parasail_hip.h"
#include "common.h"

namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr size_t kMatOff = 0, kMapOff = (size_t)64 << 10, kOutOff = (size_t)68 << 10;
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) char scratch[kScratchBytes];

namespace {

pthread_mutex_t g_sym_lock = PTHREAD_MUTEX_INITIALIZER;
char *g_sym_ptr[kMaxDevices];

// the array's address on the current device (fetched once per device)
char *scratch_ptr(char *err, size_t errn)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        snprintf(err, errn, "no current HIP device");
        return nullptr;
    }
    pthread_mutex_lock(&g_sym_lock);
    char *p = g_sym_ptr[dev];
    if (!p) {
        void *q = nullptr;
        const hipError_t rc = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
        if (rc != hipSuccess || !q)
            snprintf(err, errn, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(rc));
        else
            g_sym_ptr[dev] = p = (char *)q;
    }
    pthread_mutex_unlock(&g_sym_lock);
    return p;
}

// chunk results -> the caller's arrays (any may be NULL), with the flags
// parasail_hip_align_batch sets at width 32
__global__ void __launch_bounds__(256) k_noalloc_scatter(const phip_out_t *res, int first, int count,
                                                          int flag, int stats, int *score, int *eq,
                                                          int *er, int *matches, int *similar,
                                                          int *length, int *flags)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const phip_out_t o = res[k];
    const int p = first + k;
    const bool bad = o.status & PHIP_ST_ERR;
    if (score) score[p] = o.score;
    if (eq) eq[p] = o.end_query;
    if (er) er[p] = o.end_ref;
    if (matches) matches[p] = stats ? o.matches : 0;
    if (similar) similar[p] = stats ? o.similar : 0;
    if (length) length[p] = stats ? o.length : 0;
    if (flags) flags[p] = bad ? PARASAIL_FLAG_INVALID : flag;
}

int parse_alg(const char *name, int *alg, int *sgf)
{
    static const struct { const char *n; int alg, flags; } tab[] = {
        {"sw", PHIP_SW, 0}, {"nw", PHIP_NW, 0},
        {"sg", PHIP_SG, PHIP_S1_BEG | PHIP_S1_END | PHIP_S2_BEG | PHIP_S2_END},
        {"sg_qb", PHIP_SG, PHIP_S1_BEG}, {"sg_qe", PHIP_SG, PHIP_S1_END},
        {"sg_qx", PHIP_SG, PHIP_S1_BEG | PHIP_S1_END}, {"sg_db", PHIP_SG, PHIP_S2_BEG},
        {"sg_de", PHIP_SG, PHIP_S2_END}, {"sg_dx", PHIP_SG, PHIP_S2_BEG | PHIP_S2_END},
        {"sg_qb_de", PHIP_SG, PHIP_S1_BEG | PHIP_S2_END},
        {"sg_qe_db", PHIP_SG, PHIP_S1_END | PHIP_S2_BEG},
        {"sg_qb_db", PHIP_SG, PHIP_S1_BEG | PHIP_S2_BEG},
        {"sg_qe_de", PHIP_SG, PHIP_S1_END | PHIP_S2_END},
    };
    for (const auto &e : tab)
        if (strcmp(name, e.n) == 0) {
            *alg = e.alg;
            *sgf = e.flags;
            return 0;
        }
    return -1;
}

}  // namespace

extern "C" __attribute__((visibility("default"))) int parasail_hip_align_batch_noalloc(
    const char *alg_name, int kind, const char *s1, const int64_t *s1_off, const int *s1_len,
    const char *s2, const int64_t *s2_off, const int *s2_len, int npairs, int max_len, int open,
    int gap, const parasail_matrix_t *matrix, parasail_hip_batch_result_t *out, void *stream)
{
    char err[256];
    err[0] = 0;
    int alg = 0, sgf = 0;
    const int narr = kind == PHIP_STATS ? 8 : 2;
    const size_t per_lane = ((size_t)max_len + 1) * narr * sizeof(int);
    const size_t planes_off = kOutOff + ((size_t)1 << 20);  // room for 23,831 results
    size_t lanes = 0;
    if (!alg_name || parse_alg(alg_name, &alg, &sgf) != 0) {
        snprintf(err, sizeof(err), "unknown algorithm %s", alg_name ? alg_name : "(null)");
    } else if ((kind != PHIP_SCORE && kind != PHIP_STATS) || !s1 || !s1_off || !s1_len || !s2 ||
               !s2_off || !s2_len || npairs < 0 || open < 0 || gap < 0 || !matrix || !out ||
               out->cigars || max_len <= 0) {
        snprintf(err, sizeof(err), "invalid arguments");
    } else if (matrix->type != PARASAIL_MATRIX_TYPE_SQUARE || matrix->size <= 0 ||
               matrix->size > 128) {
        snprintf(err, sizeof(err), "a square matrix of at most 128 letters is required");
    } else {
        lanes = (kScratchBytes - planes_off) / per_lane;
        lanes = std::min<size_t>(lanes, ((size_t)1 << 20) / sizeof(phip_out_t));
        if (lanes < 1)
            snprintf(err, sizeof(err),
                     "max_len %d exceeds the %zu-byte scratch array (at most %zu for %s)", max_len,
                     kScratchBytes, (kScratchBytes - planes_off) / (narr * sizeof(int)) - 1,
                     kind == PHIP_STATS ? "_stats" : "scores");
    }
    char *base = err[0] ? nullptr : scratch_ptr(err, sizeof(err));
    if (!base) {
        fprintf(stderr, "%s: %s\n", __func__, err);
        return -1;
    }
    if (npairs == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    const size_t msz = (size_t)matrix->size * matrix->size * sizeof(int);
    if (hipMemcpyAsync(base + kMatOff, matrix->matrix, msz, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(base + kMapOff, matrix->mapper, 256 * sizeof(int), hipMemcpyHostToDevice,
                       st) != hipSuccess) {
        fprintf(stderr, "%s: matrix upload failed\n", __func__);
        return -1;
    }
    DPParams dp;
    memset(&dp, 0, sizeof(dp));
    dp.alg = alg;
    dp.kind = kind;
    dp.sg_flags = alg == PHIP_SG ? sgf : 0;
    dp.open = open;
    dp.gap = gap;
    dp.matrix = (const int *)(base + kMatOff);
    dp.msize = matrix->size;
    dp.mapper = (const int *)(base + kMapOff);
    dp.q = (const uint8_t *)s1;
    dp.qoff = s1_off;
    dp.qlen = (const int32_t *)s1_len;
    dp.t = (const uint8_t *)s2;
    dp.toff = s2_off;
    dp.tlen = (const int32_t *)s2_len;
    dp.npairs = npairs;
    int flag = alg == PHIP_SW ? PARASAIL_FLAG_SW : alg == PHIP_NW ? PARASAIL_FLAG_NW : PARASAIL_FLAG_SG;
    if (alg == PHIP_SG) {
        if (sgf & PHIP_S1_BEG) flag |= PARASAIL_FLAG_SG_S1_BEG;
        if (sgf & PHIP_S1_END) flag |= PARASAIL_FLAG_SG_S1_END;
        if (sgf & PHIP_S2_BEG) flag |= PARASAIL_FLAG_SG_S2_BEG;
        if (sgf & PHIP_S2_END) flag |= PARASAIL_FLAG_SG_S2_END;
    }
    flag |= PARASAIL_FLAG_STRIPED | PARASAIL_FLAG_LANES_64 | PARASAIL_FLAG_BITS_32;
    if (kind == PHIP_STATS) flag |= PARASAIL_FLAG_STATS;
    phip_out_t *res = (phip_out_t *)(base + kOutOff);
    int *rows = (int *)(base + planes_off);
    for (int first = 0; first < npairs; first += (int)lanes) {
        const int cnt = (int)std::min<size_t>(lanes, (size_t)(npairs - first));
        // k_simple writes pair first + k's result at out[first + k]
        dp.out = (phip_out_t *)((uintptr_t)res - (uintptr_t)first * sizeof(phip_out_t));
        hipLaunchKernelGGL(k_simple, dim3((cnt + 255) / 256), dim3(256), 0, st, dp, first, cnt,
                           (const int32_t *)nullptr, (const int *)nullptr, rows, cnt, max_len);
        hipLaunchKernelGGL(k_noalloc_scatter, dim3((cnt + 255) / 256), dim3(256), 0, st, res, first,
                           cnt, flag, kind == PHIP_STATS, out->score, out->end_query, out->end_ref,
                           out->matches, out->similar, out->length, out->flag);
        if (hipGetLastError() != hipSuccess) {
            fprintf(stderr, "%s: launch failed\n", __func__);
            return -1;
        }
    }
    return 0;
}
