// nwlong.hip -- global alignment score (parasail_nw / nw_scan_32 semantics,
// src/nw.c:29-167) of long pairs in int32, one workgroup per pair.
//
// The pair's query rows are cut into bands of 64*R rows; wave w of the
// workgroup works on bands w, w+LW, w+2LW, ...  Within a band the 64 lanes
// form the wavefront chain of the int16 kernels (lane t holds R rows and
// computes column s - t at step s; the row above comes from lane t-1 by DPP),
// here in int32 because long global scores leave int16.  The LW bands of a
// group run as a pipeline: band w starts L = 2*LK steps after band w-1 and
// takes the bottom row of band w-1 from an LDS ring; all waves meet at a
// barrier every LK steps, and every ring value a wave reads was written at
// least one barrier earlier.  The last band of a group leaves its bottom row
// in HBM for the first band of the next group.
#include <string.h>

#include <algorithm>
#include <vector>

#include "search.h"

namespace {

constexpr int LW = 15;     // waves (bands in flight) per workgroup
constexpr int LK = 128;    // steps between barriers
constexpr int LL = 2 * LK; // lag between consecutive bands
constexpr int LRING = 512; // ring slots per band boundary (power of two)
constexpr int LNEG = -(1 << 29);
constexpr int LWNEG = -(1 << 24);  // substitution value of the pad rows

__device__ __forceinline__ int shr1(int v) { return __builtin_amdgcn_mov_dpp(v, 0x138, 0xF, 0xF, false); }

}  // namespace

struct LongCfg {
    const uint8_t *q;
    const int64_t *qoff;
    const int32_t *qlen;
    const uint8_t *t;
    const int64_t *toff;
    const int32_t *tlen;
    const int *mat;     // square matrix, size x size
    int msize;
    int open, gap;
    const int32_t *pairs;
    int npairs;
    int2 *gbuf;         // per workgroup: one row of gstride entries
    int64_t gstride;
    int2 *dump;         // per thread
    phip_out_t *out;
};

// LDS: table of W+open ((size+1) rows: codes, then a pad row), then the
// LW-1 rings of (H-open, F).
template <int R>
__global__ void __launch_bounds__(LW * 64) k_nwlong(LongCfg c)
{
    extern __shared__ int lds[];
    const int msize = c.msize, open = c.open, gap = c.gap;
    const int tabn = (msize + 1) * msize;
    for (int x = threadIdx.x; x < tabn; x += blockDim.x)
        lds[x] = x < msize * msize ? c.mat[x] + open : LWNEG;
    int2 *const ring = (int2 *)(lds + ((tabn + 1) & ~1));
    __syncthreads();
    if (blockIdx.x >= (unsigned)c.npairs) return;  // whole block
    const int w = threadIdx.x >> 6, t = threadIdx.x & 63;
    const int pr = c.pairs[blockIdx.x];
    const int m = c.qlen[pr], n = c.tlen[pr];
    const uint8_t *s1 = c.q + c.qoff[pr];
    const uint8_t *s2 = c.t + c.toff[pr];
    int2 *const gb = c.gbuf + (int64_t)blockIdx.x * c.gstride;
    int2 *const dmp = c.dump + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int BR = 64 * R;
    const int nb = (m + BR - 1) / BR;
    const int ngroups = (nb + LW - 1) / LW;
    const int bm = (m - 1) / BR, rem = (m - 1) % BR, tm = rem / R, rm = rem % R;
    const int2 *ring_in = ring + (w > 0 ? w - 1 : 0) * LRING;
    int2 *ring_out = ring + (w < LW - 1 ? w : 0) * LRING;
    auto left = [&](int g) -> int { return g < 0 ? -open : -open - g * gap - open; };
    const int G = n + 63 + LL * (LW - 1);
    const int nph = (G + LK - 1) / LK;
    int score = 0;
    bool have = false;

    for (int g = 0; g < ngroups; ++g) {
        const int b = g * LW + w;
        const bool live = b < nb;
        const int first_row = b * BR + t * R;
        int qo[R], HL[R], E[R];
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int row = first_row + i;
            qo[i] = (row < m ? (int)s1[row] : msize) * msize;
            HL[i] = left(row);
            E[i] = LNEG;
        }
        int prevRH = left(first_row - 1);
        int botH = LNEG, botF = LNEG;
        // lane 0's input for column jj: the DP's row above the query, the
        // previous group's last band (HBM) or this group's band above (ring)
        auto ld_in = [&](int jj) -> int2 {
            const int jc = jj < 0 ? 0 : (jj >= n ? n - 1 : jj);
            if (w == 0) {
                if (g == 0) return make_int2(-open - jc * gap - open, LNEG);
                return gb[jc];
            }
            return ring_in[jc & (LRING - 1)];
        };
        auto ld_t = [&](int jj) -> int { return (int)s2[jj < 0 ? 0 : (jj >= n ? n - 1 : jj)]; };
        const int s0 = -LL * w;  // this band's step at group step 0
        int2 in0 = ld_in(s0), in1 = ld_in(s0 + 1);
        int tc0 = ld_t(s0 - t), tc1 = ld_t(s0 + 1 - t);

        auto step = [&](const int s, int2 &bx, int &tq) __attribute__((always_inline)) {
            const int j = s - t;
            const int xH = shr1(botH), xF = shr1(botF);
            if (j >= 0 && j < n) {
                const int code = tq;
                int W[R];
#pragma unroll
                for (int i = 0; i < R; ++i) W[i] = lds[qo[i] + code];
                const int rH = t == 0 ? bx.x : xH;
                const int rF = t == 0 ? bx.y : xF;
                int diag = prevRH;
                prevRH = rH;
                int up = rH, F = rF;
#pragma unroll
                for (int i = 0; i < R; ++i) {
                    const int t1 = diag + W[i];
                    E[i] = max(E[i] - gap, HL[i]);
                    F = max(F - gap, up);
                    const int h = max(max(t1, E[i]), F);
                    diag = HL[i];
                    HL[i] = h - open;
                    up = HL[i];
                }
                botH = HL[R - 1];
                botF = F;
                if (t == 63 && w < LW - 1) ring_out[j & (LRING - 1)] = make_int2(botH, botF);
                int2 *dst = (t == 63 && w == LW - 1) ? gb + j : dmp;
                *dst = make_int2(botH, botF);
                if (b == bm && t == tm && j == n - 1) {
#pragma unroll
                    for (int i = 0; i < R; ++i)
                        if (i == rm) score = HL[i] + open;
                    have = true;
                }
            }
            bx = ld_in(s + 2);
            tq = ld_t(j + 2);
        };
        for (int ph = 0; ph < nph; ++ph) {
            if (live) {
                const int sb = ph * LK + s0;
                for (int k = 0; k < LK; k += 2) {
                    step(sb + k, in0, tc0);
                    step(sb + k + 1, in1, tc1);
                }
            }
            __syncthreads();
        }
    }
    if (have) {
        phip_out_t o;
        memset(&o, 0, sizeof(o));
        o.score = score;
        o.end_query = m - 1;
        o.end_ref = n - 1;
        o.status = PHIP_ST_W32;
        c.out[pr] = o;
    }
}

// ------------------------------------------------------------------ host

namespace {

constexpr int kLongR = 16;

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

bool nwlong_eligible(const phip_matrix_t &mx, int alg, int kind, int check, int m, int n)
{
    // the table and the rings must fit the 64 KiB of LDS a workgroup gets
    return alg == PHIP_NW && kind == PHIP_SCORE && !check && !mx.is_pssm && mx.size <= 24 &&
           m >= 1024 && n >= 64 && (int64_t)m * n >= ((int64_t)1 << 24);
}

size_t nwlong_scratch_bytes(int npairs, int maxn)
{
    return al256((size_t)npairs * ((size_t)maxn + 64) * 8) + al256((size_t)npairs * LW * 64 * 8);
}

int nwlong_launch(const DPParams &dp, const int32_t *dpairs, int npairs, int maxn, char *scratch,
                  hipStream_t st)
{
    if (npairs <= 0) return 0;
    LongCfg c;
    memset(&c, 0, sizeof(c));
    c.q = dp.q;
    c.qoff = dp.qoff;
    c.qlen = dp.qlen;
    c.t = dp.t;
    c.toff = dp.toff;
    c.tlen = dp.tlen;
    c.mat = dp.matrix;
    c.msize = dp.msize;
    c.open = dp.open;
    c.gap = dp.gap;
    c.pairs = dpairs;
    c.npairs = npairs;
    c.gstride = (int64_t)maxn + 64;
    c.gbuf = (int2 *)scratch;
    c.dump = (int2 *)(scratch + al256((size_t)npairs * c.gstride * 8));
    c.out = dp.out;
    const int tabn = (dp.msize + 1) * dp.msize;
    const size_t lds = (size_t)((tabn + 1) & ~1) * 4 + (size_t)(LW - 1) * LRING * 8;
    hipLaunchKernelGGL(k_nwlong<kLongR>, dim3(npairs), dim3(LW * 64), lds, st, c);
    HIP_OK(hipGetLastError());
    return 0;
}
