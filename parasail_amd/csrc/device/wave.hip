// wave.hip -- packed-int16 wavefront kernels for score-only SW / NW / SG.
//
// Layout (MI355X / gfx950, wave64):
//   * a "job" is two independent pairs A and B sharing every register: A in
//     the low 16 bits, B in the high 16 bits, so each v_pk_add_i16 /
//     v_pk_max_i16 advances two DP cells (parasail's int16 lanes,
//     src/sw_striped_avx2_256_16.c:288-314, become the two halves of a VGPR);
//   * a group of T consecutive lanes of one wave owns one job; lane t holds
//     R consecutive query rows in registers (a band of T*R rows);
//   * the group sweeps the target as a wavefront: at step s lane t computes
//     column j = s - t, receiving the bottom row (H-open, F) of lane t-1 by
//     one DPP wave_shr:1 per value; the vertical-gap dependency F therefore
//     never leaves registers (no lazy-F loop, no scan);
//   * queries longer than one band run as several bands; the bottom row of
//     a band goes through a per-job HBM buffer written by lane T-1 and read
//     by lane 0 of the next band;
//   * H is kept as H-open ("Hopen"): the diagonal term H_diag + W becomes
//     Hopen_diag + (W + open), so E, F and the next diagonal share one
//     subtraction.  Scores come from an LDS copy of the substitution table
//     already offset by +open.
//
// Exactness: int16 saturating arithmetic equals the int32 recurrence while
// every value stays inside the int16 window.  The planner only sends a pair
// here when its values are provably inside (NW/SG) or when an overflow is
// detected at run time from the running maximum (SW, the reference's
// maxp test, src/sw_striped_avx2_256_16.c:350-362); such pairs are re-run by
// the exact int32 tier.
#include <string.h>

#include <algorithm>
#include <vector>

#include "wave.h"

typedef short s2v __attribute__((ext_vector_type(2)));

namespace {

__device__ __forceinline__ int pk_add(int a, int b)
{
    return __builtin_bit_cast(int, __builtin_elementwise_add_sat(__builtin_bit_cast(s2v, a),
                                                                __builtin_bit_cast(s2v, b)));
}
__device__ __forceinline__ int pk_sub(int a, int b)
{
    return __builtin_bit_cast(int, __builtin_elementwise_sub_sat(__builtin_bit_cast(s2v, a),
                                                                __builtin_bit_cast(s2v, b)));
}
__device__ __forceinline__ int pk_max(int a, int b)
{
    return __builtin_bit_cast(int, __builtin_elementwise_max(__builtin_bit_cast(s2v, a),
                                                            __builtin_bit_cast(s2v, b)));
}
__device__ __forceinline__ int lo16(int v) { return (int)(short)(v & 0xffff); }
__device__ __forceinline__ int hi16(int v) { return v >> 16; }
__device__ __forceinline__ int pack16(int lo, int hi) { return (lo & 0xffff) | (hi << 16); }
__device__ __forceinline__ int clamp16(int v) { return v < -32768 ? -32768 : (v > 32767 ? 32767 : v); }
// value of lane-1 (lane 0 keeps `old`)
__device__ __forceinline__ int shr1(int old, int v)
{
    return __builtin_amdgcn_update_dpp(old, v, 0x138, 0xF, 0xF, false);
}

}  // namespace

#define PK_NEG ((int)0x80008000)
#define W_NEG (-12000)

// One job (two pairs).  Host-built, device-read.
struct WaveJob {
    int pair[2];      // pair index, -1 if the half is empty
    int m[2], n[2];   // lengths
    int s[2];         // virtual rows above row 0 (NW/SG bottom alignment)
    int nb;           // bands
    int ncols;        // max(n)
    int cclen;        // column-code stream length: the wave's widest job
    int64_t qoff[2];  // query offsets (device codes)
    int64_t ccoff;    // column-code stream offset
    int64_t bboff;    // band buffer offset (int2 units), nb-1 bands of ncols
    int64_t cdoff[2]; // lane candidates (int4 units: T per pair for SW, 2T for NW/SG)
};

struct WaveCfg {
    const WaveJob *jobs;
    int njobs;
    const int2 *wdims;        // per wave: (max bands, max columns)
    int nwaves;
    const uint32_t *colcode;  // per job column stream: (toffA | toffB << 16)
    const short *tab;         // bank-replicated table of W+open (see k_wave16)
    int tabn;                 // table size in dwords
    int qstride, ncopy;       // bytes per query-code row; copies (32 or 1)
    int VIRT, PADQ, PADT;
    int *next_wave;           // work-queue counter
    const uint8_t *q;
    int2 *dump;               // per-thread slot for stores a lane does not need
    int2 *bandbuf;
    int4 *cand;
    int open, gap, alg, sg_flags;
};

// Score table in LDS: for query code q and target code t the int16
// W(q,t)+open sits at byte  q*qstride + (t>>1)*128*(ncopy/32) + (t&1)*2 + c*4
// with c = lane % ncopy: with 32 copies every lane of a half-wave reads its
// own bank, so the random (q, t) lookups of a wave never conflict.
template <int R, int T, bool SW>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(R <= 12 ? 3 : 2))) k_wave16(WaveCfg c)
{
    extern __shared__ int tab32[];
    for (int k = threadIdx.x; k < c.tabn; k += blockDim.x) tab32[k] = ((const int *)c.tab)[k];
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int t = lane % T;
    const int open = c.open, gap = c.gap;
    const int openv = pack16(open, open), gapv = pack16(gap, gap);
    const bool s1_beg = c.alg == PHIP_SG && (c.sg_flags & PHIP_S1_BEG);
    const bool s2_beg = c.alg == PHIP_SG && (c.sg_flags & PHIP_S2_BEG);
    const bool zero_top = SW || s2_beg;     // row 0 of the DP is all zeros
    const bool zero_left = SW || s1_beg;    // column 0 of the DP is all zeros
    const char *tabb = (const char *)tab32 + (lane % c.ncopy) * 4;
    int2 *const dump = c.dump + (threadIdx.x + blockIdx.x * blockDim.x);

    // persistent: each wave takes the next group of 64/T jobs off the queue
    // (jobs are ordered longest first, so the tail is short)
    for (;;) {
        int wv = 0;
        if (lane == 0) wv = atomicAdd(c.next_wave, 1);
        wv = __builtin_amdgcn_readfirstlane(wv);
        if (wv >= c.nwaves) break;
        const int jobi = wv * (64 / T) + lane / T;
        const bool live = jobi < c.njobs;
        WaveJob J;
        if (live) {
            J = c.jobs[jobi];
        } else {
            memset(&J, 0, sizeof(J));
            J.pair[0] = J.pair[1] = -1;
        }
        const int2 wd = c.wdims[wv];

        // SW: job-level candidate per half
        int cS[2] = {-1, -1}, cC[2] = {0, 0}, cR[2] = {0, 0};
        // NW/SG: last-column maximum (rows 0..m-2) and last-row maximum per half
        int cmx[2] = {PHIP_NEG_INF32, PHIP_NEG_INF32}, crow[2] = {0, 0};
        int rmx[2] = {PHIP_NEG_INF32, PHIP_NEG_INF32}, rj[2] = {0, 0}, corner[2] = {0, 0};

        for (int b = 0; b < wd.x; ++b) {
            const bool band_live = live && b < J.nb;
            const bool last_band = b == J.nb - 1;
            int qo[R];  // table row offsets: half A in bits 0-15, half B in 16-31
            int HL[R], E[R];
            int first_row[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) first_row[h] = b * T * R + t * R - J.s[h];
#pragma unroll
            for (int i = 0; i < R; ++i) {
                int off[2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int g = first_row[h] + i;
                    int code;
                    if (J.pair[h] < 0) code = c.PADQ;
                    else if (g < 0) code = c.VIRT;
                    else if (g < J.m[h]) code = c.q[J.qoff[h] + g];
                    else code = c.PADQ;
                    off[h] = code * c.qstride;
                }
                qo[i] = off[0] | (off[1] << 16);
            }
            // left boundary (DP column 0) as H-open, per row and half
            auto left = [&](int g) -> int {
                if (g < 0 || zero_left) return -open;      // virtual rows copy H[0][0] = 0
                return clamp16(-open - g * gap - open);     // H[g+1][0] = -open - g*gap
            };
#pragma unroll
            for (int i = 0; i < R; ++i) {
                HL[i] = pack16(left(first_row[0] + i), left(first_row[1] + i));
                E[i] = PK_NEG;
            }
            int prevRH = pack16(left(first_row[0] - 1), left(first_row[1] - 1));
            if (b == 0 && t == 0) prevRH = pack16(-open, -open);  // H[0][0] = 0
            int botH = PK_NEG, botF = PK_NEG;
            int best = (int)0xFFFFFFFF, bcol = 0;
            int snap[SW ? R : 1];
#pragma unroll
            for (int i = 0; i < (SW ? R : 1); ++i) snap[i] = 0;

            const int steps = wd.y + T - 1;
            const int64_t bbstride = (J.ncols + 15) & ~15;
            const int2 *bbin = c.bandbuf + J.bboff + (int64_t)(b - 1) * bbstride;
            int2 *bbout = c.bandbuf + J.bboff + (int64_t)b * bbstride;
            const uint32_t *cc = c.colcode + J.ccoff;
            // column codes and band-boundary rows are fetched two steps ahead
            // with unconditional loads (the stream holds pad codes up to the
            // wave's widest job, band-boundary values past the job's columns
            // only feed columns whose results are unused), so the wait for
            // the value fetched earlier stays precise
            const int2 *bbsrc = (b > 0 && band_live) ? bbin : c.bandbuf;
            const int lastcc = J.cclen > 0 ? J.cclen - 1 : 0;
            const int lastc = (b > 0 && band_live && J.ncols > 0) ? J.ncols - 1 : 0;
            auto ld_code = [&](int jj) -> uint32_t {
                return cc[jj < 0 ? 0 : (jj > lastcc ? lastcc : jj)];
            };
            auto ld_bb = [&](int jj) -> int2 {
                return bbsrc[jj < 0 ? 0 : (jj > lastc ? lastc : jj)];
            };
            uint32_t code1 = ld_code(-t), code2 = ld_code(1 - t);
            int2 bb1 = ld_bb(-t), bb2 = ld_bb(1 - t);
            for (int s = 0; s < steps; ++s) {
                const int j = s - t;
                // exchange with lane t-1 while every lane is enabled
                const int xH = shr1(botH, botH);
                const int xF = shr1(botF, botF);
                const uint32_t code = code1;
                const int2 bb = bb1;
                code1 = code2;
                bb1 = bb2;
                code2 = ld_code(j + 2);
                bb2 = ld_bb(j + 2);
                if (j < 0 || j >= wd.y) continue;  // lane outside the band's columns
                const bool in = band_live && j < J.ncols;
                int rH = xH, rF = xF;
                if (t == 0) {
                    // group head: DP row 0 (band 0) or the previous band's bottom row
                    if (b == 0) {
                        const int v = zero_top ? -open : clamp16(-open - j * gap - open);
                        rH = pack16(v, v);
                        rF = PK_NEG;
                    } else {
                        rH = bb.x;
                        rF = bb.y;
                    }
                }
                int diag = prevRH;
                prevRH = rH;
                int up = rH, F = rF;
                const int toA = code & 0xffff, toB = code >> 16;
                int cm = 0;
#pragma unroll
                for (int i = 0; i < R; ++i) {
                    const short wa = *(const short *)(tabb + (qo[i] & 0xffff) + toA);
                    const short wb = *(const short *)(tabb + ((unsigned)qo[i] >> 16) + toB);
                    const int W = pack16(wa, wb);
                    const int t1 = pk_add(diag, W);
                    E[i] = pk_max(pk_sub(E[i], gapv), HL[i]);
                    F = pk_max(pk_sub(F, gapv), up);
                    int h = pk_max(pk_max(t1, E[i]), F);
                    if (SW) h = pk_max(h, 0);
                    diag = HL[i];
                    HL[i] = pk_sub(h, openv);
                    up = HL[i];
                    if (SW) cm = i == 0 ? h : pk_max(cm, h);
                }
                botH = HL[R - 1];
                botF = F;
                {
                    // unconditional store (inactive lanes write their own dump
                    // slot) keeps the loop free of memory ops under branches
                    const bool wr = !last_band && t == T - 1 && in;
                    int2 *dst = wr ? bbout + j : dump;
                    *dst = make_int2(botH, botF);
                }
                if (SW) {
                    const int nbst = pk_max(best, cm);
                    if (nbst != best) {
                        const int d = nbst ^ best;
                        const int mask = ((d & 0xffff) ? 0xffff : 0) | ((d >> 16) ? (int)0xffff0000 : 0);
#pragma unroll
                        for (int i = 0; i < (SW ? R : 1); ++i) snap[i] = (HL[i] & mask) | (snap[i] & ~mask);
                        bcol = (pack16(j, j) & mask) | (bcol & ~mask);
                        best = nbst;
                    }
                } else if (band_live) {
                    // the last query row is the bottom row of lane T-1 in the
                    // last band (bottom alignment): its running maximum (first
                    // column on ties) and its value at the last column
                    if (last_band && t == T - 1) {
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            const int v = (h ? hi16(botH) : lo16(botH)) + open;
                            if (j < J.n[h] && v > rmx[h]) { rmx[h] = v; rj[h] = j; }
                            if (j == J.n[h] - 1) corner[h] = v;
                        }
                    }
                    // the last column: largest value over rows 0..m-2, first row
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        if (j == J.n[h] - 1) {
#pragma unroll
                            for (int i = 0; i < R; ++i) {
                                const int g = first_row[h] + i;
                                const int v = (h ? hi16(HL[i]) : lo16(HL[i])) + open;
                                if (g >= 0 && g < J.m[h] - 1 && v > cmx[h]) { cmx[h] = v; crow[h] = g; }
                            }
                        }
                    }
                }
            }
            if (SW && band_live) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    if (J.pair[h] < 0) continue;
                    const int S = h ? hi16(best) : lo16(best);
                    const int col = h ? hi16(bcol) : lo16(bcol);
                    int row = -1;
#pragma unroll
                    for (int i = (SW ? R : 1) - 1; i >= 0; --i) {
                        const int g = first_row[h] + i;
                        const int v = (h ? hi16(snap[i]) : lo16(snap[i])) + open;
                        if (g >= 0 && g < J.m[h] && v == S) row = g;
                    }
                    if (row >= 0 && (S > cS[h] || (S == cS[h] && col < cC[h]))) {
                        cS[h] = S; cC[h] = col; cR[h] = row;
                    }
                }
            }
            __threadfence();
        }
        if (live) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (J.pair[h] < 0) continue;
                if (SW) {
                    c.cand[J.cdoff[h] + t] = make_int4(cS[h], cC[h], cR[h], 0);
                } else {
                    c.cand[J.cdoff[h] + 2 * t] = make_int4(cmx[h], crow[h], rmx[h], rj[h]);
                    c.cand[J.cdoff[h] + 2 * t + 1] = make_int4(corner[h], 0, 0, 0);
                }
            }
        }
    }
}

// SW epilogue: reduce T lane candidates per pair (max score, then min column,
// then min row: src/sw.c:150-158).  ovf_limit: the int16 tier's largest safe
// score; pairs above it are appended to `redo` for the int32 tier.
__global__ void k_sw_reduce(const int4 *cand, const int64_t *cdoff, const int32_t *pairs, int n,
                            int T, int ovf_limit, phip_out_t *out, int32_t *redo, int *nredo)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int4 *c = cand + cdoff[k];
    int S = -1, col = 0, row = 0;
    for (int t = 0; t < T; ++t) {
        int4 v = c[t];
        if (v.x > S || (v.x == S && (v.y < col || (v.y == col && v.z < row)))) {
            S = v.x; col = v.y; row = v.z;
        }
    }
    phip_out_t o;
    memset(&o, 0, sizeof(o));
    o.score = S;
    o.end_query = row;
    o.end_ref = col;
    out[pairs[k]] = o;
    if (S > ovf_limit) redo[atomicAdd(nredo, 1)] = pairs[k];
}

__global__ void k_zero(int *p, int n)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0;
}

// NW / SG epilogue from the lane candidates (src/sg.c:171-235,
// src/nw.c:155-157): per lane the last-column maximum over rows 0..m-2 with
// its first row; lane T-1 also holds the last-row maximum with its first
// column and the corner H[m-1][n-1].
__global__ void k_sg_reduce(const int4 *cand, const int64_t *cdoff, const int32_t *pairs,
                            const int32_t *mlen, const int32_t *nlen, int n, int T, int alg,
                            int flags, phip_out_t *out)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int4 *c = cand + cdoff[k];
    const int m = mlen[k], nn = nlen[k];
    int cm = PHIP_NEG_INF32, ci = 0;
    for (int t = 0; t < T; ++t) {
        const int4 v = c[2 * t];
        if (v.x > cm || (v.x == cm && v.y < ci)) { cm = v.x; ci = v.y; }
    }
    const int4 last = c[2 * (T - 1)];
    const int rm = last.z, rj = last.w, corner = c[2 * (T - 1) + 1].x;
    const bool s1e = alg == PHIP_SG && (flags & PHIP_S1_END);
    const bool s2e = alg == PHIP_SG && (flags & PHIP_S2_END);
    int score = corner, eq = m - 1, er = nn - 1;
    if (s1e && s2e) {
        const int S = imax(cm, rm);
        score = S;
        if (rm == S && rj < nn - 1) { eq = m - 1; er = rj; }
        else if (m > 1 && cm == S) { eq = ci; er = nn - 1; }
    } else if (s1e) {
        if (m > 1 && cm >= corner) { score = cm; eq = ci; }
    } else if (s2e) {
        score = rm; er = rj;
    }
    phip_out_t o;
    memset(&o, 0, sizeof(o));
    o.score = score;
    o.end_query = eq;
    o.end_ref = er;
    out[pairs[k]] = o;
}

// Column-code stream: per job and column, the byte offsets of the two
// target residues' columns in the LDS table.
__global__ void k_colcode(const WaveJob *jobs, int njobs, const uint8_t *t, const int64_t *toff,
                          int PADT, int ncopy, uint32_t *colcode)
{
    const int jb = blockIdx.x;
    if (jb >= njobs) return;
    const WaveJob J = jobs[jb];
    const uint8_t *ta = J.pair[0] >= 0 ? t + toff[J.pair[0]] : nullptr;
    const uint8_t *tb = J.pair[1] >= 0 ? t + toff[J.pair[1]] : nullptr;
    for (int j = threadIdx.x; j < J.cclen; j += blockDim.x) {
        uint32_t a = (ta && j < J.n[0]) ? ta[j] : PADT;
        uint32_t bb = (tb && j < J.n[1]) ? tb[j] : PADT;
        const uint32_t oa = (a >> 1) * ncopy * 4 + (a & 1) * 2;
        const uint32_t ob = (bb >> 1) * ncopy * 4 + (bb & 1) * 2;
        colcode[J.ccoff + j] = oa | (ob << 16);
    }
}

// ------------------------------------------------------------------ host

namespace {

const int kRs[] = {4, 8, 12, 16, 20, 24, 32};
constexpr int kT = 16;

struct Shape {
    int R, nb;
};

// Band height T*R and band count for a query of m rows against n columns:
// the cheapest (rows computed x steps) cover of the query.
Shape choose(int m, int n)
{
    Shape best{32, (m + kT * 32 - 1) / (kT * 32)};
    double bestc = 1e300;
    const int maxnb = std::max(1, (m + kT * 4 - 1) / (kT * 4));
    for (int nb = 1; nb <= maxnb && nb <= 256; ++nb) {
        for (int R : kRs) {
            if ((int64_t)nb * kT * R < m) continue;
            // VALU work: ~13 ops per row and ~40 per step of each lane
            double cost = (double)nb * (n + kT - 1) * (13.0 * R + 40.0);
            if (cost < bestc) { bestc = cost; best = Shape{R, nb}; }
            break;
        }
    }
    return best;
}

typedef void (*KFn)(WaveCfg);

template <int R> KFn pick(bool sw) { return sw ? k_wave16<R, kT, true> : k_wave16<R, kT, false>; }

KFn kernel_for(int R, bool sw)
{
    switch (R) {
    case 4: return pick<4>(sw);
    case 8: return pick<8>(sw);
    case 12: return pick<12>(sw);
    case 16: return pick<16>(sw);
    case 20: return pick<20>(sw);
    case 24: return pick<24>(sw);
    default: return pick<32>(sw);
    }
}

// Whether every DP value of an m x n pair provably stays inside int16
// (NW/SG), or can be checked at run time (SW, from the running maximum).
bool int16_fits(const phip_matrix_t &mx, int alg, int open, int gap, int m, int n)
{
    if (mx.is_pssm) return false;
    if (mx.max + open > 16000 || mx.min < -16000 || open > 8000 || gap > 8000) return false;
    if (m > 30000 || n > 30000) return false;
    if (alg == PHIP_SW) return true;
    int64_t lo = -(int64_t)4 * open - (int64_t)(m + n) * gap - std::abs(mx.min) - 64;
    int64_t hi = (int64_t)std::min(m, n) * std::max(mx.max, 0) + mx.max + open + 64;
    return lo > -32000 && hi < 32000;
}

struct Launch {
    int R;
    size_t j0;
    int njr, nwv;
    size_t wd0;
};

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

struct WavePlan {
    int alg, sg_flags, open, gap;
    bool sw;
    std::vector<Launch> launches;
    int nfast = 0, nslow = 0;
    int tabrows = 0, TS = 0, VIRT = 0, PADQ = 0, PADT = 0, ovf_limit = 0;
    int tabn = 0, qstride = 0, ncopy = 1;
    int maxn = 0, simple_lanes = 0;
    char *base = nullptr;
    WaveJob *djobs = nullptr;
    short *dtab = nullptr;
    uint32_t *dcc = nullptr;
    int2 *dbb = nullptr, *dwd = nullptr;
    int4 *dcd = nullptr;
    int32_t *dpairs = nullptr, *dm = nullptr, *dn = nullptr, *dslow = nullptr, *dredo = nullptr;
    int64_t *dcdo = nullptr;
    int2 *ddump = nullptr;
    int *dcount = nullptr, *dscratch = nullptr;
};

bool wave_plan_matches(const WavePlan *p, int alg, int sg_flags, int open, int gap)
{
    return p->alg == alg && p->sg_flags == sg_flags && p->open == open && p->gap == gap;
}

void wave_plan_free(WavePlan *p)
{
    if (!p) return;
    dev_free(p->base);
    delete p;
}

WavePlan *wave_plan_create(const HostBatch &hb, const DPParams &dp, hipStream_t stream)
{
    WavePlan *P = new WavePlan();
    P->alg = dp.alg;
    P->sg_flags = dp.sg_flags;
    P->open = dp.open;
    P->gap = dp.gap;
    P->sw = dp.alg == PHIP_SW;
    const bool sw = P->sw;
    const phip_matrix_t &mx = *hb.mx;
    const int np = hb.np;
    std::vector<int> fast, slow;
    std::vector<Shape> shape(np);
    for (int i = 0; i < np; ++i) {
        P->maxn = std::max(P->maxn, (int)hb.tlen[i]);
        if (int16_fits(mx, dp.alg, dp.open, dp.gap, hb.qlen[i], hb.tlen[i])) {
            fast.push_back(i);
            shape[i] = choose(hb.qlen[i], hb.tlen[i]);
        } else {
            slow.push_back(i);
        }
    }
    // per R (one kernel instance each), longest work first so that the
    // queue ends on short jobs
    std::sort(fast.begin(), fast.end(), [&](int x, int y) {
        if (shape[x].R != shape[y].R) return shape[x].R < shape[y].R;
        if (shape[x].nb != shape[y].nb) return shape[x].nb > shape[y].nb;
        return hb.tlen[x] > hb.tlen[y];
    });
    // LDS table: rows [0,size) real, VIRT, PADQ; columns [0,size) real, PADT
    const int size = mx.size;
    P->TS = size + 1;
    P->VIRT = size;
    P->PADQ = size + 1;
    P->PADT = size;
    P->tabrows = size + 2;
    // bank-replicated layout (32 copies) when it fits comfortably in LDS
    const int tsp = (P->TS + 1) & ~1;
    P->ncopy = ((size_t)P->tabrows * (tsp / 2) * 32 * 4 <= (48u << 10)) ? 32 : 1;
    P->qstride = (tsp / 2) * P->ncopy * 4;
    P->tabn = P->tabrows * (tsp / 2) * P->ncopy;
    std::vector<short> htab((size_t)P->tabn * 2, (short)W_NEG);
    const bool s2_beg = dp.alg == PHIP_SG && (dp.sg_flags & PHIP_S2_BEG);
    if (!mx.is_pssm) {
        for (int r = 0; r < P->tabrows; ++r)
            for (int col = 0; col < tsp; ++col) {
                int v = W_NEG;
                if (col >= P->PADT || r == P->PADQ) v = W_NEG;
                else if (r == P->VIRT) v = (sw || s2_beg) ? dp.open : W_NEG;
                else v = mx.values[(size_t)r * size + col] + dp.open;
                for (int cp = 0; cp < P->ncopy; ++cp) {
                    size_t byte = (size_t)r * P->qstride + (size_t)(col >> 1) * P->ncopy * 4 +
                                  (col & 1) * 2 + cp * 4;
                    htab[byte / 2] = (short)v;
                }
            }
    }
    P->ovf_limit = 32767 - (mx.max + dp.open + 1);
    // jobs: consecutive pairs of the same shape share one job (A low, B high)
    std::vector<WaveJob> jobs;
    std::vector<int> job_R;
    std::vector<int32_t> rpairs, rm, rn;
    std::vector<int64_t> rcd;
    int64_t ccoff = 0, bboff = 0, cdoff = 0;
    size_t k = 0;
    while (k < fast.size()) {
        WaveJob J;
        memset(&J, 0, sizeof(J));
        const Shape S = shape[fast[k]];
        int halves = 1;
        if (k + 1 < fast.size() && shape[fast[k + 1]].R == S.R && shape[fast[k + 1]].nb == S.nb)
            halves = 2;
        J.nb = S.nb;
        J.pair[1] = -1;
        J.cdoff[1] = -1;
        for (int h = 0; h < halves; ++h) {
            const int p = fast[k + h];
            J.pair[h] = p;
            J.m[h] = hb.qlen[p];
            J.n[h] = hb.tlen[p];
            J.s[h] = sw ? 0 : S.nb * kT * S.R - J.m[h];
            J.qoff[h] = hb.qoff[p];
            J.ncols = std::max(J.ncols, J.n[h]);
            J.cdoff[h] = cdoff;
            cdoff += sw ? kT : 2 * kT;
            rcd.push_back(J.cdoff[h]);
            rm.push_back(J.m[h]);
            rn.push_back(J.n[h]);
            rpairs.push_back(p);
        }
        J.bboff = bboff;
        bboff += (int64_t)(J.nb - 1) * ((J.ncols + 15) & ~15);
        jobs.push_back(J);
        job_R.push_back(S.R);
        k += halves;
    }
    // launches: one per R (jobs are sorted by R); per-wave loop bounds
    const int jpw = 64 / kT;
    std::vector<int2> wdims;
    size_t j0 = 0;
    while (j0 < jobs.size()) {
        size_t j1 = j0;
        while (j1 < jobs.size() && job_R[j1] == job_R[j0]) ++j1;
        Launch L;
        L.R = job_R[j0];
        L.j0 = j0;
        L.njr = (int)(j1 - j0);
        L.nwv = (L.njr + jpw - 1) / jpw;
        L.wd0 = wdims.size();
        for (int w = 0; w < L.nwv; ++w) {
            int mb = 0, mc = 0;
            for (int q = 0; q < jpw; ++q) {
                size_t jj = j0 + (size_t)w * jpw + q;
                if (jj >= j1) break;
                mb = std::max(mb, jobs[jj].nb);
                mc = std::max(mc, jobs[jj].ncols);
            }
            wdims.push_back(make_int2(mb, mc));
            // every job of the wave streams column codes up to the wave's
            // widest job (pad codes past its own columns)
            for (int q = 0; q < jpw; ++q) {
                size_t jj = j0 + (size_t)w * jpw + q;
                if (jj >= j1) break;
                jobs[jj].cclen = mc;
                jobs[jj].ccoff = ccoff;
                ccoff += (mc + 63) & ~63;
            }
        }
        P->launches.push_back(L);
        j0 = j1;
    }
    P->nfast = (int)rpairs.size();
    P->nslow = (int)slow.size();
    // int32-tier scratch: rows of up to maxn+1 columns per lane
    const size_t per_lane = (size_t)(P->maxn + 1) * 2 * sizeof(int);
    P->simple_lanes = (int)std::max<size_t>(64, std::min<size_t>(8192, ((size_t)512 << 20) / per_lane));
    // sizes and one allocation
    const size_t nj = jobs.size();
    size_t off = 0;
    auto slot = [&](size_t bytes) { size_t o = off; off += al256(bytes ? bytes : 1); return o; };
    size_t o_jobs = slot(nj * sizeof(WaveJob)), o_tab = slot(htab.size() * 2),
           o_cc = slot((size_t)ccoff * 4), o_bb = slot((size_t)bboff * 8),
           o_dump = slot((size_t)1024 * 256 * 8), o_cd = slot((size_t)cdoff * 16),
           o_pairs = slot(rpairs.size() * 4), o_m = slot(rm.size() * 4), o_n = slot(rn.size() * 4),
           o_cdo = slot(rcd.size() * 8),
           o_wd = slot(wdims.size() * 8), o_slow = slot(slow.size() * 4),
           o_redo = slot(rpairs.size() * 4), o_count = slot(64),
           o_scr = slot(per_lane * P->simple_lanes);
    P->base = (char *)dev_alloc(off);
    if (!P->base) { delete P; return nullptr; }
    char *b = P->base;
    P->djobs = (WaveJob *)(b + o_jobs);
    P->dtab = (short *)(b + o_tab);
    P->dcc = (uint32_t *)(b + o_cc);
    P->dbb = (int2 *)(b + o_bb);
    P->ddump = (int2 *)(b + o_dump);
    P->dcd = (int4 *)(b + o_cd);
    P->dpairs = (int32_t *)(b + o_pairs);
    P->dm = (int32_t *)(b + o_m);
    P->dn = (int32_t *)(b + o_n);

    P->dcdo = (int64_t *)(b + o_cdo);
    P->dwd = (int2 *)(b + o_wd);
    P->dslow = (int32_t *)(b + o_slow);
    P->dredo = (int32_t *)(b + o_redo);
    P->dcount = (int *)(b + o_count);
    P->dscratch = (int *)(b + o_scr);
    auto up = [&](void *dst, const void *src, size_t bytes) -> int {
        if (!bytes) return 0;
        HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
        return 0;
    };
    int rc = 0;
    rc |= up(P->djobs, jobs.data(), nj * sizeof(WaveJob));
    rc |= up(P->dtab, htab.data(), htab.size() * 2);
    rc |= up(P->dpairs, rpairs.data(), rpairs.size() * 4);
    rc |= up(P->dm, rm.data(), rm.size() * 4);
    rc |= up(P->dn, rn.data(), rn.size() * 4);

    rc |= up(P->dcdo, rcd.data(), rcd.size() * 8);
    rc |= up(P->dwd, wdims.data(), wdims.size() * 8);
    rc |= up(P->dslow, slow.data(), slow.size() * 4);
    if (nj) {
        hipLaunchKernelGGL(k_colcode, dim3((unsigned)nj), dim3(256), 0, stream, P->djobs, (int)nj,
                           dp.t, dp.toff, P->PADT, P->ncopy, P->dcc);
        if (hipGetLastError() != hipSuccess) rc = -1;
    }
    // the pageable host vectors die here: wait for the copies
    if (hipStreamSynchronize(stream) != hipSuccess) rc = -1;
    if (rc) {
        phip_set_error("wave_plan_create: upload failed");
        wave_plan_free(P);
        return nullptr;
    }
    return P;
}

namespace {

// Side streams on which the per-R kernels of one pass run concurrently, so
// one kernel's tail overlaps the next one's start.
constexpr int kSide = 4;
hipStream_t g_side[kSide];
hipEvent_t g_fork, g_join[kSide];
bool g_side_ok = false;

int side_streams()
{
    if (g_side_ok) return 0;
    for (int k = 0; k < kSide; ++k) {
        HIP_OK(hipStreamCreateWithFlags(&g_side[k], hipStreamNonBlocking));
        HIP_OK(hipEventCreateWithFlags(&g_join[k], hipEventDisableTiming));
    }
    HIP_OK(hipEventCreateWithFlags(&g_fork, hipEventDisableTiming));
    g_side_ok = true;
    return 0;
}

}  // namespace

int wave_plan_run(WavePlan *P, const DPParams &dp, hipStream_t stream)
{
    if (side_streams() != 0) return -1;
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, stream, P->dcount, 16);
    HIP_OK(hipEventRecord(g_fork, stream));
    const int nside = std::min<int>(kSide, (int)P->launches.size());
    for (int k = 0; k < nside; ++k) HIP_OK(hipStreamWaitEvent(g_side[k], g_fork, 0));
    int li = 0;
    for (const Launch &L : P->launches) {
        WaveCfg cfg;
        memset(&cfg, 0, sizeof(cfg));
        cfg.jobs = P->djobs + L.j0;
        cfg.njobs = L.njr;
        cfg.wdims = P->dwd + L.wd0;
        cfg.nwaves = L.nwv;
        cfg.next_wave = P->dcount + 1 + li;
        cfg.colcode = P->dcc;
        cfg.tab = P->dtab;
        cfg.tabn = P->tabn;
        cfg.qstride = P->qstride;
        cfg.ncopy = P->ncopy;
        cfg.VIRT = P->VIRT;
        cfg.PADQ = P->PADQ;
        cfg.PADT = P->PADT;
        cfg.q = dp.q;
        cfg.dump = P->ddump;
        cfg.bandbuf = P->dbb;
        cfg.cand = P->dcd;
        cfg.open = P->open;
        cfg.gap = P->gap;
        cfg.alg = P->alg;
        cfg.sg_flags = P->sg_flags;
        const int wpb = 4;
        const int blocks = std::min((L.nwv + wpb - 1) / wpb, 1024);
        hipLaunchKernelGGL(kernel_for(L.R, P->sw), dim3(blocks), dim3(64 * wpb),
                           (size_t)cfg.tabn * 4, g_side[li % kSide], cfg);
        HIP_OK(hipGetLastError());
        ++li;
    }
    for (int k = 0; k < nside; ++k) {
        HIP_OK(hipEventRecord(g_join[k], g_side[k]));
        HIP_OK(hipStreamWaitEvent(stream, g_join[k], 0));
    }
    if (P->nfast) {
        if (P->sw) {
            hipLaunchKernelGGL(k_sw_reduce, dim3((P->nfast + 255) / 256), dim3(256), 0, stream,
                               P->dcd, P->dcdo, P->dpairs, P->nfast, kT, P->ovf_limit, dp.out,
                               P->dredo, P->dcount);
        } else {
            hipLaunchKernelGGL(k_sg_reduce, dim3((P->nfast + 255) / 256), dim3(256), 0, stream,
                               P->dcd, P->dcdo, P->dpairs, P->dm, P->dn, P->nfast, kT, P->alg,
                               P->sg_flags, dp.out);
        }
        HIP_OK(hipGetLastError());
    }
    // int32 tier: pairs outside the int16 window, then SW pairs that overflowed
    const int lanes = P->simple_lanes;
    const dim3 sg((lanes + 255) / 256), sb(256);
    if (P->nslow) {
        hipLaunchKernelGGL(k_simple, sg, sb, 0, stream, dp, 0, P->nslow, P->dslow, (const int *)nullptr,
                           P->dscratch, lanes, P->maxn);
        HIP_OK(hipGetLastError());
    }
    if (P->sw && P->nfast) {
        hipLaunchKernelGGL(k_simple, sg, sb, 0, stream, dp, 0, P->nfast, P->dredo, P->dcount,
                           P->dscratch, lanes, P->maxn);
        HIP_OK(hipGetLastError());
    }
    return 0;
}
