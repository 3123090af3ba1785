// wave.hip -- packed-int16 wavefront kernels for score-only SW / NW / SG.
//
// Layout (MI355X / gfx950, wave64):
//   * a "job" is two independent pairs A and B sharing every register: A in
//     the low 16 bits, B in the high 16 bits, so each v_pk_add_i16 /
//     v_pk_max_i16 advances two DP cells (parasail's int16 lanes,
//     src/sw_striped_avx2_256_16.c:288-314, become the two halves of a VGPR);
//   * a chain of T consecutive lanes of one wave owns one job (T = 16: four
//     jobs per wave; T = 64: the whole wave, for the largest jobs); lane t
//     holds R consecutive query rows in registers (a band of T*R rows);
//   * the chain sweeps the target as a wavefront: at step s lane t computes
//     column j = s - t, receiving the bottom row (H-open, F) of lane t-1 by
//     one DPP wave_shr:1 per value; the vertical-gap dependency F therefore
//     never leaves registers (no lazy-F loop, no scan);
//   * queries longer than one band run as several bands; the bottom row of
//     a band goes through a per-job HBM row written by lane T-1 and read by
//     lane 0 of the next band (band 0 reads the DP's row 0 from a table);
//   * NW/SG keep H as H-open ("Hopen"): the diagonal term H_diag + W
//     becomes Hopen_diag + (W + open), so E, F and the next diagonal share
//     one subtraction (the LDS table holds W + open).  SW keeps every value
//     clamped at 0 instead -- H, max(H - open, 0) and E, F >= 0 with
//     unsigned saturating subtraction -- so max(diag + W, E, F) is already
//     the local max(..., 0) and needs no clamp of its own (the table holds
//     W); values below 0 never reach a local score, so the result is the
//     same.
//   * one persistent launch per pass: waves take work descriptors off a
//     queue ordered by decreasing latency, so the longest jobs start first
//     and the short ones fill the machine around them; R (rows per lane) is
//     chosen per wave and dispatched inside the kernel.
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

#include "dp32.h"
#include "pk.h"
#include "wave.h"

namespace {

__host__ __device__ __forceinline__ int clamp16(int v) { return v < -32768 ? -32768 : (v > 32767 ? 32767 : v); }

}  // namespace

// largest rows per lane (the SW snapshot area in LDS is sized for it)
constexpr int kSnapRows = 20;
// rows between a row's table reads and its use (2 kLA - 2 reads in flight
// behind each: at most 15 can be waited on exactly)
constexpr int kLA = 8;

// two consecutive column codes of a job's stream, read as one 16-byte load
// (at any 8-byte boundary)
struct __attribute__((aligned(8))) CodePair {
    unsigned x, y, z, w;
};

// One job (two pairs).  Host-built, device-read.
struct WaveJob {
    int pair[2];      // pair index, -1 if the half is empty
    int m[2], n[2];   // lengths
    int s[2];         // virtual rows above row 0 (NW/SG bottom alignment)
    int nb, pad;      // bands of this job
    int64_t qoff[2];  // query offsets (device codes)
    int64_t ccoff;    // column-code stream: T pads, the wave's width, pads
    int64_t bboff;    // band rows (int2), bbw entries each
    int64_t cdoff[2]; // lane candidates (int4 units, T per pair)
};

// One queue entry: up to 64/T jobs that share R, T, the band count and the
// loop bounds.
struct WaveDesc {
    int job0, njobs, R, T;
    int nb, width, bbw, prio;
    int ext;  // NW/SG: some pair's saturation window needs the cells' extrema
};

struct WaveCfg {
    const WaveJob *jobs;
    const WaveDesc *waves;
    int nwaves;
    const uint2 *colcode;     // per job column stream: x = the dwords holding
                              // the two residues' table entries (toffA |
                              // toffB << 16), y = the v_perm selector taking
                              // each entry's half
    const short *tab;         // bank-replicated table of W+open (see k_wave)
    int tabn;                 // table size in dwords
    int qstride, ncopy;       // bytes per query-code row; copies (32 or 1)
    int VIRT, PADQ;
    int *next_wave;           // work-queue counter
    const uint8_t *q;
    int2 *bandbuf;
    const int2 *toprow;       // DP row 0 as (H-open, F) per column, both halves
    int4 *cand;
    int open, gap, alg, sg_flags;
};

// Per-pair reference for the epilogues.
struct PairRef {
    int64_t cdoff;    // lane candidates
    int64_t lastrow;  // NW/SG: band row holding the DP's last row (int2 units)
    int pair, m, n, T, half, pad;
};

namespace {

// All bands of one queue entry, R rows per lane.  EXT (NW/SG): also the
// extrema the saturation windows test -- the largest H and the smallest E
// and F (H is never below either) -- per lane over its pairs' columns.  The
// virtual rows above row 0 change neither: their H and E repeat row 0 (H 0,
// E -open when row 0 is free) and their F is row 1's (-open when free), so
// their H is at most 0 and their E and F are values of the boundary row or
// of real cells, which the reduction counts anyway.
template <int R, bool SW, bool EXT>
__device__ __forceinline__ void wave_run(const WaveCfg &c, const WaveDesc &wd, const char *tabb,
                                         int lane, short *snp)
{
    constexpr int LA = kLA < R ? kLA : R;  // read-ahead distance in rows
    const int T = wd.T;
    const int t = lane & (T - 1);
    const int jslot = T == 64 ? 0 : lane >> 4;
    const WaveJob J = c.jobs[wd.job0 + (jslot < wd.njobs ? jslot : 0)];
    const bool live = jslot < wd.njobs;
    const int open = c.open, gap = c.gap;
    const int openv = pk::pack(open, open), gapv = pk::pack(gap, gap);
    const bool s1_beg = c.alg == PHIP_SG && (c.sg_flags & PHIP_S1_BEG);
    const bool zero_left = SW || s1_beg;  // column 0 of the DP is all zeros
    const int width = wd.width;
    const int steps = width + T - 1;

    // SW: job-level candidate per half; NW/SG: last-column maximum over rows
    // 0..m-2 (first row on ties) per half
    int cS[2] = {-1, -1}, cC[2] = {0, 0}, cR[2] = {0, 0};
    int cmx[2] = {PHIP_NEG_INF32, PHIP_NEG_INF32}, crow[2] = {0, 0};
    int vmx = 0, vmn = 0;  // EXT: packed largest H, smallest E/F (from 0)

    for (int b = 0; b < wd.nb; ++b) {
        const bool band_live = live && b < J.nb;
        int qo[R];  // table row offsets: half A in bits 0-15, half B in 16-31
        // NW/SG: HL = H - open.  SW: every value is clamped at 0 (values
        // below 0 never reach a local score), Hs = H, HL = max(H - open, 0),
        // E and F >= 0, so H = max(diag + W, E, F) needs no clamp of its own
        int HL[R], E[R];
        int Hs[SW ? R : 1];
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
                if (J.pair[h] < 0 || !live) code = c.PADQ;
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
            if (SW) {
                HL[i] = 0;
                Hs[i] = 0;
                E[i] = 0;
            } else {
                HL[i] = pk::pack(left(first_row[0] + i), left(first_row[1] + i));
                E[i] = PK_NEG;
            }
        }
        // diagonal of the lane's first row: SW the H of the row above (0 in
        // column 0), NW/SG its H-open
        int prevRH = SW ? 0 : pk::pack(left(first_row[0] - 1), left(first_row[1] - 1));
        if (!SW && b == 0 && t == 0) prevRH = pk::pack(-open, -open);  // H[0][0] = 0
        int botH = SW ? 0 : PK_NEG, botF = SW ? 0 : PK_NEG;
        int best = (int)0xFFFFFFFF;
        // NW/SG: the lane's rows at each half's last column (SW keeps its
        // snapshot, the rows at the column of the running maximum, in LDS)
        int snap[SW ? 1 : R];
#pragma unroll
        for (int i = 0; i < (SW ? 1 : R); ++i) snap[i] = 0;
        int bcolA = 0, bcolB = 0;
        const int lastA = J.n[0] - 1, lastB = J.n[1] - 1;
        // EXT: the last column whose values count, per half (none past the
        // job's bands, whose rows are not the pair's)
        const int eA = band_live ? lastA : -1, eB = band_live ? lastB : -1;

        // column codes: ccl[k] is the code pair of column k - t; the stream
        // holds T leading and 8 trailing pad codes around the wave's width
        const uint2 *ccl = c.colcode + J.ccoff + (T - t);
        // row above the band, indexed by step (lane 0 reads column == step)
        const int2 *bbs = (b == 0 || !band_live) ? c.toprow
                                                : c.bandbuf + J.bboff + (int64_t)(b - 1) * wd.bbw;
        const bool wr_lane = t == T - 1 && band_live && (!SW || b < J.nb - 1);
        int2 *const bbr = c.bandbuf + J.bboff + (int64_t)b * wd.bbw;

        // the rows' scores against one column: W + open (SW: W) of both
        // halves from the LDS table, at the code pair's byte offsets.  Row
        // k's two table dwords are read LA rows before row k uses them (for
        // the rows below LA, during the previous column's last rows), so at
        // most 2 LA reads are in flight and each use waits for its own
        // reads only; the v_perm that takes each half's entry runs at the use
        unsigned wa[R], wb[R];
        auto issue = [&](const int k, const uint2 code) __attribute__((always_inline)) {
            wa[k] = *(const unsigned *)(tabb + (qo[k] & 0xffff) + (code.x & 0xffff));
            wb[k] = *(const unsigned *)(tabb + ((unsigned)qo[k] >> 16) + (code.x >> 16));
        };
        // after row k's entries are used: the read LA rows further down
        // (this column's, or the next column's top rows)
        auto refill = [&](const int k, const uint2 cur, const uint2 nxt) __attribute__((always_inline)) {
            if (k + LA < R) issue(k + LA, cur);
            else issue(k + LA - R, nxt);
        };

        // one DP column.  p holds this column's code pair (x, y) and the
        // next column's (z, w), fetched two steps earlier; bx, the row
        // above for lane 0, is fetched two steps ahead the same way.
        // FAST: every lane's column lies in [0, width) (the steady state)
        auto step = [&](auto fastc, const int sv, CodePair &p, int2 &bx) __attribute__((always_inline)) {
            constexpr bool FAST = decltype(fastc)::value;
            const int j = sv - t;
            const uint2 cc = make_uint2(p.x, p.y), cn = make_uint2(p.z, p.w);
            const int2 bxs = bx;
            // this step's codes and row above (loaded two steps ago) are
            // waited for before the loads for two steps ahead are issued:
            // vector loads may return out of order, so the wait for one is
            // a wait for every load in flight (vmcnt(0)), which would
            // otherwise include the ones just issued
            asm volatile("" ::"v"(cc.x), "v"(cc.y), "v"(cn.x), "v"(bxs.x), "v"(bxs.y));
            __builtin_amdgcn_sched_barrier(0);
            p = *(const CodePair *)(ccl + sv + 2);
            bx = bbs[sv + 2];
            auto wval = [&](const int k) __attribute__((always_inline)) {
                return (int)__builtin_amdgcn_perm(wb[k], wa[k], cc.y);
            };
            const int xH = pk::shr1(botH);  // every lane enabled
            const int xF = pk::shr1(botF);
            if (FAST || (j >= 0 && j < width)) {
                const int rH = t == 0 ? bxs.x : xH;
                const int rF = t == 0 ? bxs.y : xF;
                int t1 = pk::add(prevRH, wval(0));
                refill(0, cc, cn);
                prevRH = rH;
                // SW passes H between lanes and bands, the others H - open
                int up = SW ? pk::subu(rH, openv) : rH, F = rF;
                int cm = 0;
                int smx = 0, smn = 0;  // EXT: the column's largest H, smallest E/F
                // each HL[i] (SW: Hs[i]) is read (E, next row's diagonal)
                // before it is rewritten, so it stays in its register
#pragma unroll
                for (int i = 0; i < R; ++i) {
                    if (SW) {
                        E[i] = pk::max(pk::subu(E[i], gapv), HL[i]);
                        const int t1n = i + 1 < R ? pk::add(Hs[i], wval(i + 1)) : 0;
                        if (i + 1 < R) refill(i + 1, cc, cn);
                        F = pk::max(pk::subu(F, gapv), up);
                        const int h = pk::max(pk::max(t1, E[i]), F);
                        Hs[i] = h;
                        HL[i] = pk::subu(h, openv);
                        up = HL[i];
                        t1 = t1n;
                    } else {
                        E[i] = pk::max(pk::sub(E[i], gapv), HL[i]);
                        const int t1n = i + 1 < R ? pk::add(HL[i], wval(i + 1)) : 0;
                        if (i + 1 < R) refill(i + 1, cc, cn);
                        F = pk::max(pk::sub(F, gapv), up);
                        const int h = pk::max(pk::max(t1, E[i]), F);
                        HL[i] = pk::sub(h, openv);
                        up = HL[i];
                        t1 = t1n;
                        if (EXT) {
                            const int ef = pk::min(E[i], F);
                            smx = i == 0 ? h : pk::max(smx, h);
                            smn = i == 0 ? ef : pk::min(smn, ef);
                        }
                    }
                }
                if (EXT && !SW) {
                    // a half past its target's last column keeps no value
                    const int M = (j <= eA ? 0xffff : 0) | (j <= eB ? (int)0xffff0000 : 0);
                    vmx = pk::max(vmx, (smx & M) | (PK_NEG & ~M));
                    vmn = pk::min(vmn, (smn & M) | (PK_POS & ~M));
                }
                botH = SW ? Hs[R - 1] : HL[R - 1];
                botF = F;
                if (wr_lane) bbr[j] = make_int2(botH, botF);
                if (SW) {
                    // the rows' maximum as a tree of three-way maxima
                    // (independent maxima issue back to back; a serial chain
                    // would wait on every result).  Every H is >= 0, and
                    // below 0x7C00 unless the pair passes the overflow limit
                    // (then the maximum is at or above 0x7C00 as well)
                    int mx[R];
#pragma unroll
                    for (int i = 0; i < R; ++i) mx[i] = Hs[SW ? i : 0];
                    int nm = R;
#pragma unroll
                    for (int lvl = 0; lvl < 5; ++lvl) {
                        if (nm == 1) break;
                        int k = 0;
#pragma unroll
                        for (int i = 0; i < nm; i += 3) {
                            if (i + 2 < nm) mx[k++] = pk::max3u15(mx[i], mx[i + 1], mx[i + 2]);
                            else if (i + 1 < nm) mx[k++] = pk::max(mx[i], mx[i + 1]);
                            else mx[k++] = mx[i];
                        }
                        nm = k;
                    }
                    cm = mx[0];
                    const int nbst = pk::max(best, cm);
                    if (nbst != best) {
                        // the half whose maximum rose stores its rows in LDS
                        // (per-thread slots, rows 256 slots apart)
                        const int d = nbst ^ best;
                        if (d & 0xffff) {
#pragma unroll
                            for (int i = 0; i < R; ++i) snp[i * 256] = (short)Hs[SW ? i : 0];
                            bcolA = j;
                        }
                        if (d >> 16) {
#pragma unroll
                            for (int i = 0; i < R; ++i)
                                snp[(kSnapRows + i) * 256] = (short)(Hs[SW ? i : 0] >> 16);
                            bcolB = j;
                        }
                        best = nbst;
                    }
                } else {
                    const bool ha = j == lastA, hb = j == lastB;
                    if (ha || hb) {
                        const int mask = (ha ? 0xffff : 0) | (hb ? (int)0xffff0000 : 0);
#pragma unroll
                        for (int i = 0; i < R; ++i) snap[i] = (HL[i] & mask) | (snap[i] & ~mask);
                    }
                }
            } else {
                // a lane outside the band's columns reads the next column's
                // top rows, which it needs if that column is its first
#pragma unroll
                for (int k = 0; k < LA; ++k) issue(k, cn);
            }
        };
        // code pairs: p0 for even steps, p1 for odd ones; step sv works on
        // column sv (relative to the lane) and reads ahead into sv + 1
#pragma unroll
        for (int k = 0; k < LA; ++k) issue(k, ccl[0]);
        CodePair p0 = *(const CodePair *)(ccl + 0), p1 = *(const CodePair *)(ccl + 1);
        int2 b0 = bbs[0], b1 = bbs[1];
        const std::integral_constant<bool, true> fastc;
        const std::integral_constant<bool, false> slowc;
        // pairs of steps whose every lane is inside the band's columns: the
        // even s with s >= T - 1 and s + 1 < width
        const int lo = min(T, steps), hi = max(lo, width - 1);
        int s = 0;
        for (; s < lo; s += 2) {
            step(slowc, s, p0, b0);
            step(slowc, s + 1, p1, b1);
        }
        for (; s < hi; s += 2) {
            step(fastc, s, p0, b0);
            step(fastc, s + 1, p1, b1);
        }
        for (; s < steps; s += 2) {
            step(slowc, s, p0, b0);
            step(slowc, s + 1, p1, b1);
        }
        if (SW && band_live) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (J.pair[h] < 0) continue;
                const int S = h ? pk::hi(best) : pk::lo(best);
                const int col = h ? bcolB : bcolA;
                int row = -1;
#pragma unroll
                for (int i = R - 1; i >= 0; --i) {
                    const int g = first_row[h] + i;
                    const int v = (int)snp[(h * kSnapRows + i) * 256];
                    if (g >= 0 && g < J.m[h] && v == S) row = g;
                }
                // a maximum at or above 0x7C00 (past the overflow limit,
                // where the rows' maximum may differ from every row) still
                // reaches the reduction, which sends the pair to the int32 tier
                if (row < 0 && S >= 0x7C00) row = 0;
                if (row >= 0 && (S > cS[h] || (S == cS[h] && col < cC[h]))) {
                    cS[h] = S; cC[h] = col; cR[h] = row;
                }
            }
        }
        if (!SW && band_live) {
            // the last column: largest value over rows 0..m-2, first row
#pragma unroll
            for (int h = 0; h < 2; ++h) {
#pragma unroll
                for (int i = 0; i < R; ++i) {
                    const int g = first_row[h] + i;
                    const int v = (h ? pk::hi(snap[i]) : pk::lo(snap[i])) + open;
                    if (g >= 0 && g < J.m[h] - 1 && v > cmx[h]) { cmx[h] = v; crow[h] = g; }
                }
            }
        }
        // lane T-1's row is read by lane 0 of the same wave in the next band
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    if (live) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (J.pair[h] < 0) continue;
            if (SW) c.cand[J.cdoff[h] + t] = make_int4(cS[h], cC[h], cR[h], 0);
            else
                c.cand[J.cdoff[h] + t] = make_int4(cmx[h], crow[h], h ? pk::hi(vmn) : pk::lo(vmn),
                                                   h ? pk::hi(vmx) : pk::lo(vmx));
        }
    }
}

// One queue entry at its rows per lane.
template <bool SW, bool EXT>
__device__ __forceinline__ void wave_rows(const WaveCfg &c, const WaveDesc &wd, const char *tabb,
                                          int lane, short *snp)
{
    switch (wd.R) {
    case 4: wave_run<4, SW, EXT>(c, wd, tabb, lane, snp); break;
    case 6: wave_run<6, SW, EXT>(c, wd, tabb, lane, snp); break;
    case 8: wave_run<8, SW, EXT>(c, wd, tabb, lane, snp); break;
    case 10: wave_run<10, SW, EXT>(c, wd, tabb, lane, snp); break;
    case 12: wave_run<12, SW, EXT>(c, wd, tabb, lane, snp); break;
    case 14: wave_run<14, SW, EXT>(c, wd, tabb, lane, snp); break;
    case 16: wave_run<16, SW, EXT>(c, wd, tabb, lane, snp); break;
    case 18: wave_run<18, SW, EXT>(c, wd, tabb, lane, snp); break;
    default: wave_run<20, SW, EXT>(c, wd, tabb, lane, snp); break;
    }
}

}  // namespace

// Score table in LDS: for query code q and target code t the int16
// W(q,t)+open sits at byte  q*qstride + (t>>1)*128*(ncopy/32) + (t&1)*2 + c*4
// with c = lane % ncopy: with 32 copies every lane of a half-wave reads its
// own bank, so the random (q, t) lookups of a wave never conflict.
//
// SW end snapshots follow the table in LDS: 2 x kSnapRows rows of 256
// int16 slots, one per thread (half A rows, then half B rows).
template <bool SW, bool EXT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_wave(WaveCfg c)
{
    extern __shared__ int tab32[];
    for (int k = threadIdx.x; k < c.tabn; k += blockDim.x) tab32[k] = ((const int *)c.tab)[k];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const char *tabb = (const char *)tab32 + (lane % c.ncopy) * 4;
    short *const snp = (short *)(tab32 + ((c.tabn + 3) & ~3)) + threadIdx.x;
    // persistent: each wave takes the next entry off the queue (entries are
    // ordered by decreasing latency, so the tail is short)
    for (;;) {
        int wv = 0;
        if (lane == 0) wv = atomicAdd(c.next_wave, 1);
        wv = __builtin_amdgcn_readfirstlane(wv);
        if (wv >= c.nwaves) break;
        const WaveDesc wd = c.waves[wv];
        if (wd.prio) __builtin_amdgcn_s_setprio(2);
        if (EXT && wd.ext) {
            wave_rows<SW, true>(c, wd, tabb, lane, snp);
        } else {
            wave_rows<SW, false>(c, wd, tabb, lane, snp);
        }
        if (wd.prio) __builtin_amdgcn_s_setprio(0);
    }
}

// SW epilogue: reduce T lane candidates per pair (max score, then min column,
// then min row: src/sw.c:150-158).  ovf_limit: the int16 tier's largest safe
// score; pairs above it are appended to `redo` for the int32 tier.
__global__ void __launch_bounds__(256) k_sw_reduce(const int4 *cand, const PairRef *refs, int n,
                                                   int ovf_limit, phip_out_t *out, int32_t *redo,
                                                   int *nredo)
{
    // 16 lanes per pair; the order is total, so the lanes' partial picks
    // combine to the serial pick in any order
    const int k = blockIdx.x * 16 + (threadIdx.x >> 4);
    const int l = threadIdx.x & 15;
    int S = -1, col = 0, row = 0;
    if (k < n) {
        const PairRef r = refs[k];
        for (int t = l; t < r.T; t += 16) {
            const int4 v = cand[r.cdoff + t];
            if (v.x > S || (v.x == S && (v.y < col || (v.y == col && v.z < row)))) {
                S = v.x; col = v.y; row = v.z;
            }
        }
    }
    for (int d = 8; d >= 1; d >>= 1) {
        const int oS = __shfl_xor(S, d), oc = __shfl_xor(col, d), orow = __shfl_xor(row, d);
        if (oS > S || (oS == S && (oc < col || (oc == col && orow < row)))) {
            S = oS; col = oc; row = orow;
        }
    }
    if (k >= n || l) return;
    const int pair = refs[k].pair;
    phip_out_t o;
    memset(&o, 0, sizeof(o));
    o.score = S;
    o.end_query = row;
    o.end_ref = col;
    out[pair] = o;
    if (S > ovf_limit) redo[atomicAdd(nredo, 1)] = pair;
}

__global__ void k_zero(int *p, int n)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0;
}

// NW / SG epilogue, one wave per job (src/sg.c:171-235, src/nw.c:155-157).
// A job's halves have adjacent refs (half 0, then half 1 when present) and
// share the band row at `lastrow`, which holds the DP's last row as H-open
// for both: the wave of the half-0 ref settles both pairs, the half-1
// ref's wave has nothing to do.  The lane candidates hold each half's
// last-column maximum over rows 0..m-2 with its first row; the last row is
// scanned (its maximum, first column on ties) only when its end gap is
// free (SG with s2_end); otherwise the corner H[m-1][n-1] is all it needs.
__global__ void __launch_bounds__(256) k_sg_reduce(const int4 *cand, const int2 *bandbuf,
                                                   const PairRef *refs, int n, int alg, int flags,
                                                   int open, int gap, bool ext, phip_out_t *out)
{
    const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (k >= n) return;
    const PairRef r0 = refs[k];
    if (r0.half) return;
    const bool two = k + 1 < n && refs[k + 1].half;
    const PairRef r1 = two ? refs[k + 1] : r0;
    const bool s1e = alg == PHIP_SG && (flags & PHIP_S1_END);
    const bool s2e = alg == PHIP_SG && (flags & PHIP_S2_END);
    // last-column candidates, per half
    int cm0 = PHIP_NEG_INF32, ci0 = 0, cm1 = PHIP_NEG_INF32, ci1 = 0;
    // ext: the lanes' smallest E/F and largest H, per half (0 where no lane)
    int lo0 = 0, hi0 = 0, lo1 = 0, hi1 = 0;
    if ((s1e || ext) && lane < r0.T) {
        const int4 v = cand[r0.cdoff + lane];
        cm0 = s1e ? v.x : PHIP_NEG_INF32;
        ci0 = v.y;
        lo0 = v.z;
        hi0 = v.w;
        if (two) {
            const int4 w = cand[r1.cdoff + lane];
            cm1 = s1e ? w.x : PHIP_NEG_INF32;
            ci1 = w.y;
            lo1 = w.z;
            hi1 = w.w;
        }
    }
    if (ext) {
        for (int d = 32; d >= 1; d >>= 1) {
            lo0 = imin(lo0, __shfl_xor(lo0, d));
            hi0 = imax(hi0, __shfl_xor(hi0, d));
            lo1 = imin(lo1, __shfl_xor(lo1, d));
            hi1 = imax(hi1, __shfl_xor(hi1, d));
        }
    }
    // last row: one pass over the shared band row serves both halves
    const int2 *row = bandbuf + r0.lastrow;
    int rm0 = PHIP_NEG_INF32, rj0 = 0, rm1 = PHIP_NEG_INF32, rj1 = 0;
    if (s2e) {
        const int nmax = two ? imax(r0.n, r1.n) : r0.n;
        for (int j = lane; j < nmax; j += 64) {
            const int x = row[j].x;
            const int v0 = pk::lo(x) + open, v1 = pk::hi(x) + open;
            if (j < r0.n && v0 > rm0) { rm0 = v0; rj0 = j; }
            if (two && j < r1.n && v1 > rm1) { rm1 = v1; rj1 = j; }
        }
    }
    if (s1e || s2e) {
        for (int d = 32; d >= 1; d >>= 1) {
            int ov = __shfl_xor(cm0, d), oi = __shfl_xor(ci0, d);
            if (ov > cm0 || (ov == cm0 && oi < ci0)) { cm0 = ov; ci0 = oi; }
            ov = __shfl_xor(cm1, d); oi = __shfl_xor(ci1, d);
            if (ov > cm1 || (ov == cm1 && oi < ci1)) { cm1 = ov; ci1 = oi; }
            ov = __shfl_xor(rm0, d); oi = __shfl_xor(rj0, d);
            if (ov > rm0 || (ov == rm0 && oi < rj0)) { rm0 = ov; rj0 = oi; }
            ov = __shfl_xor(rm1, d); oi = __shfl_xor(rj1, d);
            if (ov > rm1 || (ov == rm1 && oi < rj1)) { rm1 = ov; rj1 = oi; }
        }
    }
    if (lane > 1 || (lane == 1 && !two)) return;
    // lane 0 settles half 0, lane 1 half 1
    const PairRef &r = lane ? r1 : r0;
    const int cm = lane ? cm1 : cm0, ci = lane ? ci1 : ci0;
    const int rm = lane ? rm1 : rm0, rj = lane ? rj1 : rj0;
    const int x = row[r.n - 1].x;
    const int corner = (lane ? pk::hi(x) : pk::lo(x)) + open;
    const int m = r.m, nn = r.n;
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
    if (ext) {
        // the cells' extrema with the DP's boundary row and column
        // (-open - (len-1) gap at their ends unless that gap is free)
        int vmin = lane ? lo1 : lo0;
        if (!(alg == PHIP_SG && (flags & PHIP_S2_BEG))) vmin = imin(vmin, -open - (nn - 1) * gap);
        if (!(alg == PHIP_SG && (flags & PHIP_S1_BEG))) vmin = imin(vmin, -open - (m - 1) * gap);
        o.min_val = vmin;
        o.max_val = lane ? hi1 : hi0;
    }
    out[r.pair] = o;
}

// Column-code stream: per job, the byte offsets of the two target residues'
// columns in the LDS table, for columns -T .. len-T-1 (pads outside the
// pairs' columns).
__global__ void k_colcode(const WaveJob *jobs, const int *clen, const int *cT, int njobs,
                          const uint8_t *t, const int64_t *toff, int PADT, int ncopy,
                          uint2 *colcode)
{
    const int jb = blockIdx.x;
    if (jb >= njobs) return;
    const WaveJob J = jobs[jb];
    const int L = clen[jb], T = cT[jb];
    const uint8_t *ta = J.pair[0] >= 0 ? t + toff[J.pair[0]] : nullptr;
    const uint8_t *tb = J.pair[1] >= 0 ? t + toff[J.pair[1]] : nullptr;
    for (int k = threadIdx.x; k < L; k += blockDim.x) {
        const int j = k - T;
        uint32_t a = (ta && j >= 0 && j < J.n[0]) ? ta[j] : PADT;
        uint32_t bb = (tb && j >= 0 && j < J.n[1]) ? tb[j] : PADT;
        // the table holds two columns' entries per dword (even column in
        // the low half); the selector takes bytes 0-1 or 2-3 of A's dword
        // into the low half and of B's into the high half
        const uint32_t oa = (a >> 1) * ncopy * 4, ob = (bb >> 1) * ncopy * 4;
        const uint32_t sel = 0x05040100u + (a & 1) * 0x0202u + (bb & 1) * 0x02020000u;
        colcode[J.ccoff + k] = make_uint2(oa | (ob << 16), sel);
    }
}

// ------------------------------------------------------------------ host

namespace {

const int kRs[] = {4, 6, 8, 10, 12, 14, 16, 18, kSnapRows};
constexpr int kRmax = 20;

struct Shape {
    int R, nb;
    double cost;  // time of the job's wave (cost of a step times steps)
};

// Cost of one step of a wave with R rows per lane: measured on the device
// (phip_wave_autotune: ns per wave-step, [1] SW, [0] NW/SG) or, untuned,
// the instruction model -- ~13 VALU instructions per row and ~30 per step,
// 4 cycles each.  Only ratios matter to the planner.
double g_step_cost[2][kRmax + 1];
bool g_tuned = false;

double step_cost(bool sw, int R)
{
    if (g_tuned && g_step_cost[sw][R] > 0) return g_step_cost[sw][R];
    return (13.0 * R + 30.0) * 4.0;
}

double job_cycles(bool sw, int T, int R, int nb, int n)
{
    return (double)nb * (n + T - 1) * step_cost(sw, R);
}

// Band height T*R and band count for a query of m rows against n columns:
// the cheapest (rows computed x steps) cover of the query.  The columns
// scale every candidate's cost alike (n + T - 1 steps a band), so the
// choice depends on m alone; choose_m memoises it per query length.
Shape choose(bool sw, int T, int m, int n)
{
    Shape best{kRmax, (m + T * kRmax - 1) / (T * kRmax), 0};
    best.cost = job_cycles(sw, T, best.R, best.nb, n);
    const int maxnb = std::max(1, (m + T * 4 - 1) / (T * 4));
    for (int nb = 1; nb <= maxnb && nb <= 512; ++nb) {
        for (int R : kRs) {
            if ((int64_t)nb * T * R < m) continue;
            const double cost = job_cycles(sw, T, R, nb, n);
            if (cost < best.cost) best = Shape{R, nb, cost};
            break;
        }
    }
    return best;
}

// choose() for T = 16 through a table indexed by m (filled on first use);
// the cost for n columns is nb * (n + T - 1) * step_cost(R)
struct ShapeTable {
    bool sw;
    std::vector<Shape> by_m;
    Shape get(int m, int n)
    {
        if ((size_t)m >= by_m.size()) by_m.resize((size_t)m + 1, Shape{0, 0, 0});
        Shape &s = by_m[m];
        if (s.nb == 0) s = choose(sw, 16, m, 1);
        Shape r = s;
        r.cost = job_cycles(sw, 16, r.R, r.nb, n);
        return r;
    }
};

typedef void (*KFn)(WaveCfg);

// Whether every DP value of an m x n pair provably stays inside int16
// (NW/SG), or can be checked at run time (SW, from the running maximum).
bool int16_fits(const phip_matrix_t &mx, int alg, int open, int gap, int m, int n)
{
    if (mx.is_pssm) return false;
    if (mx.max + open > 16000 || mx.min < -16000 || open > 8000 || gap > 8000) return false;
    if (m > 30000 || n > 30000) return false;
    if (alg == PHIP_SW) return true;
    // the virtual rows above a query repeat row 0 (H -open - (j-1) gap) by
    // horizontal gaps from H = 0, which reopen more cheaply than they
    // extend when gap > open
    if (gap > open) return false;
    int64_t lo = -(int64_t)4 * open - (int64_t)(m + n) * gap - std::abs(mx.min) - 64;
    int64_t hi = (int64_t)std::min(m, n) * std::max(mx.max, 0) + mx.max + open + 64;
    return lo > -32000 && hi < 32000;
}

// NW/SG with a saturation check: whether the pair's window at dp.check is
// decided without the cells' extrema -- it provably holds, or the DP's
// boundary row or column alone leaves it and the next window (16 bits;
// after 16, 32) provably holds.  The reduction reports the boundary's
// extrema, and the cells' only when some pair of the plan is not settled.
bool window_settled(const phip_matrix_t &mx, const DPParams &dp, int m, int n)
{
    if (dp32_window_holds(mx, dp.alg, PHIP_SCORE, dp.check, dp.open, dp.gap, m, n)) return true;
    const int64_t wmin = -((int64_t)1 << (dp.check - 1));
    const int64_t neg = (-dp.open < mx.min ? wmin + dp.open : wmin - (int64_t)mx.min) + 1;
    const bool s1_beg = dp.alg == PHIP_SG && (dp.sg_flags & PHIP_S1_BEG);
    const bool s2_beg = dp.alg == PHIP_SG && (dp.sg_flags & PHIP_S2_BEG);
    const int64_t top = s2_beg ? 0 : -(int64_t)dp.open - (int64_t)(n - 1) * dp.gap;
    const int64_t left = s1_beg ? 0 : -(int64_t)dp.open - (int64_t)(m - 1) * dp.gap;
    return std::min(top, left) < neg &&
           (dp.check >= 16 || dp32_window_holds(mx, dp.alg, PHIP_SCORE, 16, dp.open, dp.gap, m, n));
}

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

int device_simds()
{
    static int simds = 0;
    if (!simds) {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0)
            cus = 256;
        simds = 4 * cus;
    }
    return simds;
}

}  // namespace

struct WavePlan {
    int alg, sg_flags, open, gap;
    int check;    // saturation window whose extrema the results carry (0: none)
    bool ext;     // NW/SG: some pair's window needs the cells' extrema
    bool sw;
    int nwaves = 0, nblocks = 0;
    int nfast = 0, nslow = 0;
    int VIRT = 0, PADQ = 0, PADT = 0, ovf_limit = 0;
    int tabn = 0, qstride = 0, ncopy = 1;
    int maxn = 0, simple_lanes = 0;
    char *base = nullptr;
    WaveJob *djobs = nullptr;
    WaveDesc *dwaves = nullptr;
    short *dtab = nullptr;
    uint2 *dcc = nullptr;
    int2 *dbb = nullptr, *dtop = nullptr;
    int4 *dcd = nullptr;
    PairRef *drefs = nullptr;
    int32_t *dslow = nullptr, *dredo = nullptr;
    int *dcount = nullptr, *dscratch = nullptr;
    LongPlan lp;                 // long global pairs (nwlong.hip)
    char *dlong = nullptr;
};

std::vector<int> radix_order(const std::vector<uint32_t> &keys)
{
    const size_t N = keys.size();
    std::vector<uint32_t> k(keys), k2(N);
    std::vector<int> ix(N), ix2(N);
    for (size_t i = 0; i < N; ++i) ix[i] = (int)i;
    std::vector<uint32_t> cnt(65537);
    for (int shift = 0; shift < 32; shift += 16) {
        std::fill(cnt.begin(), cnt.end(), 0u);
        for (size_t i = 0; i < N; ++i) ++cnt[((k[i] >> shift) & 0xFFFF) + 1];
        for (int d = 0; d < 65536; ++d) cnt[d + 1] += cnt[d];
        for (size_t i = 0; i < N; ++i) {
            const size_t pos = cnt[(k[i] >> shift) & 0xFFFF]++;
            k2[pos] = k[i];
            ix2[pos] = ix[i];
        }
        k.swap(k2);
        ix.swap(ix2);
    }
    return ix;
}

bool wave_plan_matches(const WavePlan *p, int alg, int sg_flags, int open, int gap, int check)
{
    return p->alg == alg && p->sg_flags == sg_flags && p->open == open && p->gap == gap &&
           p->check == check;
}

bool wave_covers(const HostBatch &hb, const DPParams &dp)
{
    const phip_matrix_t &mx = *hb.mx;
    const int tsp0 = (mx.size + 2) & ~1;
    if ((size_t)(mx.size + 2) * (tsp0 / 2) * 4 > 65535u) return false;
    for (int i = 0; i < hb.np; ++i)
        if (!int16_fits(mx, dp.alg, dp.open, dp.gap, hb.qlen[i], hb.tlen[i])) return false;
    return true;
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
    P->check = dp.check;
    P->ext = false;
    P->sw = dp.alg == PHIP_SW;
    const bool sw = P->sw;
    const phip_matrix_t &mx = *hb.mx;
    const int np = hb.np;
    std::vector<int> fast, slow;
    std::vector<Shape> shape(np);
    std::vector<int> pT(np, 16);
    std::vector<char> unsettled(np, 0);
    // LDS table geometry (below); its byte offsets must fit the 16-bit
    // halves of the packed row offsets, or every pair takes the int32 tier
    const int tsp0 = (mx.size + 2) & ~1;
    const bool table_ok = (size_t)(mx.size + 2) * (tsp0 / 2) * 4 <= 65535u;
    double total = 0;
    std::vector<int> lng;
    ShapeTable st{sw, {}};
    for (int i = 0; i < np; ++i) {
        P->maxn = std::max(P->maxn, (int)hb.tlen[i]);
        if (table_ok && int16_fits(mx, dp.alg, dp.open, dp.gap, hb.qlen[i], hb.tlen[i])) {
            fast.push_back(i);
            if (!sw && dp.check && !window_settled(mx, dp, hb.qlen[i], hb.tlen[i])) unsettled[i] = 1;
            shape[i] = st.get(hb.qlen[i], hb.tlen[i]);
            total += shape[i].cost;
        } else if (nwlong_eligible(mx, dp.alg, dp.kind, dp.check, hb.qlen[i], hb.tlen[i])) {
            lng.push_back(i);
        } else {
            slow.push_back(i);
        }
    }
    // machine time of the whole pass at 16 lanes per job (4 jobs, 8 pairs
    // share a wave's instructions; ~2 waves interleave per SIMD): a job whose
    // own latency is a large share of it runs on a whole wave instead
    const double pass = total / 8.0 / device_simds();
    for (int i : fast) {
        if (shape[i].cost > 0.25 * pass && hb.qlen[i] > 16 * 8) {
            const Shape s64 = choose(sw, 64, hb.qlen[i], hb.tlen[i]);
            shape[i] = s64;
            pT[i] = 64;
        }
    }
    // class (T, R), then longest first: T 64 before 16, R ascending, bands
    // and then columns descending (nb <= 512, R <= 20, n <= 30000 for a pair
    // int16_fits admits)
    {
        std::vector<uint32_t> key(fast.size());
        for (size_t x = 0; x < fast.size(); ++x) {
            const int i = fast[x];
            key[x] = (uint32_t)(pT[i] == 64 ? 0 : 1) << 30 | (uint32_t)shape[i].R << 25 |
                     (uint32_t)(1023 - shape[i].nb) << 15 | (uint32_t)(32767 - hb.tlen[i]);
        }
        const std::vector<int> ord = radix_order(key);
        std::vector<int> sorted(fast.size());
        for (size_t x = 0; x < fast.size(); ++x) sorted[x] = fast[ord[x]];
        fast.swap(sorted);
    }
    // LDS table: rows [0,size) real, VIRT, PADQ; columns [0,size) real, PADT
    const int size = mx.size;
    const int TS = size + 1;
    P->VIRT = size;
    P->PADQ = size + 1;
    P->PADT = size;
    const int tabrows = size + 2;
    // bank-replicated layout (32 copies) when it fits comfortably in LDS
    const int tsp = (TS + 1) & ~1;
    P->ncopy = ((size_t)tabrows * (tsp / 2) * 32 * 4 <= (48u << 10)) ? 32 : 1;
    P->qstride = (tsp / 2) * P->ncopy * 4;
    P->tabn = tabrows * (tsp / 2) * P->ncopy;
    std::vector<short> htab((size_t)P->tabn * 2, (short)W_NEG);
    const bool s2_beg = dp.alg == PHIP_SG && (dp.sg_flags & PHIP_S2_BEG);
    const bool zero_top = sw || s2_beg;
    if (!mx.is_pssm) {
        for (int r = 0; r < tabrows; ++r)
            for (int col = 0; col < tsp; ++col) {
                int v = W_NEG;
                if (col >= P->PADT || r == P->PADQ) v = W_NEG;
                else if (r == P->VIRT) v = zero_top ? dp.open : W_NEG;
                else v = mx.values[(size_t)r * size + col] + (sw ? 0 : dp.open);
                for (int cp = 0; cp < P->ncopy; ++cp) {
                    size_t byte = (size_t)r * P->qstride + (size_t)(col >> 1) * P->ncopy * 4 +
                                  (col & 1) * 2 + cp * 4;
                    htab[byte / 2] = (short)v;
                }
            }
    }
    // SW: every H of a pair below the limit stays below 0x7C00 (the rows'
    // maximum is taken on binary16 patterns); NW/SG: below int16's bound
    P->ovf_limit = (sw ? 0x7BFF : 32767) - (mx.max + dp.open + 1);
    // jobs: consecutive pairs of the same class and band count share one
    // job (A low, B high); then entries of 64/T jobs
    std::vector<WaveJob> jobs;
    std::vector<int> jT, jR;
    std::vector<char> jext;
    std::vector<double> jcost;
    jobs.reserve(fast.size());
    jT.reserve(fast.size());
    jR.reserve(fast.size());
    jext.reserve(fast.size());
    jcost.reserve(fast.size());
    size_t k = 0;
    while (k < fast.size()) {
        WaveJob J;
        memset(&J, 0, sizeof(J));
        const int p0 = fast[k];
        int halves = 1;
        if (k + 1 < fast.size()) {
            const int p1 = fast[k + 1];
            if (pT[p1] == pT[p0] && shape[p1].R == shape[p0].R && shape[p1].nb == shape[p0].nb)
                halves = 2;
        }
        J.nb = shape[p0].nb;
        J.pair[1] = -1;
        J.cdoff[1] = -1;
        double cost = 0;
        char ext = 0;
        for (int h = 0; h < halves; ++h) {
            const int p = fast[k + h];
            ext |= unsettled[p];
            J.pair[h] = p;
            J.m[h] = hb.qlen[p];
            J.n[h] = hb.tlen[p];
            J.s[h] = sw ? 0 : J.nb * pT[p0] * shape[p0].R - J.m[h];
            J.qoff[h] = hb.qoff[p];
            cost = std::max(cost, shape[p].cost);
        }
        jobs.push_back(J);
        jT.push_back(pT[p0]);
        jR.push_back(shape[p0].R);
        jext.push_back(ext);
        jcost.push_back(cost);
        k += halves;
    }
    std::vector<WaveDesc> waves;
    std::vector<double> wcost;
    waves.reserve(jobs.size());
    wcost.reserve(jobs.size());
    {
        size_t j0 = 0;
        while (j0 < jobs.size()) {
            const int T = jT[j0], R = jR[j0];
            WaveDesc w;
            memset(&w, 0, sizeof(w));
            w.job0 = (int)j0;
            w.T = T;
            w.R = R;
            double cst = 0;
            size_t j1 = j0;
            // one band count per wave: the wave runs its most bands over its
            // widest columns, so a many-band job beside a wide one would
            // multiply the two (one such wave once took twice a pass)
            while (j1 < jobs.size() && (int)(j1 - j0) < 64 / T && jT[j1] == T && jR[j1] == R &&
                   jobs[j1].nb == jobs[j0].nb) {
                w.nb = std::max(w.nb, jobs[j1].nb);
                w.width = std::max(w.width, std::max(jobs[j1].n[0], jobs[j1].n[1]));
                cst = std::max(cst, jcost[j1]);
                w.ext |= jext[j1];
                ++j1;
            }
            P->ext = P->ext || w.ext;
            w.njobs = (int)(j1 - j0);
            w.bbw = (w.width + T + 8 + 15) & ~15;
            w.prio = T == 64;
            waves.push_back(w);
            wcost.push_back(cst);
            j0 = j1;
        }
    }
    // queue order: decreasing latency
    std::vector<int> order(waves.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return wcost[x] > wcost[y]; });
    std::vector<WaveDesc> qw;
    std::vector<WaveJob> qj;
    std::vector<int> clen, cT;
    std::vector<PairRef> refs;
    qw.reserve(waves.size());
    qj.reserve(jobs.size());
    clen.reserve(jobs.size());
    cT.reserve(jobs.size());
    refs.reserve(fast.size());
    int64_t ccoff = 0, bboff = 0, cdoff = 0;
    int maxw = 0;
    for (int wi : order) {
        WaveDesc w = waves[wi];
        const int j0 = w.job0;
        w.job0 = (int)qj.size();
        maxw = std::max(maxw, w.width);
        for (int q = 0; q < w.njobs; ++q) {
            WaveJob J = jobs[j0 + q];
            J.ccoff = ccoff;
            const int L = (w.width + 2 * w.T + 8 + 15) & ~15;
            ccoff += L;
            clen.push_back(L);
            cT.push_back(w.T);
            J.bboff = bboff;
            const int rows = sw ? J.nb - 1 : J.nb;
            bboff += (int64_t)rows * w.bbw;
            for (int h = 0; h < 2; ++h) {
                if (J.pair[h] < 0) continue;
                J.cdoff[h] = cdoff;
                PairRef r;
                memset(&r, 0, sizeof(r));
                r.cdoff = cdoff;
                r.lastrow = sw ? 0 : J.bboff + (int64_t)(J.nb - 1) * w.bbw;
                r.pair = J.pair[h];
                r.m = J.m[h];
                r.n = J.n[h];
                r.T = w.T;
                r.half = h;
                refs.push_back(r);
                cdoff += w.T;
            }
            qj.push_back(J);
        }
        qw.push_back(w);
    }
    P->nwaves = (int)qw.size();
    P->nfast = (int)refs.size();
    P->nslow = (int)slow.size();
    if (!nwlong_plan(hb, lng, dp.open, P->lp)) { delete P; return nullptr; }
    P->nblocks = std::min((P->nwaves + 3) / 4, 1024);
    // DP row 0 as H-open per column (both halves), F = -inf
    const int ntop = maxw + 64 + 16;
    std::vector<int2> top(ntop);
    for (int j = 0; j < ntop; ++j) {
        // SW: H = 0 and F clamped at 0; NW/SG: H - open and F = -inf
        const int v = sw ? 0 : zero_top ? -dp.open : clamp16(-dp.open - j * dp.gap - dp.open);
        top[j] = make_int2(pk::pack(v, v), sw ? 0 : PK_NEG);
    }
    // int32-tier scratch: rows of up to maxn+1 columns per lane
    const size_t per_lane = (size_t)(P->maxn + 1) * 2 * sizeof(int);
    P->simple_lanes = (int)std::max<size_t>(64, std::min<size_t>(8192, ((size_t)512 << 20) / per_lane));
    // sizes and one allocation
    const size_t nj = qj.size();
    size_t off = 0;
    auto slot = [&](size_t bytes) { size_t o = off; off += al256(bytes ? bytes : 1); return o; };
    size_t o_jobs = slot(nj * sizeof(WaveJob)), o_waves = slot(qw.size() * sizeof(WaveDesc)),
           o_tab = slot(htab.size() * 2), o_cc = slot((size_t)ccoff * 8),
           o_clen = slot(nj * 4), o_cT = slot(nj * 4), o_bb = slot((size_t)bboff * 8),
           o_top = slot((size_t)ntop * 8),
           o_cd = slot((size_t)cdoff * 16), o_refs = slot(refs.size() * sizeof(PairRef)),
           o_slow = slot(slow.size() * 4), o_redo = slot(refs.size() * 4), o_count = slot(64),
           o_scr = slot(per_lane * P->simple_lanes), o_long = slot(P->lp.bytes);
    P->base = (char *)dev_alloc(off);
    if (!P->base) { delete P; return nullptr; }
    char *b = P->base;
    P->djobs = (WaveJob *)(b + o_jobs);
    P->dwaves = (WaveDesc *)(b + o_waves);
    P->dtab = (short *)(b + o_tab);
    P->dcc = (uint2 *)(b + o_cc);
    P->dbb = (int2 *)(b + o_bb);
    P->dtop = (int2 *)(b + o_top);
    P->dcd = (int4 *)(b + o_cd);
    P->drefs = (PairRef *)(b + o_refs);
    P->dslow = (int32_t *)(b + o_slow);
    P->dredo = (int32_t *)(b + o_redo);
    P->dcount = (int *)(b + o_count);
    P->dscratch = (int *)(b + o_scr);
    P->dlong = b + o_long;
    int *dclen = (int *)(b + o_clen), *dcT = (int *)(b + o_cT);
    auto up = [&](void *dst, const void *src, size_t bytes) -> int {
        if (!bytes) return 0;
        HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
        return 0;
    };
    int rc = 0;
    rc |= up(P->djobs, qj.data(), nj * sizeof(WaveJob));
    rc |= up(P->dwaves, qw.data(), qw.size() * sizeof(WaveDesc));
    rc |= up(P->dtab, htab.data(), htab.size() * 2);
    rc |= up(dclen, clen.data(), nj * 4);
    rc |= up(dcT, cT.data(), nj * 4);
    rc |= up(P->dtop, top.data(), (size_t)ntop * 8);
    rc |= up(P->drefs, refs.data(), refs.size() * sizeof(PairRef));
    rc |= up(P->dslow, slow.data(), slow.size() * 4);
    rc |= up(P->dlong, P->lp.items.data(), P->lp.items.size());
    if (nj) {
        hipLaunchKernelGGL(k_colcode, dim3((unsigned)nj), dim3(256), 0, stream, P->djobs, dclen,
                           dcT, (int)nj, dp.t, dp.toff, P->PADT, P->ncopy, P->dcc);
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
hipEvent_t g_kw0, g_kw1;
bool g_kw_created = false, g_kw_recorded = false;
}  // namespace

double wave_last_kernel_ms()
{
    float ms = 0.f;
    if (!g_kw_recorded || hipEventElapsedTime(&ms, g_kw0, g_kw1) != hipSuccess) return 0.0;
    return ms;
}

int wave_plan_run(WavePlan *P, const DPParams &dp, hipStream_t stream, hipEvent_t e0, hipEvent_t e1)
{
    if (!g_kw_created) {
        HIP_OK(hipEventCreate(&g_kw0));
        HIP_OK(hipEventCreate(&g_kw1));
        g_kw_created = true;
    }
    g_kw_recorded = false;
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, stream, P->dcount, 16);
    if (P->nwaves) {
        WaveCfg cfg;
        memset(&cfg, 0, sizeof(cfg));
        cfg.jobs = P->djobs;
        cfg.waves = P->dwaves;
        cfg.nwaves = P->nwaves;
        cfg.next_wave = P->dcount + 1;
        cfg.colcode = P->dcc;
        cfg.tab = P->dtab;
        cfg.tabn = P->tabn;
        cfg.qstride = P->qstride;
        cfg.ncopy = P->ncopy;
        cfg.VIRT = P->VIRT;
        cfg.PADQ = P->PADQ;
        cfg.q = dp.q;
        cfg.bandbuf = P->dbb;
        cfg.toprow = P->dtop;
        cfg.cand = P->dcd;
        cfg.open = P->open;
        cfg.gap = P->gap;
        cfg.alg = P->alg;
        cfg.sg_flags = P->sg_flags;
        KFn fn = P->sw ? k_wave<true, false> : P->ext ? k_wave<false, true> : k_wave<false, false>;
        HIP_OK(hipEventRecord(g_kw0, stream));
        if (e0) HIP_OK(hipEventRecordWithFlags(e0, stream, hipEventRecordExternal));
        const size_t lds = (size_t)((cfg.tabn + 3) & ~3) * 4 + (P->sw ? 2 * kSnapRows * 256 * 2 : 0);
        hipLaunchKernelGGL(fn, dim3(P->nblocks), dim3(256), lds, stream, cfg);
        HIP_OK(hipGetLastError());
        HIP_OK(hipEventRecord(g_kw1, stream));
        if (e1) HIP_OK(hipEventRecordWithFlags(e1, stream, hipEventRecordExternal));
        g_kw_recorded = true;
        if (P->sw) {
            hipLaunchKernelGGL(k_sw_reduce, dim3((P->nfast + 15) / 16), dim3(256), 0, stream,
                               P->dcd, P->drefs, P->nfast, P->ovf_limit, dp.out, P->dredo,
                               P->dcount);
        } else {
            hipLaunchKernelGGL(k_sg_reduce, dim3((P->nfast + 3) / 4), dim3(256), 0, stream, P->dcd,
                               P->dbb, P->drefs, P->nfast, P->alg, P->sg_flags, P->open, P->gap,
                               P->check != 0, dp.out);
        }
        HIP_OK(hipGetLastError());
    }
    if (P->lp.nitems && nwlong_launch(dp, P->lp, P->dlong, stream) != 0) return -1;
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

// ------------------------------------------------------------ autotuning
//
// The planner's cost of a step with R rows per lane, measured: for each R a
// batch whose every pair plans to one band of R rows per lane on 16-lane
// chains (query 16 R rows, target 384 columns) and holds 8 waves per SIMD
// runs on the wavefront kernel; its event-timed duration per wave-step is
// R's cost.  SW and NW (SG runs NW's kernel) are measured apart.  Pairs
// are random over the matrix's first 20 letters.  Calibration batches use
// the instruction model, so that their shapes are the ones intended.
extern "C" int phip_wave_autotune(const phip_matrix_t *mx, int open, int gap)
{
    if (phip_init() != 0) return -1;
    if (!mx || mx->is_pssm || mx->size < 2 || open < 0 || gap < 0) {
        phip_set_error("phip_wave_autotune: needs a square matrix and gaps >= 0");
        return -1;
    }
    const int n = 384, T = 16;
    const int waves = 8 * device_simds();
    const int np = 8 * waves;  // 4 jobs of 2 pairs a wave
    const int letters = std::min(20, mx->size - 1);
    double cost[2][kRmax + 1] = {};
    const bool was = g_tuned;
    g_tuned = false;
    uint64_t x = 0x9E3779B97F4A7C15ull;
    auto rnd = [&]() -> uint8_t {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        return (uint8_t)(x % (uint64_t)letters);
    };
    std::vector<uint8_t> q, t((size_t)np * n);
    for (uint8_t &c : t) c = rnd();
    std::vector<int64_t> qoff(np), toff(np);
    std::vector<int32_t> qlen(np), tlen(np, n);
    std::vector<phip_out_t> out(np);
    int rc = 0;
    for (int sw = 0; sw < 2 && rc == 0; ++sw) {
        for (int R : kRs) {
            const int m = T * R;
            q.resize((size_t)np * m);
            for (uint8_t &c : q) c = rnd();
            for (int i = 0; i < np; ++i) {
                qoff[i] = (int64_t)i * m;
                toff[i] = (int64_t)i * n;
                qlen[i] = m;
            }
            phip_batch_t b;
            memset(&b, 0, sizeof(b));
            b.alg = sw ? PHIP_SW : PHIP_NW;
            b.kind = PHIP_SCORE;
            b.open = open;
            b.gap = gap;
            b.matrix = *mx;
            b.npairs = np;
            b.q = q.data();
            b.qoff = qoff.data();
            b.qlen = qlen.data();
            b.t = t.data();
            b.toff = toff.data();
            b.tlen = tlen.data();
            double best = 0;
            for (int rep = 0; rep < 2 && rc == 0; ++rep) {
                rc = phip_align(&b, out.data(), nullptr);
                const double ms = phip_last_main_kernel_ms();
                if (rep == 0 || ms < best) best = ms;
            }
            if (rc) break;
            cost[sw][R] = best * 1e6 / ((double)waves * (n + T - 1));
        }
    }
    if (rc) {
        g_tuned = was;
        return -1;
    }
    memcpy(g_step_cost, cost, sizeof(cost));
    g_tuned = true;
    return 0;
}

// The measured costs (ns per wave-step, device-wide) per rows-per-lane
// choice, SW and NW/SG; the number of choices, 0 if untuned.
extern "C" int phip_wave_autotune_table(int *rows, double *sw_cost, double *nw_cost, int cap)
{
    int k = 0;
    for (int R : kRs) {
        if (k < cap) {
            if (rows) rows[k] = R;
            if (sw_cost) sw_cost[k] = g_tuned ? g_step_cost[1][R] : 0;
            if (nw_cost) nw_cost[k] = g_tuned ? g_step_cost[0][R] : 0;
        }
        ++k;
    }
    return g_tuned ? k : 0;
}

// Back to the instruction model.
extern "C" void phip_wave_autotune_reset(void) { g_tuned = false; }
