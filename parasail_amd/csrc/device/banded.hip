// banded.hip -- banded global alignment (parasail_nw_banded,
// src/nw_banded.c:29-189) for many pairs, one wave per pair.
//
// The reference keeps the band in diagonal coordinates: slot i of column j
// is row colOff+i (colOff grows by one per column), H/E of slot i at column
// j come from slots i (diagonal) and i+1 (left) of column j-1 and slot i-1
// of column j (up); slots whose row lies outside the query keep whatever
// they held, and the two edge slots stay at -inf.  Here the slots of a
// column are split over the 64 lanes of a wave (R consecutive slots per
// lane, in registers) and the lanes sweep the columns as a wavefront: lane
// t works on column j = s - t at step s.  Up comes from lane t-1's last slot
// of the same column (computed one step earlier, one DPP wave_shr:1); left
// for the lane's last slot comes from lane t+1's first slot of the previous
// column, which lane t+1 computes in the same step -- the lanes run that
// slot first and hand it over with one DPP wave_shl:1 before the rest.
// Every array value, including the stale and edge ones, is the reference's,
// so the score read at the end is too.
#include <string.h>

#include <algorithm>
#include <vector>

#include "search.h"

#define BNEG (INT32_MIN / 2)  // NEG_INF_32, src/nw_banded.c:16

namespace {

__device__ __forceinline__ int shr1(int v) { return __builtin_amdgcn_mov_dpp(v, 0x138, 0xF, 0xF, false); }
__device__ __forceinline__ int shl1(int v) { return __builtin_amdgcn_mov_dpp(v, 0x130, 0xF, 0xF, false); }
// v of lane t-1; lane 0 gets old
__device__ __forceinline__ int shr1_or(int old, int v)
{
    return __builtin_amdgcn_update_dpp(old, v, 0x138, 0xF, 0xF, false);
}
// v of lane t+1; lane 63 gets old
__device__ __forceinline__ int shl1_or(int old, int v)
{
    return __builtin_amdgcn_update_dpp(old, v, 0x130, 0xF, 0xF, false);
}

}  // namespace

struct BandCfg {
    const uint8_t *q;
    const int64_t *qoff;
    const int32_t *qlen;
    const uint8_t *t;
    const int64_t *toff;
    const int32_t *tlen;
    const int *mat;        // device, row-major (square: size x size; PSSM: m x size)
    int msize, is_pssm;
    int open, gap, k;
    int lring;             // bottom-row slots per wave (k_banded_rows)
    const int32_t *pairs;  // pairs of this launch
    const int32_t *rsel;   // slots per lane of each (k_banded_chain)
    int npairs;
    phip_out_t *out;
};

// Slots per lane: R (colLen <= 64 R).  LTAB: the square matrix is staged
// in LDS; otherwise substitution values come from the matrix or PSSM in
// global memory (one select of the row index, no branch).
//
// Two kinds of step.  Near the ends of the pair some lanes are outside the
// target, or some slots hold rows outside the query; there every slot is
// updated only if valid, as in the reference.  In between (every lane on a
// column of the target, every band slot on a row of the query), all slots
// are updated unconditionally and the slots past the band's end are capped
// at -inf; their values only ever reach the band as the left neighbour of
// its last slot, where any value <= -inf/2 loses the max to a real path just
// as the reference's -inf does.
//
// Several waves of one workgroup may share a pair: lane g = 64 w + t of an
// L = 64 NWV lane chain.  Inside a wave the exchanges above are DPP moves;
// across waves they go through LDS: lane 0 of wave w publishes its first
// slot (barrier A, mid-step) for lane 63 of wave w-1, and lane 63 of wave w
// its bottom row (barrier B, end of step) for lane 0 of wave w+1.
template <int R, bool LTAB>
__device__ __forceinline__ void band_chain(const BandCfg &c, const int *tab, const int pr,
                                           const int w, const int t, const int NWV, int2 *lx,
                                           int2 *ux)
{
    const int m = c.qlen[pr], n = c.tlen[pr];
    const uint8_t *s1 = c.q + c.qoff[pr];
    const uint8_t *s2 = c.t + c.toff[pr];
    const int open = c.open, gap = c.gap, k = c.k, msize = c.msize;
    const int low = (m > n) ? n - m - k : m - n - k;
    const int colLen = k - low + 1;
    const int colOff0 = (m > n) ? -k : low;
    const int L = 64 * NWV, g = 64 * w + t;  // lanes of the chain; this lane
    const int tR = g * R;
    const int *mrow = c.mat;

    // the reference's initial arrays (src/nw_banded.c:110-138); qm holds
    // the query code of each slot's row times the matrix width
    int Hb[R], Eb[R], qm[R], cap[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = tR + r;
        int h = 0, e = 0;
        if (i >= -colOff0 + 1) { h = -open - (i - (-colOff0 + 1)) * gap; e = BNEG; }
        if (i == -colOff0 - 1) { h = -open; e = BNEG; }
        if (i == -colOff0) { h = 0; e = BNEG; }
        if (i >= colLen) { h = BNEG; e = BNEG; }
        Hb[r] = h;
        Eb[r] = e;
        cap[r] = i < colLen ? INT32_MAX : BNEG;
        // the lane's codes shift one row per step from step 0 on, when its
        // column is -g: start from that column's rows
        const int pos = colOff0 - g + i;
        qm[r] = (int)s1[pos < 0 ? 0 : (pos >= m ? m - 1 : pos)] * msize;
    }
    // clamped, unconditional loads: the residue of a row outside the query
    // is never used (its slot is skipped)
    auto ldq = [&](int pos) -> int { return (int)s1[pos < 0 ? 0 : (pos >= m ? m - 1 : pos)]; };
    auto ldt = [&](int j) -> int { return (int)s2[j < 0 ? 0 : (j >= n ? n - 1 : j)]; };
    int botH = BNEG, botF = BNEG;
    // codes fetched two steps ahead: the target residue of the lane's column
    // and the query residue entering its last slot at the next column
    int tq0 = ldt(-g), tq1 = ldt(1 - g);
    int nq0 = ldq(colOff0 - g + tR + R), nq1 = ldq(colOff0 + 1 - g + tR + R);

    auto step = [&](auto fastc, const int s, int &tq, int &nq) __attribute__((always_inline)) {
        constexpr bool FAST = decltype(fastc)::value;
        const int j = s - g;
        const bool act = j >= 0 && j < n;
        const int xH = shr1(botH), xF = shr1(botF);
        // slot -1 of lane 0 of the chain is the -inf edge; lane 0 of a later
        // wave takes the bottom row the wave above published last step
        int Hup = xH, F = xF;
        if (t == 0) {
            Hup = BNEG;
            F = BNEG;
            if (w > 0) {
                const int2 u = ux[w - 1];
                Hup = u.x;
                F = u.y;
            }
        }
        const int colOff = colOff0 + j;
        if (!FAST && colOff0 + s - (L - 1) < 0) {
            // some lane's column still has rows above the query: the slot
            // just above row 0 holds that column's top boundary
            const int ib = -colOff - 1 - tR;
            const int bval = -open - j * gap;
            const bool bst = act && colOff < 0;
#pragma unroll
            for (int r = 0; r < R; ++r) Hb[r] = (bst && r == ib) ? bval : Hb[r];
        }
        const int tc = tq;
        int W[R];
        bool val[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int pos = colOff + tR + r;
            val[r] = act & (pos >= 0) & (pos < m) & (tR + r < colLen);
            if (LTAB) {
                W[r] = tab[qm[r] + tc];
            } else {
                const int pc = pos < 0 ? 0 : (pos >= m ? m - 1 : pos);
                W[r] = mrow[(size_t)(c.is_pssm ? pc * msize : qm[r]) + tc];
            }
        }
        auto slot = [&](int r, int Hn, int En) __attribute__((always_inline)) {
            const int Hdag = Hb[r] + W[r];
            const int E = max(Hn - open, En - gap);
            const int Fn = max(Hup - open, F - gap);
            const int H = max(max(E, Fn), Hdag);
            if (FAST) {
                Eb[r] = E;
                F = Fn;
                Hb[r] = min(H, cap[r]);
            } else {
                const bool v = val[r];
                Eb[r] = v ? E : Eb[r];
                F = v ? Fn : F;
                Hb[r] = v ? H : Hb[r];
            }
            Hup = Hb[r];
        };
        slot(0, Hb[1], Eb[1]);
        if (NWV > 1) {
            if (t == 0 && w > 0) lx[w] = make_int2(Hb[0], Eb[0]);
            lds_barrier();  // A
        }
        // lane g+1's first slot, just computed for column j-1, is the left
        // neighbour of this lane's last slot
        int xHn = shl1(Hb[0]), xEn = shl1(Eb[0]);
        if (t == 63) {
            xHn = BNEG;
            xEn = BNEG;
            if (w + 1 < NWV) {
                const int2 v = lx[w + 1];
                xHn = v.x;
                xEn = v.y;
            }
        }
#pragma unroll
        for (int r = 1; r < R; ++r) slot(r, r + 1 < R ? Hb[r + 1] : xHn, r + 1 < R ? Eb[r + 1] : xEn);
        botH = Hup;
        botF = F;
        if (NWV > 1) {
            if (t == 63 && w + 1 < NWV) ux[w] = make_int2(botH, botF);
            lds_barrier();  // B
        }
        // next column: every slot moves one row down the query
#pragma unroll
        for (int r = 0; r < R - 1; ++r) qm[r] = qm[r + 1];
        qm[R - 1] = nq * msize;
        tq = ldt(j + 2);
        nq = ldq(colOff + 2 + tR + R);
    };
    const std::integral_constant<bool, false> slow;
    const std::integral_constant<bool, true> fast;
    // steps where every lane is on a column of the target and every band
    // slot on a row of the query, rounded inwards to even steps
    const int tl = (colLen - 1) / R;
    const int sa = (max(L - 1, -colOff0) + 1) & ~1;
    const int sb = max(sa, min(n, m + tl + 1 - colOff0 - colLen) & ~1);
    const int steps = n + L - 1;
    int s = 0;
    for (; s < sa; s += 2) {
        step(slow, s, tq0, nq0);
        step(slow, s + 1, tq1, nq1);
    }
    for (; s < sb; s += 2) {
        step(fast, s, tq0, nq0);
        step(fast, s + 1, tq1, nq1);
    }
    for (; s < steps; s += 2) {
        step(slow, s, tq0, nq0);
        step(slow, s + 1, tq1, nq1);
    }
    const int sidx = (m > n) ? -low : k;
    int score = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) score = (tR + r == sidx) ? Hb[r] : score;
    if (sidx >= tR && sidx < tR + R) {
        phip_out_t o;
        memset(&o, 0, sizeof(o));
        o.score = score;
        o.end_query = m - 1;
        o.end_ref = n - 1;
        o.status = PHIP_ST_W32;
        c.out[pr] = o;
    }
}

// One workgroup per pair (blockDim/64 waves, one lane chain); c.rsel gives
// each pair's slots per lane (at most RMAX).  LDS: the square matrix
// (LTAB), then the two 8-entry exchange rows of band_chain.
template <int RMAX, bool LTAB>
__global__ void __launch_bounds__(512) k_banded_chain(BandCfg c)
{
    extern __shared__ int tab[];
    if (LTAB)
        for (int x = threadIdx.x; x < c.msize * c.msize; x += blockDim.x) tab[x] = c.mat[x];
    int2 *const lx = (int2 *)(tab + (LTAB ? (c.msize * c.msize + 1) & ~1 : 0));
    int2 *const ux = lx + 8;
    __syncthreads();
    if (blockIdx.x >= (unsigned)c.npairs) return;  // whole block
    const int NWV = blockDim.x >> 6, t = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int pr = c.pairs[blockIdx.x];
#define BAND_R(RR) band_chain<(RMAX >= RR ? RR : 2), LTAB>(c, tab, pr, w, t, NWV, lx, ux)
    switch (c.rsel[blockIdx.x]) {
    case 2: BAND_R(2); break;
    case 3: BAND_R(3); break;
    case 4: BAND_R(4); break;
    case 5: if (RMAX >= 5) BAND_R(5); break;
    case 6: if (RMAX >= 6) BAND_R(6); break;
    case 8: if (RMAX >= 8) BAND_R(8); break;
    case 10: if (RMAX >= 10) BAND_R(10); break;
    case 12: if (RMAX >= 12) BAND_R(12); break;
    case 16: if (RMAX >= 16) BAND_R(16); break;
    case 20: if (RMAX >= 20) BAND_R(20); break;
    case 24: if (RMAX >= 24) BAND_R(24); break;
    case 32: if (RMAX >= 32) BAND_R(32); break;
    default: break;
    }
#undef BAND_R
}

// ------------------------------------------------------------ segments
//
// Long pairs in the diagonal coordinates above, cut along the target into
// segments that run at the same time, one 64-lane chain each -- the method
// of Maleki, Musuvathi and Mytkowicz, "Parallelizing dynamic programming
// through rank convergence" (PPoPP 2014).  Between two columns the band's
// state (H and E of every slot; F restarts at -inf in every column) maps to
// the next column's by a max-plus map: every new value is a max of old
// values plus constants, so adding d to every real value of a column adds d
// to every real value after it, and the -inf slots stay -inf.  A segment
// other than the first starts X columns before its own first column, from
// 0 in every band slot; once it has gone far enough its state differs from
// the true one by a single constant in every real slot.  The join checks
// that exactly, at the column before each segment's own first column: the
// previous segment's state there and this segment's must differ by one
// constant d in every real slot, with -inf (<= -inf/2) in the same slots.
// Then this segment's values are the true ones minus the sum of the d's up
// to it, and the pair's score is the last segment's plus that sum.  A pair
// whose check fails anywhere is aligned again without segments.
//
// A segment's run covers steps [c0, c0 + len): lane t works on column
// s - t, as in band_chain.  At its first step the lanes' columns lie before
// c0; from the arbitrary start they just compute more of the warm-up.
// Every lane stores its slots' (H, E) at the column cs (the state the join
// compares with the previous segment's end) and at the column ce (the
// state the next segment is compared with).
struct SegDesc {
    int32_t pair, R;
    int32_t c0;    // first step; 0 for a pair's first segment (the reference's start)
    int32_t len;   // steps of the run (even)
    int32_t cs;    // column of the stored start state, -1 for the first segment
    int32_t ce;    // column of the stored end state, -1 for the last segment
    int64_t voff;  // int2 offset of the states: start at voff, end at voff + 64 R
};

struct SegPair {
    int32_t pair, first, count, colLen;
};

template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}
// f(integral_constant<int, 0>), ..., f(integral_constant<int, N - 1>)
template <int N, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

__device__ __forceinline__ int med3(int a, int b, int c) { return max(min(a, b), min(max(a, b), c)); }

// The square matrix is staged in LDS (tab).  RP: tab also holds, after the
// matrix, each query letter's profile -- W + open of every target letter in
// 4-bit fields -- and the steps where every band slot is a row of the query
// and every lane a column of the target run R at a time with it:
//  * H is kept as H - open, so E = max(Hl - open, El - gap) and the same for
//    F take two instructions each, and the diagonal term is H - open plus
//    the profile's W + open;
//  * a slot's row moves one down the query per step, so slot r at step
//    s + u of a block of R steps reads its profile from qp[(r + u) % R]: the
//    register slot 0 has just used takes the profile of the row entering
//    slot R-1 at the next step, and no register is copied;
//  * the slots past the band's end are held at -inf by H = med3(max(E, F),
//    H_diag, cap) with cap = -inf there (their H_diag stays within the
//    host's score bound of -inf, so the median stays <= -inf/2) and
//    cap = INT32_MAX elsewhere, where the median is the max.
template <int R, bool RP>
__device__ __forceinline__ void seg_chain(const BandCfg &c, const int *tab, const SegDesc &d, int2 *vec,
                                          int *raw, const int si, const int t)
{
    const int pr = d.pair;
    const int m = c.qlen[pr], n = c.tlen[pr];
    const uint8_t *s1 = c.q + c.qoff[pr];
    const uint8_t *s2 = c.t + c.toff[pr];
    const int open = c.open, gap = c.gap, k = c.k, msize = c.msize;
    const int low = (m > n) ? n - m - k : m - n - k;
    const int colLen = k - low + 1;
    const int colOff0 = (m > n) ? -k : low;
    const int tR = t * R;
    const int s0 = d.c0;
    const bool first = d.cs < 0;

    int Hb[R], Eb[R], qm[R], cap[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = tR + r;
        int h = 0, e = BNEG;
        if (first) {
            // the reference's initial arrays (src/nw_banded.c:110-138)
            e = 0;
            if (i >= -colOff0 + 1) { h = -open - (i - (-colOff0 + 1)) * gap; e = BNEG; }
            if (i == -colOff0 - 1) { h = -open; e = BNEG; }
            if (i == -colOff0) { h = 0; e = BNEG; }
        }
        if (i >= colLen) { h = BNEG; e = BNEG; }
        Hb[r] = h;
        Eb[r] = e;
        cap[r] = i < colLen ? INT32_MAX : BNEG;
        const int pos = colOff0 + s0 - t + i;  // the slot's row at the lane's first column
        qm[r] = (int)s1[pos < 0 ? 0 : (pos >= m ? m - 1 : pos)] * msize;
    }
    auto ldq = [&](int pos) -> int { return (int)s1[pos < 0 ? 0 : (pos >= m ? m - 1 : pos)]; };
    auto ldt = [&](int j) -> int { return (int)s2[j < 0 ? 0 : (j >= n ? n - 1 : j)]; };
    int botH = BNEG, botF = BNEG;
    int tq0 = ldt(s0 - t), tq1 = ldt(s0 + 1 - t);
    int nq0 = ldq(colOff0 + s0 - t + tR + R), nq1 = ldq(colOff0 + s0 + 1 - t + tR + R);
    int2 *const vst = vec + d.voff + tR, *const ven = vec + d.voff + 64 * R + tR;

    auto step = [&](auto fastc, const int s, const bool chk, int &tq, int &nq) __attribute__((always_inline)) {
        constexpr bool FAST = decltype(fastc)::value;
        const int j = s - t;
        const bool act = j >= 0 && j < n;
        // slot -1 of lane 0 is the band's -inf edge
        int Hup = shr1_or(BNEG, botH), F = shr1_or(BNEG, botF);
        const int colOff = colOff0 + j;
        if (!FAST && first && colOff0 + s - 63 < 0) {
            // the slot just above row 0 holds the column's top boundary
            const int ib = -colOff - 1 - tR;
            const int bval = -open - j * gap;
            const bool bst = act && colOff < 0;
#pragma unroll
            for (int r = 0; r < R; ++r) Hb[r] = (bst && r == ib) ? bval : Hb[r];
        }
        const int tc = tq;
        int W[R];
        bool val[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int pos = colOff + tR + r;
            val[r] = act & (pos >= 0) & (pos < m) & (tR + r < colLen);
            W[r] = tab[qm[r] + tc];
        }
        auto slot = [&](int r, int Hn, int En) __attribute__((always_inline)) {
            const int Hdag = Hb[r] + W[r];
            const int E = max(Hn - open, En - gap);
            const int Fn = max(Hup - open, F - gap);
            if (FAST) {
                Eb[r] = E;
                F = Fn;
                Hb[r] = med3(max(E, Fn), Hdag, cap[r]);
            } else {
                const bool v = val[r];
                Eb[r] = v ? E : Eb[r];
                F = v ? Fn : F;
                Hb[r] = v ? max(max(E, Fn), Hdag) : Hb[r];
            }
            Hup = Hb[r];
        };
        slot(0, Hb[1], Eb[1]);
        // lane t+1's first slot, just computed for column j-1, is the left
        // neighbour of this lane's last slot; lane 63's is the -inf edge
        const int xHn = shl1_or(BNEG, Hb[0]), xEn = shl1_or(BNEG, Eb[0]);
#pragma unroll
        for (int r = 1; r < R; ++r) slot(r, r + 1 < R ? Hb[r + 1] : xHn, r + 1 < R ? Eb[r + 1] : xEn);
        botH = Hup;
        botF = F;
        if (chk) {
            if (j == d.cs) {
#pragma unroll
                for (int r = 0; r < R; ++r) vst[r] = make_int2(Hb[r], Eb[r]);
            }
            if (j == d.ce) {
#pragma unroll
                for (int r = 0; r < R; ++r) ven[r] = make_int2(Hb[r], Eb[r]);
            }
        }
#pragma unroll
        for (int r = 0; r < R - 1; ++r) qm[r] = qm[r + 1];
        qm[R - 1] = nq * msize;
        tq = ldt(j + 2);
        nq = ldq(colOff + 2 + tR + R);
    };
    const std::integral_constant<bool, false> slow;
    const std::integral_constant<bool, true> fast;
    // steps whose every lane is on a column of the target and every band
    // slot on a row of the query (as in band_chain)
    const int tl = (colLen - 1) / R;
    const int sa = (max(63, -colOff0) + 1) & ~1;
    const int sb = max(sa, min(n, m + tl + 1 - colOff0 - colLen) & ~1);
    const int s1e = s0 + d.len;
    // the stores fall in the 64 steps from cs and from ce
    auto chk_at = [&](int s) -> bool {
        return (unsigned)(s + 1 - d.cs) < 65u || (unsigned)(s + 1 - d.ce) < 65u;
    };
    int s = s0;
    for (const int e = min(sa, s1e); s < e; s += 2) {
        const bool chk = chk_at(s);
        step(slow, s, chk, tq0, nq0);
        step(slow, s + 1, chk, tq1, nq1);
    }
    if (RP) {
        // blocks of R fast steps; the rest of the fast steps run as slow
        // steps below, which compute the same values there
        const int e = s + (min(sb, s1e) - s) / R * R;
        if (s < e) {
            const int *prof = tab + ((msize * msize + 3) & ~3);
            int qp[R];
            // H - open from here on, the bottom row handed down included
#pragma unroll
            for (int r = 0; r < R; ++r) {
                Hb[r] -= open;
                qp[r] = prof[ldq(colOff0 + s - t + tR + r)];
            }
            botH -= open;
            // the row entering slot R-1 after step s' is colOff0 + s' - t +
            // tR + R; its letter and the step's target letter are loaded
            // two steps ahead, unclamped: in these steps the lanes' columns
            // are at most n and the pad slots' rows at most m + 64 R, and
            // both residue arrays end with zeroed slack
            const uint8_t *const tcol = s2 - t;
            const uint8_t *const qrow = s1 + (colOff0 - t + tR + R);
            int ta = tcol[s], tb = tcol[s + 1];
            int na = qrow[s], nb = qrow[s + 1];
            // the DPP results: lane 0 (up) and lane 63 (left) keep -inf
            int xH = BNEG, xF = BNEG, xHn = BNEG, xEn = BNEG;
            for (; s < e; s += R) {
                const uint8_t *const tblk = tcol + s, *const qblk = qrow + s;
                static_for<R>([&](auto uc) __attribute__((always_inline)) {
                    constexpr int U = decltype(uc)::value;
                    const int sp = s + U;
                    int &tq = (U & 1) ? tb : ta;
                    int &nq = (U & 1) ? nb : na;
                    const int tc4 = tq << 2;
                    xH = shr1_or(xH, botH);
                    xF = shr1_or(xF, botF);
                    int Hup = xH, F = xF;
                    auto slot = [&](int r, int HLn, int En) __attribute__((always_inline)) {
                        const int Wp = (int)__builtin_amdgcn_ubfe((unsigned)qp[(r + U) % R], tc4, 4);
                        const int Hdag = Hb[r] + Wp;
                        const int E = max(HLn, En - gap);
                        F = max(Hup, F - gap);
                        Eb[r] = E;
                        Hb[r] = med3(max(E, F), Hdag, cap[r]) - open;
                        Hup = Hb[r];
                    };
                    slot(0, Hb[1], Eb[1]);
                    qp[U % R] = prof[nq];
                    xHn = shl1_or(xHn, Hb[0]);
                    xEn = shl1_or(xEn, Eb[0]);
#pragma unroll
                    for (int r = 1; r < R; ++r)
                        slot(r, r + 1 < R ? Hb[r + 1] : xHn, r + 1 < R ? Eb[r + 1] : xEn);
                    botH = Hup;
                    botF = F;
                    if (chk_at(sp & ~1)) {
                        const int j = sp - t;
                        if (j == d.cs) {
#pragma unroll
                            for (int r = 0; r < R; ++r) vst[r] = make_int2(Hb[r] + open, Eb[r]);
                        }
                        if (j == d.ce) {
#pragma unroll
                            for (int r = 0; r < R; ++r) ven[r] = make_int2(Hb[r] + open, Eb[r]);
                        }
                    }
                    tq = tblk[U + 2];
                    nq = qblk[U + 2];
                });
            }
            // back to the slow step's state at step s
#pragma unroll
            for (int r = 0; r < R; ++r) {
                Hb[r] += open;
                qm[r] = ldq(colOff0 + s - t + tR + r) * msize;
            }
            botH += open;
            tq0 = ldt(s - t);
            tq1 = ldt(s + 1 - t);
            nq0 = ldq(colOff0 + s - t + tR + R);
            nq1 = ldq(colOff0 + s + 1 - t + tR + R);
        }
    } else {
        for (const int e = min(sb, s1e); s < e; s += 2) {
            const bool chk = chk_at(s);
            step(fast, s, chk, tq0, nq0);
            step(fast, s + 1, chk, tq1, nq1);
        }
    }
    for (; s < s1e; s += 2) {
        const bool chk = chk_at(s);
        step(slow, s, chk, tq0, nq0);
        step(slow, s + 1, chk, tq1, nq1);
    }
    if (d.ce < 0) {
        // the last segment: the score as band_chain reads it, relative to
        // this segment's start
        const int sidx = (m > n) ? -low : k;
        int score = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) score = (tR + r == sidx) ? Hb[r] : score;
        if (sidx >= tR && sidx < tR + R) raw[si] = score;
    }
}

// One wave per segment, 4 segments per workgroup; slots per lane from the
// descriptor (at most RMAX).  LDS: the square matrix, then (RP) each query
// letter's 4-bit profile.
template <int RMAX, bool RP>
__global__ void __launch_bounds__(256) k_banded_seg(BandCfg c, const SegDesc *segs, const int32_t *order,
                                                    int nseg, int2 *vec, int *raw)
{
    extern __shared__ int tab[];
    const int ms = c.msize;
    for (int x = threadIdx.x; x < ms * ms; x += blockDim.x) tab[x] = c.mat[x];
    if (RP) {
        int *prof = tab + ((ms * ms + 3) & ~3);
        for (int x = threadIdx.x; x < ms; x += blockDim.x) {
            int pf = 0;
            for (int y = 0; y < ms; ++y) pf |= (c.mat[x * ms + y] + c.open) << (4 * y);
            prof[x] = pf;
        }
    }
    __syncthreads();
    const int li = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (li >= nseg) return;
    const int t = threadIdx.x & 63;
    const int si = __builtin_amdgcn_readfirstlane(order[li]);
    const SegDesc d = segs[si];
#define SEG_R(RR) seg_chain<(RMAX >= RR ? RR : 8), RP>(c, tab, d, vec, raw, si, t)
    switch (__builtin_amdgcn_readfirstlane(d.R)) {
    case 8: SEG_R(8); break;
    case 10: if (RMAX >= 10) SEG_R(10); break;
    case 12: if (RMAX >= 12) SEG_R(12); break;
    case 14: if (RMAX >= 14) SEG_R(14); break;
    case 16: if (RMAX >= 16) SEG_R(16); break;
    case 20: if (RMAX >= 20) SEG_R(20); break;
    case 24: if (RMAX >= 24) SEG_R(24); break;
    case 32: if (RMAX >= 32) SEG_R(32); break;
    default: break;
    }
#undef SEG_R
}

// One wave per segmented pair: the join described above.  A pair whose
// states do not line up gets status PHIP_ST_ERR (aligned again by the host).
__global__ void __launch_bounds__(64) k_seg_join(BandCfg c, const SegPair *sp, const SegDesc *segs,
                                                 const int2 *vec, const int *raw)
{
    const SegPair p = sp[blockIdx.x];
    const int t = threadIdx.x;
    const int thr = BNEG / 2;
    int sum = 0;
    bool ok = true;
    for (int q = 1; q < p.count && ok; ++q) {
        const SegDesc a = segs[p.first + q - 1], b = segs[p.first + q];
        const int2 *ea = vec + a.voff + 64 * a.R;
        const int2 *sb = vec + b.voff;
        int lo = INT32_MAX, hi = INT32_MIN, bad = 0;
        for (int x = t; x < p.colLen; x += 64) {
            const int2 u = ea[x], v = sb[x];
            const bool uh = u.x > thr, vh = v.x > thr, ue = u.y > thr, ve = v.y > thr;
            bad |= (uh != vh) | (ue != ve);
            if (uh && vh) { lo = min(lo, u.x - v.x); hi = max(hi, u.x - v.x); }
            if (ue && ve) { lo = min(lo, u.y - v.y); hi = max(hi, u.y - v.y); }
        }
        for (int o = 32; o > 0; o >>= 1) {
            lo = min(lo, __shfl_xor(lo, o));
            hi = max(hi, __shfl_xor(hi, o));
            bad |= __shfl_xor(bad, o);
        }
        ok = !bad && lo == hi && lo != INT32_MAX;
        sum += lo;
    }
    if (t == 0) {
        const int m = c.qlen[p.pair], n = c.tlen[p.pair];
        phip_out_t o;
        memset(&o, 0, sizeof(o));
        o.score = ok ? raw[p.first + p.count - 1] + sum : 0;
        o.end_query = m - 1;
        o.end_ref = n - 1;
        o.status = ok ? PHIP_ST_W32 : PHIP_ST_ERR;
        c.out[p.pair] = o;
    }
}

// Single-lane fallback for bands wider than 64 x 32 slots (one pair per
// block): the reference loop over a band kept in HBM scratch.
__global__ void __launch_bounds__(64) k_banded_one(BandCfg c, int *scratch, int64_t stride)
{
    if (threadIdx.x != 0 || blockIdx.x >= (unsigned)c.npairs) return;
    const int pr = c.pairs[blockIdx.x];
    const int m = c.qlen[pr], n = c.tlen[pr];
    const uint8_t *s1 = c.q + c.qoff[pr];
    const uint8_t *s2 = c.t + c.toff[pr];
    const int open = c.open, gap = c.gap, k = c.k;
    const int low = (m > n) ? n - m - k : m - n - k;
    const int colLen = k - low + 1;
    int colOff = (m > n) ? -k : low;
    int *H = scratch + (int64_t)blockIdx.x * 2 * stride + 1;
    int *E = H + stride;
    for (int i = -1; i <= colLen; ++i) { H[i] = 0; E[i] = 0; }
    for (int i = -colOff + 1, j = 0; i < colLen; ++i, ++j) { H[i] = -open - j * gap; E[i] = BNEG; }
    H[-colOff - 1] = -open;
    E[-colOff - 1] = BNEG;
    H[-colOff] = 0;
    E[-colOff] = BNEG;
    H[-1] = BNEG;
    E[-1] = BNEG;
    H[colLen] = BNEG;
    E[colLen] = BNEG;
    for (int j = 0; j < n; ++j) {
        int F = BNEG;
        if (colOff < 0) H[-colOff - 1] = -open - j * gap;
        const int lo = max(0, -colOff), hi = min(colLen, m - colOff);
        for (int i = lo; i < hi; ++i) {
            const int pos = colOff + i;
            const int matval = c.is_pssm ? c.mat[(size_t)c.msize * pos + s2[j]]
                                         : c.mat[(size_t)c.msize * s1[pos] + s2[j]];
            const int H_dag = H[i] + matval;
            const int E_opn = H[i + 1] - open, E_ext = E[i + 1] - gap;
            const int F_opn = H[i - 1] - open, F_ext = F - gap;
            E[i] = max(E_opn, E_ext);
            F = max(F_opn, F_ext);
            H[i] = max(max(E[i], F), H_dag);
        }
        colOff += 1;
    }
    phip_out_t o;
    memset(&o, 0, sizeof(o));
    o.score = (m > n) ? H[-low] : H[k];
    o.end_query = m - 1;
    o.end_ref = n - 1;
    o.status = PHIP_ST_W32;
    c.out[pr] = o;
}

// ------------------------------------------------------------ row bands
//
// Long pairs: the same band in (row, column) coordinates.  Row r of the
// query lies in the band for the columns [r + c0, r + c0 + colLen - 1],
// c0 = 1 - colOff0 - colLen, and the reference's band (stale slots and -inf
// edges included) is the global recurrence of src/nw_banded.c:143-168 over
// those cells with every cell outside the band at -inf, the top row
// H(-1, j) = -open - j*gap where row -1 lies in the band, H(-1, -1) = 0, and
// the left column H(r, -1) = -open - r*gap for the rows in the band at
// column -1.  Here every dependency points down or right, so the query's
// rows are cut into bands of 64 R rows, lane t of a wave holds R rows of
// one band and computes column a + s - t at step s (a: the band's first
// column), the row above coming from lane t-1 by DPP and, for lane 0, from
// the band above through LDS.  A band's lanes meet their rows' colLen band
// columns in turn (R + 1 steps apart), so a band takes about
// 64 R + colLen + 64 steps; cells outside the band are masked to -inf.
//
// The NWV waves of a workgroup take the bands of one pair in turn: band b
// runs on wave b % NWV, starting G >= RB_L steps (on the workgroup's common
// step count) after band b-1 reached the same column, and the waves meet at
// a barrier every RB_K steps.  Wave w's lane 63 stores its bottom row at step
// s in slot s of the wave's LDS row (two steps per store); lane 0 of the
// next band loads, two steps ahead, what was stored at least
// G - 64 R - 68 >= RB_K steps earlier, so every value it loads was stored
// before the last barrier.  A wave's next band starts NWV G steps
// after its last one, at least a band's steps plus RB_K: each wave holds
// one band per stretch between barriers, and overwrites a slot long after
// the next band loaded it.
constexpr int RB_K = 32;          // steps between barriers
constexpr int RB_L = RB_K + 68;   // lag of a band behind the band above
constexpr int RNEG = -(1 << 29);  // -inf: the host bounds every score above it
static_assert(RB_L >= RB_K + 67, "band lag too short for the barrier spacing");

// steps between consecutive row bands of 64 R = br rows taking send steps
// each, on nwv waves (a multiple of 4, like every band's steps: the step
// loop runs four steps a turn)
__host__ __device__ __forceinline__ int rows_spacing(int br, int send, int nwv)
{
    const int q = ((send + RB_K + nwv - 1) / nwv + 3) & ~3;
    return br + RB_L > q ? br + RB_L : q;
}

__device__ __forceinline__ int rol1(int v) { return __builtin_amdgcn_mov_dpp(v, 0x134, 0xF, 0xF, false); }
template <int R, bool RP>
__device__ __forceinline__ void band_rows(const BandCfg &c, const int *tab, const int pr, const int w,
                                          const int t, const int NWV, int2 *rings)
{
    constexpr int BR = 64 * R;
    const int m = c.qlen[pr], n = c.tlen[pr];
    const uint8_t *s1 = c.q + c.qoff[pr];
    const uint8_t *s2 = c.t + c.toff[pr];
    const int open = c.open, gap = c.gap, k = c.k, msize = c.msize;
    const int low = (m > n) ? n - m - k : m - n - k;
    const int colLen = k - low + 1;
    const int colOff0 = (m > n) ? -k : low;
    const int c0 = 1 - colOff0 - colLen;  // first band column of row r: r + c0
    const int lb = colOff0 + colLen - 2;  // rows <= lb lie in the band at column -1
    const int nb = (m + BR - 1) / BR;
    // a band's steps: lane 63 passes its last band column, then stores
    // -inf for the R + 2 columns lane 0 of the next band still loads
    // (rounded up to a multiple of 4)
    const int send = (BR + colLen + R + 64 + 3) & ~3;
    int2 *const ring_out = rings + w * c.lring;
    // slot of the column lane 0 computes at step s: BR + s + 63
    const int2 *const ring_in = rings + (w == 0 ? NWV - 1 : w - 1) * c.lring + BR + 63;
    // steps between the starts of consecutive bands: at least the lag RB_L
    // behind the band above, and 1/NWV of a band's steps plus RB_K, so that
    // a wave is free again when its next band starts
    const int G = rows_spacing(BR, send, NWV);
    const int nph = ((nb - 1) * G + send + RB_K - 1) / RB_K;
    const int tp = t * (R + 1);

    int HL[R], E[R], qp[R];  // H - open and E of the lane's rows; their scores
    // fast bands: whether row 0 was in the band 1, 2, ... steps ago (row i
    // is in the band at step s iff row 0 was at step s - i)
    bool vh[R > 1 ? R - 1 : 1];
    int prevRH = RNEG, botH = RNEG, botF = RNEG, tq = 0, leftTop = RNEG, r0 = 0, a = 0;
    // cbuf: codes of the 64 columns from lane 0's, rotated one lane a step
    // (lane 0 holds lane 0's column); cnext: the next 64, as loaded
    int cbuf = 0, cnext = 0;
    int2 in0 = make_int2(RNEG, RNEG), in1 = in0;
    auto ldc = [&](int j) -> int { return (int)s2[j < 0 ? 0 : (j >= n ? n - 1 : j)]; };
    int b = w;
    // loaded one band ahead, so that a band starts without waiting on HBM:
    // the query codes of the lane's rows in the wave's next band and that
    // band's first 128 column codes
    int nqc[R], ncb0, ncb1;
    auto prefetch = [&](int bn) {
        const int rn = bn * BR + t * R;
#pragma unroll
        for (int i = 0; i < R; ++i) nqc[i] = (int)s1[rn + i < m ? rn + i : m - 1];
        const int an = bn * BR + c0;
        ncb0 = ldc(an + t);
        ncb1 = ldc(an + 64 + t);
    };
    prefetch(b);
    // lane 0's row above at step s for band 0: the top row of the DP
    auto top = [&](int s) -> int2 {
        const int j = a + s;
        return make_int2(j >= 0 && j <= -1 - colOff0 ? -open - j * gap - open : RNEG, RNEG);
    };
    auto ld_in = [&](int s) -> int2 { return b == 0 ? top(s) : ring_in[s]; };

    auto step = [&](auto slowc, const int s, const int2 in) __attribute__((always_inline)) {
        constexpr bool SLOW = decltype(slowc)::value;
        // lane 0 takes the new column's code and the row above from LDS;
        // every other lane its left neighbour's, one step old
        tq = shr1_or(cbuf, tq);
        cbuf = rol1(cbuf);
        const int xH = shr1_or(in.x, botH);
        const int xF = shr1_or(in.y, botF);
        int W[R];
#pragma unroll
        for (int i = 0; i < R; ++i) W[i] = RP ? (int)__builtin_amdgcn_ubfe((unsigned)qp[i], tq, 4) : tab[qp[i] + tq];
        const int p = s - tp;  // this lane's column minus its first row's first band column
        int t1 = prevRH + W[0];
        prevRH = xH;
        bool act = true;
        if (SLOW) {
            const int j = a + s - t;
            if (j == 0) t1 = leftTop + W[0];
            act = j >= 0 && j < n;
        }
        if (act) {
            int up = xH, F = xF;
            bool v = false;
            const bool v0 = (unsigned)p < (unsigned)colLen;
#pragma unroll
            for (int i = 0; i < R; ++i) {
                if (SLOW)
                    v = (unsigned)(p - i) < (unsigned)colLen && r0 + i < m;
                else
                    v = i == 0 ? v0 : vh[i - 1];  // row i now = row 0 i steps ago
                E[i] = max(E[i] - gap, HL[i]);
                const int t1n = i + 1 < R ? HL[i] + W[i + 1] : 0;
                F = max(F - gap, up);
                const int h = max(max(t1, E[i]), F);
                HL[i] = v ? h - open : RNEG;
                up = HL[i];
                t1 = t1n;
            }
            botH = HL[R - 1];
            botF = v ? F : RNEG;
            if (!SLOW && R > 1) {
#pragma unroll
                for (int k = R - 2; k > 0; --k) vh[k] = vh[k - 1];
                vh[0] = v0;
            }
        }
    };
    // lane 0's row above at step s + 2, loaded into the register of step s
    // once step s has used it
    auto ld_next = [&](auto slowc, const int s) -> int2 {
        constexpr bool SLOW = decltype(slowc)::value;
        return SLOW && b == 0 ? top(s + 2) : ring_in[s + 2];
    };
    auto run = [&](auto slowc, int s, int e) __attribute__((always_inline)) {
        // the step counters are the wave's own (scalar registers)
        s = __builtin_amdgcn_readfirstlane(s);
        e = __builtin_amdgcn_readfirstlane(e);
        while (s < e) {
            if ((s & 63) == 0 && s != 0) {
                cbuf = RP ? 4 * cnext : cnext;
                cnext = ldc(a + s + 64 + t);
            }
            const int ce = min(e, (s | 63) + 1);
            for (; s < ce; s += 4) {
#pragma unroll
                for (int u = 0; u < 4; u += 2) {
                    step(slowc, s + u, in0);
                    const int h0 = botH, f0 = botF;
                    in0 = ld_next(slowc, s + u);
                    // keep each load a whole step ahead of its use (the
                    // scheduler would otherwise pair the two loads late)
                    __builtin_amdgcn_sched_barrier(0);
                    step(slowc, s + u + 1, in1);
                    in1 = ld_next(slowc, s + u + 1);
                    if (t == 63) *(int4 *)(ring_out + s + u) = make_int4(h0, f0, botH, botF);
                }
            }
        }
    };
    const std::integral_constant<bool, false> fastc;
    const std::integral_constant<bool, true> slowc;

    for (int ph = 0; ph < nph; ++ph) {
        const int P = ph * RB_K;
        if (b < nb && b * G + send <= P) b += NWV;
        if (b < nb) {
            const int Tb = b * G;
            const int s = max(0, P - Tb), e = min(send, P + RB_K - Tb);
            if (s < e) {
                a = b * BR + c0;
                if (s == 0) {
                    // the band's rows at column -1: the left column or -inf
                    r0 = b * BR + t * R;
#pragma unroll
                    for (int i = 0; i < R; ++i) {
                        const int r = r0 + i, rc = r < m ? r : m - 1;
                        if (RP && c.is_pssm) {
                            // a PSSM row: its profile from HBM
                            const int *mr = c.mat + (size_t)rc * msize;
                            int pf = 0;
                            for (int x = 0; x < msize; ++x) pf |= (mr[x] + open) << (4 * x);
                            qp[i] = pf;
                        } else {
                            // the profile of the row's letter, or its table row
                            qp[i] = RP ? tab[nqc[i]] : nqc[i] * msize;
                        }
                        HL[i] = r < m && r <= lb ? -open - r * gap - open : RNEG;
                        E[i] = RNEG;
                    }
#pragma unroll
                    for (int k = 0; k < (R > 1 ? R - 1 : 1); ++k) vh[k] = false;  // steps < 0
                    const int rt = r0 - 1;
                    leftTop = rt == -1 ? -open : (rt <= lb ? -open - rt * gap - open : RNEG);
                    botH = botF = RNEG;
                    tq = 0;
                    in0 = ld_in(0);
                    in1 = ld_in(1);
                    prevRH = t == 0 ? ld_in(-1).x : RNEG;
                    cbuf = RP ? 4 * ncb0 : ncb0;
                    cnext = ncb1;
                    prefetch(b + NWV);
                }
                // the first bands (columns below 0, the top row, the left
                // column) and the last (rows past the query, the score)
                // check every cell; the others only the band's edges
                if (a >= 1 && b < nb - 1)
                    run(fastc, s, e);
                else
                    run(slowc, s, e);
                if (e == send && b == nb - 1) {
                    // lanes stop at column n-1: the score is H(m-1, n-1)
                    const int rr = m - 1 - b * BR;
                    if (t == rr / R) {
                        int h = 0;
#pragma unroll
                        for (int i = 0; i < R; ++i) h = i == rr % R ? HL[i] : h;
                        phip_out_t o;
                        memset(&o, 0, sizeof(o));
                        o.score = h + open;
                        o.end_query = m - 1;
                        o.end_ref = n - 1;
                        o.status = PHIP_ST_W32;
                        c.out[pr] = o;
                    }
                }
            }
        }
        lds_barrier();
    }
}

// One workgroup (NWV = blockDim/64 waves) per pair; c.rsel gives each
// pair's rows per lane.  LDS: the W + open table (!RP) or the register
// profile of each letter (RP: W + open of each target letter in 4-bit
// fields), then NWV rows of c.lring bottom-row slots, set to -inf first.
template <bool RP>
__global__ void __launch_bounds__(256) k_banded_rows(BandCfg c)
{
    extern __shared__ int tab[];
    const int tabn = ((RP ? c.msize : c.msize * c.msize) + 3) & ~3;
    if (!RP) {
        for (int x = threadIdx.x; x < c.msize * c.msize; x += blockDim.x) tab[x] = c.mat[x] + c.open;
    } else if (!c.is_pssm) {
        for (int x = threadIdx.x; x < c.msize; x += blockDim.x) {
            int pf = 0;
            for (int y = 0; y < c.msize; ++y) pf |= (c.mat[x * c.msize + y] + c.open) << (4 * y);
            tab[x] = pf;
        }
    }
    int2 *const rings = (int2 *)(tab + tabn);
    const int NWV = blockDim.x >> 6;
    for (int x = threadIdx.x; x < NWV * c.lring; x += blockDim.x) rings[x] = make_int2(RNEG, RNEG);
    __syncthreads();
    if (blockIdx.x >= (unsigned)c.npairs) return;  // whole block
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), t = threadIdx.x & 63;
    const int pr = c.pairs[blockIdx.x];
    switch (c.rsel[blockIdx.x]) {
    case 1: band_rows<1, RP>(c, tab, pr, w, t, NWV, rings); break;
    case 2: band_rows<2, RP>(c, tab, pr, w, t, NWV, rings); break;
    case 3: band_rows<3, RP>(c, tab, pr, w, t, NWV, rings); break;
    case 4: band_rows<4, RP>(c, tab, pr, w, t, NWV, rings); break;
    case 6: band_rows<6, RP>(c, tab, pr, w, t, NWV, rings); break;
    case 8: band_rows<8, RP>(c, tab, pr, w, t, NWV, rings); break;
    default: break;
    }
}

// ------------------------------------------------------------------ host

namespace {

typedef void (*BFn)(BandCfg);

// instances by the largest slots per lane they hold (registers grow with it)
template <bool LTAB>
BFn chain_kernel_t(int rmax)
{
    if (rmax <= 6) return k_banded_chain<6, LTAB>;
    if (rmax <= 8) return k_banded_chain<8, LTAB>;
    if (rmax <= 16) return k_banded_chain<16, LTAB>;
    return k_banded_chain<32, LTAB>;
}

BFn chain_kernel(int rmax, bool ltab)
{
    return ltab ? chain_kernel_t<true>(rmax) : chain_kernel_t<false>(rmax);
}

const int kBandRs[] = {2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32};

// k_banded_seg: slots per lane from this set (even: the profile rotation
// keeps a step's two prefetch registers on the same parity)
const int kSegRs[] = {8, 10, 12, 14, 16, 20, 24, 32};
// bytes after each residue array: the segment kernel's loads reach up to
// 64 x 32 + 2 bytes past a pair's query and 2 past its target
constexpr size_t kResidueSlack = 4096;
// segment columns: 0 no segments, 1 sized to fill the device once, more:
// about that many columns a segment
int g_seg_cols = 1;
int g_seg_stats[3];  // last batch: pairs on k_banded_seg, segments, pairs redone without
// warm-up columns before a segment's own: three band widths plus 256.  On
// long mutated DNA pairs (10% substitutions, 2% indels, k = 256) the states
// of 3,169 segment starts lined up after 0.8 band widths on average and
// after 2.6 at most.
int64_t seg_warmup(int64_t colLen) { return (3 * colLen + 256 + 63) & ~(int64_t)63; }

struct SegCand {
    int pair, R;
    int64_t n, colLen, X, Lmin;
};
// instruction issues of one step of a one-wave chain with R slots a lane
int64_t seg_step_cost(int R) { return 9 * R + 12; }

// Segments per candidate pair.  Fixed columns: about cols own columns a
// segment.  Automatic: the least run cost D whose segments number at most
// waves (one round of the device's resident waves), every segment's run
// (warm-up, own columns, 64 steps of lane fill) costing at most D.
std::vector<int64_t> seg_plan(const std::vector<SegCand> &cs, int cols, int64_t waves)
{
    std::vector<int64_t> S(cs.size(), 1);
    auto count = [&](const SegCand &c, int64_t D) -> int64_t {
        const int64_t cost = seg_step_cost(c.R);
        if ((c.n + 64) * cost <= D) return 1;
        const int64_t L = (D / cost - c.X - 64) & ~(int64_t)1;
        if (L < c.Lmin) return std::max<int64_t>(1, (c.n - c.X) / c.Lmin);  // the finest split
        return (c.n - c.X + L - 1) / L;
    };
    if (cols >= 2) {
        for (size_t x = 0; x < cs.size(); ++x)
            S[x] = std::max<int64_t>(1, (cs[x].n - cs[x].X + cols / 2) / cols);
    } else {
        int64_t lo = 1, hi = 1;
        for (const SegCand &c : cs) hi = std::max(hi, (c.n + 64) * seg_step_cost(c.R));
        while (lo < hi) {
            const int64_t mid = lo + (hi - lo) / 2;
            int64_t tot = 0;
            for (const SegCand &c : cs) tot += count(c, mid);
            if (tot <= waves) hi = mid;
            else lo = mid + 1;
        }
        for (size_t x = 0; x < cs.size(); ++x) S[x] = count(cs[x], lo);
    }
    // own columns of at least Lmin each, or one segment
    for (size_t x = 0; x < cs.size(); ++x) {
        if (S[x] >= 2) S[x] = std::min<int64_t>(S[x], (cs[x].n - cs[x].X) / cs[x].Lmin);
        if (S[x] < 2) S[x] = 1;
    }
    return S;
}

// k_banded_rows: four waves per pair, rows per lane from this set
constexpr int kRowsWaves = 4;
const int kRowsRs[] = {1, 2, 3, 4, 6, 8};
constexpr size_t kLdsMax = 160 * 1024;

// steps of one row band (band_rows' send)
int64_t rows_band_steps(int R, int64_t colLen) { return (64 * R + colLen + R + 64 + 3) & ~(int64_t)3; }
// bottom-row slots a wave's LDS row needs: the band's steps, plus the
// BR + 66 further slots lane 0 of the next band loads past them
int64_t rows_slots(int R, int64_t colLen) { return rows_band_steps(R, colLen) + 64 * R + 66; }

// the pair's steps on the row-band kernel, with R rows per lane
int64_t rows_steps(int R, int64_t colLen, int64_t m)
{
    const int64_t send = rows_band_steps(R, colLen);
    const int64_t nb = (m + 64 * R - 1) / (64 * R);
    return (nb - 1) * rows_spacing(64 * R, (int)send, kRowsWaves) + send;
}

// rows per lane for an m-row query: the least steps times the measured
// cost of a step (about 27 + 10 R instruction issues)
int rows_pick(int64_t colLen, int64_t m)
{
    int best = 0;
    int64_t best_cost = 0;
    for (int R : kRowsRs) {
        const int64_t cost = rows_steps(R, colLen, m) * (27 + 10 * R);
        if (!best || cost < best_cost) {
            best = R;
            best_cost = cost;
        }
    }
    return best;
}

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

// inside the locked region: record the failure and leave the do-block
#define BOK(expr)                                                                 \
    if ((expr) != hipSuccess) {                                                   \
        phip_set_error("%s:%d %s failed", __FILE__, __LINE__, #expr);            \
        break;                                                                    \
    }

static int banded_impl(const phip_batch_t *b, int k, phip_out_t *out, bool allow_seg)
{
    const int np = b->npairs;
    if (np <= 0) return 0;
    // lane chains of 64 x nwv lanes (nwv waves of a workgroup per pair:
    // more for few pairs, which otherwise leave most SIMDs idle) for bands
    // up to 64 x nwv x 32 slots, the single-lane loop for wider ones; gaps
    // large enough to overflow int32 over a band fall back too
    const int nwv = np <= 512 ? 4 : (np <= 1024 ? 2 : 1);
    std::vector<int32_t> chain, rsel(np, 0), wide, rows;
    std::vector<int64_t> work(np, 0);
    int rmax = 0;
    int64_t qtot = 0, ttot = 0, maxcol = 0, rows_send = 0;  // rows_send: LDS slots per wave
    // row bands (k_banded_rows) for pairs of >= 512 query rows whose band
    // fits the rings, when every W + open fits a 4-bit register field
    // (square or PSSM, <= 8 letters) or the square matrix an LDS table, and
    // every score of the pair stays far above the kernel's -inf
    const bool is_pssm = b->matrix.is_pssm;
    const int msz = b->matrix.size;
    bool rows_rp = msz <= 8;
    int64_t wabs = 0;
    for (int64_t x = 0; x < (int64_t)msz * b->matrix.length; ++x) {
        const int v = b->matrix.values[x];
        wabs = std::max<int64_t>(wabs, v < 0 ? -(int64_t)v : v);
        if (v + b->open < 0 || v + b->open > 15) rows_rp = false;
    }
    const bool rows_ok = rows_rp || (!is_pssm && msz <= 64);
    // segments (k_banded_seg): square matrices staged in LDS
    const bool seg_ok = !is_pssm && msz <= 64;
    std::vector<SegCand> cands;
    std::vector<SegDesc> sdesc;
    std::vector<SegPair> spairs;
    int64_t vtot = 0;
    int seg_rmax = 0;
    for (int i = 0; i < np; ++i) {
        const int m = b->qlen[i], n = b->tlen[i];
        if (m <= 0 || n <= 0) {
            phip_set_error("phip_nw_banded_batch: pair %d is empty", i);
            return -1;
        }
        qtot = std::max<int64_t>(qtot, b->qoff[i] + m);
        ttot = std::max<int64_t>(ttot, b->toff[i] + n);
        const int64_t low = (m > n) ? (int64_t)n - m - k : (int64_t)m - n - k;
        const int64_t colLen = k - low + 1;
        maxcol = std::max(maxcol, colLen);
        const bool bounded = ((int64_t)m + n + 2) * ((int64_t)b->open + b->gap + wabs) < (1 << 27);
        // one-wave lane chains over segments (k_banded_seg) for bands of
        // 384 to 2,048 slots (8 to 32 per lane, at least 3/4 of them filled)
        if (allow_seg && seg_ok && bounded && colLen >= 384 && colLen <= 64 * 32 && g_seg_cols > 0) {
            SegCand sc;
            sc.pair = i;
            sc.n = n;
            sc.colLen = colLen;
            sc.R = 0;
            for (int r : kSegRs)
                if (colLen <= 64 * r) { sc.R = r; break; }
            sc.X = seg_warmup(colLen);
            const int64_t colOff0 = (m > n) ? -(int64_t)k : low;
            // a segment's own columns: at least its warm-up, and every step
            // of a later segment's run a fast one
            sc.Lmin = std::max(sc.X, (std::max<int64_t>(63, -colOff0) + 1) & ~(int64_t)1);
            cands.push_back(sc);
            continue;
        }
        if (rows_ok && m >= 512 && m <= 1 << 24 && colLen <= 1 << 14 &&
            ((int64_t)m + n + 2) * ((int64_t)b->open + b->gap + wabs) < (1 << 27)) {
            const int R = rows_pick(colLen, m);
            const int64_t steps = R ? rows_band_steps(R, colLen) : 0;
            const int64_t slots = R ? std::max(rows_slots(R, colLen), rows_send) : 0;
            const size_t tabb = (size_t)(((rows_rp ? msz : msz * msz) + 3) & ~3) * 4;
            if (R && tabb + (size_t)kRowsWaves * slots * 8 <= kLdsMax) {
                rows.push_back(i);
                rsel[i] = R;
                rows_send = slots;
                work[i] = rows_steps(R, colLen, m) * (27 + 10 * R);
                continue;
            }
        }
        int R = 0;
        for (int r : kBandRs)
            if (colLen <= 64 * nwv * r) { R = r; break; }
        if (R == 0 || b->open > (1 << 20) || b->gap > (1 << 20)) {
            wide.push_back(i);
        } else {
            chain.push_back(i);
            rsel[i] = R;
            work[i] = (int64_t)(n + 64 * nwv) * (R + 4);
            rmax = std::max(rmax, R);
        }
    }
    if (!cands.empty()) {
        // segments of each candidate pair: the first from the reference's
        // start, the others from X warm-up columns before their own L
        // columns (the last owns the rest), each storing the states the join
        // compares
        int cus = 0, clk = 0;
        if (phip_device_info(&cus, &clk) != 0 || cus <= 0) cus = 256;
        const std::vector<int64_t> S = seg_plan(cands, g_seg_cols, (int64_t)cus * 8);
        for (size_t x = 0; x < cands.size(); ++x) {
            const SegCand &sc = cands[x];
            const int64_t n = sc.n, X = sc.X, Sx = S[x];
            const int64_t L = Sx >= 2 ? ((n - X) / Sx) & ~(int64_t)1 : 0;
            SegPair p;
            p.pair = sc.pair;
            p.first = (int)sdesc.size();
            p.count = (int)Sx;
            p.colLen = (int)sc.colLen;
            spairs.push_back(p);
            for (int64_t q = 0; q < Sx; ++q) {
                const int64_t J = q == 0 ? 0 : X + q * L, Jn = q + 1 == Sx ? n : X + (q + 1) * L;
                SegDesc d;
                d.pair = sc.pair;
                d.R = sc.R;
                d.c0 = (int)(q == 0 ? 0 : J - X);
                d.len = (int)(q + 1 == Sx ? ((n + 64 - d.c0) & ~(int64_t)1) : Jn + 64 - d.c0);
                d.cs = (int)(q == 0 ? -1 : J - 1);
                d.ce = (int)(q + 1 == Sx ? -1 : Jn - 1);
                d.voff = vtot;
                vtot += 2 * 64 * sc.R;
                sdesc.push_back(d);
            }
            seg_rmax = std::max(seg_rmax, sc.R);
        }
    }
    // one launch each; the longest chains start first
    std::stable_sort(chain.begin(), chain.end(), [&](int x, int y) { return work[x] > work[y]; });
    std::stable_sort(rows.begin(), rows.end(), [&](int x, int y) { return work[x] > work[y]; });
    const size_t mbytes = (size_t)b->matrix.size * b->matrix.length * 4;
    const int64_t stride = maxcol + 2;
    size_t off = 0;
    auto slot = [&](size_t bytes) { size_t o = off; off += al256(bytes ? bytes : 1); return o; };
    const size_t o_mat = slot(mbytes), o_q = slot((size_t)qtot + kResidueSlack),
                 o_t = slot((size_t)ttot + kResidueSlack),
                 o_qo = slot((size_t)np * 8), o_to = slot((size_t)np * 8), o_ql = slot((size_t)np * 4),
                 o_tl = slot((size_t)np * 4), o_pairs = slot((size_t)np * 4), o_rsel = slot((size_t)np * 4),
                 o_out = slot((size_t)np * sizeof(phip_out_t)),
                 o_scr = slot(wide.empty() ? 0 : (size_t)wide.size() * 2 * stride * 4);
    const size_t nseg = sdesc.size();
    const size_t o_sd = slot(nseg * sizeof(SegDesc)), o_so = slot(nseg * 4),
                 o_sp = slot(spairs.size() * sizeof(SegPair)), o_vec = slot((size_t)vtot * sizeof(int2)),
                 o_raw = slot(nseg * 4);
    // a pair's segments are consecutive descriptors (the join walks them in
    // order); they launch costliest first, 4 to a workgroup, through sord
    std::vector<int32_t> sord(nseg);
    for (size_t x = 0; x < nseg; ++x) sord[x] = (int32_t)x;
    std::stable_sort(sord.begin(), sord.end(), [&](int32_t x, int32_t y) {
        return (int64_t)sdesc[x].len * (sdesc[x].R + 6) > (int64_t)sdesc[y].len * (sdesc[y].R + 6);
    });
    phip_lock();
    char *base = (char *)dev_alloc(off);
    if (!base) {
        phip_unlock();
        return -1;
    }
    hipStream_t st = phip_stream();
    int rc = -1;
    do {
        BandCfg c;
        memset(&c, 0, sizeof(c));
        c.q = (const uint8_t *)(base + o_q);
        c.qoff = (const int64_t *)(base + o_qo);
        c.qlen = (const int32_t *)(base + o_ql);
        c.t = (const uint8_t *)(base + o_t);
        c.toff = (const int64_t *)(base + o_to);
        c.tlen = (const int32_t *)(base + o_tl);
        c.mat = (const int *)(base + o_mat);
        c.msize = b->matrix.size;
        c.is_pssm = b->matrix.is_pssm;
        c.open = b->open;
        c.gap = b->gap;
        c.k = k;
        c.out = (phip_out_t *)(base + o_out);
        int32_t *dpairs = (int32_t *)(base + o_pairs);
        bool ok = true;
        ok &= hipMemcpyAsync((void *)c.mat, b->matrix.values, mbytes, hipMemcpyHostToDevice, st) == hipSuccess;
        if (qtot) ok &= hipMemcpyAsync((void *)c.q, b->q, (size_t)qtot, hipMemcpyHostToDevice, st) == hipSuccess;
        ok &= hipMemcpyAsync((void *)c.t, b->t, (size_t)ttot, hipMemcpyHostToDevice, st) == hipSuccess;
        // the slack after each residue array (read by k_banded_seg's
        // unclamped loads past a pair's end): code 0
        ok &= hipMemsetAsync((void *)(c.q + qtot), 0, kResidueSlack, st) == hipSuccess;
        ok &= hipMemsetAsync((void *)(c.t + ttot), 0, kResidueSlack, st) == hipSuccess;
        ok &= hipMemcpyAsync((void *)c.qoff, b->qoff, (size_t)np * 8, hipMemcpyHostToDevice, st) == hipSuccess;
        ok &= hipMemcpyAsync((void *)c.toff, b->toff, (size_t)np * 8, hipMemcpyHostToDevice, st) == hipSuccess;
        ok &= hipMemcpyAsync((void *)c.qlen, b->qlen, (size_t)np * 4, hipMemcpyHostToDevice, st) == hipSuccess;
        ok &= hipMemcpyAsync((void *)c.tlen, b->tlen, (size_t)np * 4, hipMemcpyHostToDevice, st) == hipSuccess;
        int32_t *drsel = (int32_t *)(base + o_rsel);
        std::vector<int32_t> order(chain), rs;
        order.insert(order.end(), wide.begin(), wide.end());
        order.insert(order.end(), rows.begin(), rows.end());
        for (int i : chain) rs.push_back(rsel[i]);
        for (size_t x = 0; x < wide.size(); ++x) rs.push_back(0);
        for (int i : rows) rs.push_back(rsel[i]);
        ok &= hipMemcpyAsync(dpairs, order.data(), order.size() * 4, hipMemcpyHostToDevice, st) == hipSuccess;
        if (!rs.empty())
            ok &= hipMemcpyAsync(drsel, rs.data(), rs.size() * 4, hipMemcpyHostToDevice, st) == hipSuccess;
        if (nseg) {
            ok &= hipMemcpyAsync(base + o_sd, sdesc.data(), nseg * sizeof(SegDesc), hipMemcpyHostToDevice,
                                 st) == hipSuccess;
            ok &= hipMemcpyAsync(base + o_so, sord.data(), nseg * 4, hipMemcpyHostToDevice, st) == hipSuccess;
            ok &= hipMemcpyAsync(base + o_sp, spairs.data(), spairs.size() * sizeof(SegPair),
                                 hipMemcpyHostToDevice, st) == hipSuccess;
        }
        if (!ok) {
            phip_set_error("phip_nw_banded_batch: upload failed");
            break;
        }
        BOK(hipEventRecord(phip_event(0), st));
        const bool ltab = !b->matrix.is_pssm && b->matrix.size <= 64;
        const size_t lds = (ltab ? (size_t)((b->matrix.size * b->matrix.size + 1) & ~1) * 4 : 0) + 16 * 8;
        if (!chain.empty()) {
            BandCfg cc = c;
            cc.pairs = dpairs;
            cc.rsel = drsel;
            cc.npairs = (int)chain.size();
            hipLaunchKernelGGL(chain_kernel(rmax, ltab), dim3(cc.npairs), dim3(64 * nwv), lds, st, cc);
            BOK(hipGetLastError());
        }
        if (!wide.empty()) {
            BandCfg cc = c;
            cc.pairs = dpairs + chain.size();
            cc.npairs = (int)wide.size();
            hipLaunchKernelGGL(k_banded_one, dim3(cc.npairs), dim3(64), 0, st, cc, (int *)(base + o_scr), stride);
            BOK(hipGetLastError());
        }
        if (!rows.empty()) {
            BandCfg cc = c;
            const size_t base_i = chain.size() + wide.size();
            cc.pairs = dpairs + base_i;
            cc.rsel = drsel + base_i;
            cc.npairs = (int)rows.size();
            const int lr = (int)rows_send;  // even: band steps and 64 R + 66 are
            cc.lring = lr;
            const size_t tabb = (size_t)(((rows_rp ? msz : msz * msz) + 3) & ~3) * 4;
            const size_t lds_r = tabb + (size_t)kRowsWaves * lr * 8;
            auto fn = rows_rp ? k_banded_rows<true> : k_banded_rows<false>;
            BOK(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_r));
            hipLaunchKernelGGL(fn, dim3(cc.npairs), dim3(64 * kRowsWaves), lds_r, st, cc);
            BOK(hipGetLastError());
        }
        if (nseg) {
            const size_t lds_s = (size_t)(((msz * msz + 3) & ~3) + (rows_rp ? msz : 0)) * 4;
            auto fn = rows_rp ? (seg_rmax <= 16 ? k_banded_seg<16, true> : k_banded_seg<32, true>)
                              : (seg_rmax <= 16 ? k_banded_seg<16, false> : k_banded_seg<32, false>);
            const SegDesc *dsd = (const SegDesc *)(base + o_sd);
            hipLaunchKernelGGL(fn, dim3((unsigned)((nseg + 3) / 4)), dim3(256), lds_s, st, c, dsd,
                               (const int32_t *)(base + o_so), (int)nseg, (int2 *)(base + o_vec),
                               (int *)(base + o_raw));
            BOK(hipGetLastError());
            hipLaunchKernelGGL(k_seg_join, dim3((unsigned)spairs.size()), dim3(64), 0, st, c,
                               (const SegPair *)(base + o_sp), dsd, (const int2 *)(base + o_vec),
                               (const int *)(base + o_raw));
            BOK(hipGetLastError());
        }
        BOK(hipEventRecord(phip_event(1), st));
        BOK(hipMemcpyAsync(out, c.out, (size_t)np * sizeof(phip_out_t), hipMemcpyDeviceToHost, st));
        BOK(hipStreamSynchronize(st));
        phip_note_kernel_time();
        rc = 0;
    } while (0);
    dev_free(base);
    phip_unlock();
    if (rc != 0 || spairs.empty()) return rc;
    // pairs whose segments did not line up: again, without segments
    std::vector<int64_t> fqo, fto;
    std::vector<int32_t> fql, ftl, fidx;
    for (const SegPair &p : spairs)
        if (out[p.pair].status & PHIP_ST_ERR) {
            fidx.push_back(p.pair);
            fqo.push_back(b->qoff[p.pair]);
            fto.push_back(b->toff[p.pair]);
            fql.push_back(b->qlen[p.pair]);
            ftl.push_back(b->tlen[p.pair]);
        }
    g_seg_stats[0] = (int)spairs.size();
    g_seg_stats[1] = (int)nseg;
    g_seg_stats[2] = (int)fidx.size();
    if (fidx.empty()) return 0;
    const double kms = phip_last_kernel_ms();
    phip_batch_t sub = *b;
    sub.npairs = (int)fidx.size();
    sub.qoff = fqo.data();
    sub.toff = fto.data();
    sub.qlen = fql.data();
    sub.tlen = ftl.data();
    std::vector<phip_out_t> so(fidx.size());
    rc = banded_impl(&sub, k, so.data(), false);
    if (rc != 0) return rc;
    for (size_t x = 0; x < fidx.size(); ++x) out[fidx[x]] = so[x];
    phip_add_kernel_ms(kms);
    return 0;
}

extern "C" int phip_nw_banded_batch(const phip_batch_t *b, int k, phip_out_t *out)
{
    if (phip_init() != 0) return -1;
    // s1 is required even with a PSSM, as in the reference (src/nw_banded.c:50)
    if (b->npairs > 0 && (k <= 0 || b->open < 0 || b->gap < 0 || !b->q || !b->t)) {
        phip_set_error("phip_nw_banded_batch: bad arguments");
        return -1;
    }
    g_seg_stats[0] = g_seg_stats[1] = g_seg_stats[2] = 0;
    return banded_impl(b, k, out, true);
}

extern "C" int phip_nw_banded_segment_cols(int cols)
{
    const int prev = g_seg_cols;
    if (cols >= 0) g_seg_cols = cols;
    return prev;
}

extern "C" void phip_nw_banded_last_stats(int *pairs, int *segments, int *redone)
{
    if (pairs) *pairs = g_seg_stats[0];
    if (segments) *segments = g_seg_stats[1];
    if (redone) *redone = g_seg_stats[2];
}
