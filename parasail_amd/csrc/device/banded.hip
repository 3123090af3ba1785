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
            __syncthreads();  // A
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
            __syncthreads();  // B
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
    const int NWV = blockDim.x >> 6, w = threadIdx.x >> 6, t = threadIdx.x & 63;
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

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

// inside the locked region: record the failure and leave the do-block
#define BOK(expr)                                                                 \
    if ((expr) != hipSuccess) {                                                   \
        phip_set_error("%s:%d %s failed", __FILE__, __LINE__, #expr);            \
        break;                                                                    \
    }

extern "C" int phip_nw_banded_batch(const phip_batch_t *b, int k, phip_out_t *out)
{
    if (phip_init() != 0) return -1;
    const int np = b->npairs;
    if (np <= 0) return 0;
    // s1 is required even with a PSSM, as in the reference (src/nw_banded.c:50)
    if (k <= 0 || b->open < 0 || b->gap < 0 || !b->q || !b->t) {
        phip_set_error("phip_nw_banded_batch: bad arguments");
        return -1;
    }
    // lane chains of 64 x nwv lanes (nwv waves of a workgroup per pair:
    // more for few pairs, which otherwise leave most SIMDs idle) for bands
    // up to 64 x nwv x 32 slots, the single-lane loop for wider ones; gaps
    // large enough to overflow int32 over a band fall back too
    const int nwv = np <= 512 ? 4 : (np <= 1024 ? 2 : 1);
    std::vector<int32_t> chain, rsel(np, 0), wide;
    std::vector<int64_t> work(np, 0);
    int rmax = 0;
    int64_t qtot = 0, ttot = 0, maxcol = 0;
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
    // one launch; the longest chains start first
    std::stable_sort(chain.begin(), chain.end(), [&](int x, int y) { return work[x] > work[y]; });
    const size_t mbytes = (size_t)b->matrix.size * b->matrix.length * 4;
    const int64_t stride = maxcol + 2;
    size_t off = 0;
    auto slot = [&](size_t bytes) { size_t o = off; off += al256(bytes ? bytes : 1); return o; };
    const size_t o_mat = slot(mbytes), o_q = slot((size_t)qtot), o_t = slot((size_t)ttot),
                 o_qo = slot((size_t)np * 8), o_to = slot((size_t)np * 8), o_ql = slot((size_t)np * 4),
                 o_tl = slot((size_t)np * 4), o_pairs = slot((size_t)np * 4), o_rsel = slot((size_t)np * 4),
                 o_out = slot((size_t)np * sizeof(phip_out_t)),
                 o_scr = slot(wide.empty() ? 0 : (size_t)wide.size() * 2 * stride * 4);
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
        ok &= hipMemcpyAsync((void *)c.qoff, b->qoff, (size_t)np * 8, hipMemcpyHostToDevice, st) == hipSuccess;
        ok &= hipMemcpyAsync((void *)c.toff, b->toff, (size_t)np * 8, hipMemcpyHostToDevice, st) == hipSuccess;
        ok &= hipMemcpyAsync((void *)c.qlen, b->qlen, (size_t)np * 4, hipMemcpyHostToDevice, st) == hipSuccess;
        ok &= hipMemcpyAsync((void *)c.tlen, b->tlen, (size_t)np * 4, hipMemcpyHostToDevice, st) == hipSuccess;
        int32_t *drsel = (int32_t *)(base + o_rsel);
        std::vector<int32_t> order(chain), rs;
        order.insert(order.end(), wide.begin(), wide.end());
        for (int i : chain) rs.push_back(rsel[i]);
        ok &= hipMemcpyAsync(dpairs, order.data(), order.size() * 4, hipMemcpyHostToDevice, st) == hipSuccess;
        if (!rs.empty())
            ok &= hipMemcpyAsync(drsel, rs.data(), rs.size() * 4, hipMemcpyHostToDevice, st) == hipSuccess;
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
        BOK(hipEventRecord(phip_event(1), st));
        BOK(hipMemcpyAsync(out, c.out, (size_t)np * sizeof(phip_out_t), hipMemcpyDeviceToHost, st));
        BOK(hipStreamSynchronize(st));
        phip_note_kernel_time();
        rc = 0;
    } while (0);
    dev_free(base);
    phip_unlock();
    return rc;
}
