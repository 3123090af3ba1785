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
    int npairs;
    phip_out_t *out;
};

// Band slots per lane: R (colLen <= 64 R).
template <int R>
__global__ void __launch_bounds__(256) k_banded_chain(BandCfg c)
{
    extern __shared__ int tab[];
    const bool ltab = !c.is_pssm && c.msize <= 64;
    if (ltab)
        for (int x = threadIdx.x; x < c.msize * c.msize; x += blockDim.x) tab[x] = c.mat[x];
    __syncthreads();
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= c.npairs) return;  // whole wave
    const int t = threadIdx.x & 63;
    const int pr = c.pairs[w];
    const int m = c.qlen[pr], n = c.tlen[pr];
    const uint8_t *s1 = c.q + c.qoff[pr];
    const uint8_t *s2 = c.t + c.toff[pr];
    const int open = c.open, gap = c.gap, k = c.k, msize = c.msize;
    const int low = (m > n) ? n - m - k : m - n - k;
    const int colLen = k - low + 1;
    const int colOff0 = (m > n) ? -k : low;
    const int tR = t * R;
    const int *mrow = c.mat;

    // the reference's initial arrays (src/nw_banded.c:110-138)
    int Hb[R], Eb[R], qc[R];
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
        // the lane's codes shift one row per step from step 0 on, when its
        // column is -t: start from that column's rows
        const int pos = colOff0 - t + i;
        qc[r] = s1[pos < 0 ? 0 : (pos >= m ? m - 1 : pos)];
    }
    // clamped, unconditional loads: the residue of a row outside the query
    // is never used (its slot is skipped)
    auto ldq = [&](int pos) -> int { return (int)s1[pos < 0 ? 0 : (pos >= m ? m - 1 : pos)]; };
    auto ldt = [&](int j) -> int { return (int)s2[j < 0 ? 0 : (j >= n ? n - 1 : j)]; };
    int botH = BNEG, botF = BNEG;
    // codes fetched two steps ahead: the target residue of the lane's column
    // and the query residue entering its last slot at the next column
    int tq0 = ldt(-t), tq1 = ldt(1 - t);
    int nq0 = ldq(colOff0 - t + tR + R), nq1 = ldq(colOff0 + 1 - t + tR + R);

    auto step = [&](const int s, int &tq, int &nq) __attribute__((always_inline)) {
        const int j = s - t;
        const bool act = j >= 0 && j < n;
        const int xH = shr1(botH), xF = shr1(botF);
        int Hup = t == 0 ? BNEG : xH;  // slot -1 of lane 0 is the -inf edge
        int F = t == 0 ? BNEG : xF;
        const int colOff = colOff0 + j;
        if (colOff0 + s - 63 < 0) {
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
            val[r] = act && pos >= 0 && pos < m && tR + r < colLen;
            if (c.is_pssm) {
                const int pc = pos < 0 ? 0 : (pos >= m ? m - 1 : pos);
                W[r] = mrow[(size_t)msize * pc + tc];
            } else {
                W[r] = ltab ? tab[qc[r] * msize + tc] : mrow[qc[r] * msize + tc];
            }
        }
        auto slot = [&](int r, int Hn, int En) __attribute__((always_inline)) {
            const int Hdag = Hb[r] + W[r];
            const int E = max(Hn - open, En - gap);
            const int Fn = max(Hup - open, F - gap);
            const int H = max(max(E, Fn), Hdag);
            const bool v = val[r];
            Eb[r] = v ? E : Eb[r];
            F = v ? Fn : F;
            Hb[r] = v ? H : Hb[r];
            Hup = Hb[r];
        };
        slot(0, Hb[1], Eb[1]);
        // lane t+1's first slot, just computed for column j-1, is the left
        // neighbour of this lane's last slot
        int xHn = shl1(Hb[0]), xEn = shl1(Eb[0]);
        if (t == 63) { xHn = BNEG; xEn = BNEG; }
#pragma unroll
        for (int r = 1; r < R; ++r) slot(r, r + 1 < R ? Hb[r + 1] : xHn, r + 1 < R ? Eb[r + 1] : xEn);
        botH = Hup;
        botF = F;
        // next column: every slot moves one row down the query
#pragma unroll
        for (int r = 0; r < R - 1; ++r) qc[r] = qc[r + 1];
        qc[R - 1] = nq;
        tq = ldt(j + 2);
        nq = ldq(colOff + 2 + tR + R);
    };
    const int steps = n + 63;
    for (int s = 0; s < steps; s += 2) {
        step(s, tq0, nq0);
        step(s + 1, tq1, nq1);
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

BFn chain_kernel(int R)
{
    switch (R) {
    case 2: return k_banded_chain<2>;
    case 4: return k_banded_chain<4>;
    case 8: return k_banded_chain<8>;
    case 12: return k_banded_chain<12>;
    case 16: return k_banded_chain<16>;
    case 24: return k_banded_chain<24>;
    default: return k_banded_chain<32>;
    }
}

const int kBandRs[] = {2, 4, 8, 12, 16, 24, 32};

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
    // lane chains for bands up to 64 x 32 slots (bucketed by slots per
    // lane), the single-lane loop for wider ones; gaps large enough to
    // overflow int32 over a band fall back too
    std::vector<std::vector<int32_t>> bins(sizeof(kBandRs) / sizeof(kBandRs[0]));
    std::vector<int32_t> wide;
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
        int bin = -1;
        for (size_t x = 0; x < bins.size(); ++x)
            if (colLen <= 64 * kBandRs[x]) { bin = (int)x; break; }
        if (bin < 0 || b->open > (1 << 20) || b->gap > (1 << 20)) wide.push_back(i);
        else bins[bin].push_back(i);
    }
    const size_t mbytes = (size_t)b->matrix.size * b->matrix.length * 4;
    const int64_t stride = maxcol + 2;
    size_t off = 0;
    auto slot = [&](size_t bytes) { size_t o = off; off += al256(bytes ? bytes : 1); return o; };
    const size_t o_mat = slot(mbytes), o_q = slot((size_t)qtot), o_t = slot((size_t)ttot),
                 o_qo = slot((size_t)np * 8), o_to = slot((size_t)np * 8), o_ql = slot((size_t)np * 4),
                 o_tl = slot((size_t)np * 4), o_pairs = slot((size_t)np * 4),
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
        std::vector<int32_t> order;
        for (auto &v : bins) order.insert(order.end(), v.begin(), v.end());
        order.insert(order.end(), wide.begin(), wide.end());
        ok &= hipMemcpyAsync(dpairs, order.data(), order.size() * 4, hipMemcpyHostToDevice, st) == hipSuccess;
        if (!ok) {
            phip_set_error("phip_nw_banded_batch: upload failed");
            break;
        }
        BOK(hipEventRecord(phip_event(0), st));
        size_t at = 0;
        const size_t lds = (!b->matrix.is_pssm && b->matrix.size <= 64)
                               ? (size_t)b->matrix.size * b->matrix.size * 4 : 0;
        for (size_t x = 0; x < bins.size(); ++x) {
            if (!bins[x].empty()) {
                BandCfg cc = c;
                cc.pairs = dpairs + at;
                cc.npairs = (int)bins[x].size();
                hipLaunchKernelGGL(chain_kernel(kBandRs[x]), dim3((cc.npairs + 3) / 4), dim3(256), lds, st, cc);
                if (hipGetLastError() != hipSuccess) { ok = false; break; }
            }
            at += bins[x].size();
        }
        if (!ok) {
            phip_set_error("phip_nw_banded_batch: launch failed");
            break;
        }
        if (!wide.empty()) {
            BandCfg cc = c;
            cc.pairs = dpairs + at;
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
