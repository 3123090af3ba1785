// dp32.hip -- _stats and _trace of SW / NW / SG (src/sw_stats.c:204-261,
// src/sg_stats.c:227-309, src/sg_trace.c:144-150, src/sw_trace.c:145-147 and
// the matching nw files) as wavefront lane chains in int32, and the CIGAR
// walk of src/cigar_template.c:58-263 on the device.
//
// One pair per job of T lanes (T = 16, 32 or 64; 64/T jobs per wave, all
// of one shape); lane t holds query rows t*R .. t*R+R-1 in registers, the
// whole query in one band (m <= T*R), and computes column s - t at step s.
// The row above comes from lane t-1 by DPP; H is kept as H - open.  Each
// cell's statistics (matches, similar, length) are packed in one word --
// length in bits 0-11, similar in 12-21, matches in 22-31 -- so one select
// moves all three and the diagonal move adds one table constant; pairs with
// min(m, n) > 1023 or m + n > 4095 take the general tier.  Query rows past
// m read a pad row of the table (substitution -2^24), so their values stay
// below every real value of their column and never win an end.
//
// Traceback codes (parasail.h:31-41 values) go to HBM in wavefront order:
// pair base + ((s*T + t)*R + r) bytes for row t*R+r at step s, so each
// step of a job is one contiguous store of T*R bytes.  k_cigar reads that
// layout (or row-major, for pairs of the general tier) and writes each
// pair's CIGAR ops.
#include <string.h>

#include <algorithm>
#include <vector>

#include "dp32.h"

namespace {

constexpr int kPadSub = -(1 << 24);
constexpr int kNeg = PHIP_NEG_INF32;

__device__ __forceinline__ int shr1(int v) { return __builtin_amdgcn_mov_dpp(v, 0x138, 0xF, 0xF, false); }

}  // namespace

// One job: a pair, its shape in the wave and where its trace goes.
struct Dp32Job {
    int pair;
    int status;      // PHIP_ST_* reported with the result
    int64_t troff;   // byte offset of the pair's trace (kind == TRACE)
};

// One wave: njobs consecutive jobs of lanes T and rows R, steps columns.
struct Dp32Wave {
    int job0, njobs, T, R, steps, pad0, pad1, pad2;
};

struct Dp32Cfg {
    const uint8_t *q;
    const int64_t *qoff;
    const int32_t *qlen;
    const uint8_t *t;
    const int64_t *toff;
    const int32_t *tlen;
    const int *mat;
    int msize;
    int open, gap, alg, sg_flags;
    const Dp32Job *jobs;
    const Dp32Wave *waves;
    int nwaves;
    int8_t *trace;
    phip_out_t *out;
    int *dtab[4];  // TAB: score, matches, similar, length tables of a one-pair launch
};

// KIND: PHIP_SCORE, PHIP_STATS or PHIP_TRACE.  EXT: also the exact extrema
// of H, E and F over the pair's cells and boundaries, the smallest cell H
// and (_stats) the largest length count of a cell's H, for the caller's
// saturation test (k_simple's `check`).  TAB (one pair per launch): 1 writes
// every cell's H (and, for _stats, its H's counts) row-major into dtab[],
// 2 only the last row (n entries) followed by the last column (m entries),
// the PARASAIL_TABLE / PARASAIL_ROWCOL outputs of src/sw.c:159-170.
template <int R, int KIND, int ALG, bool EXT, int TAB>
__device__ __forceinline__ void dp32_run(const Dp32Cfg &c, const Dp32Wave &wd, const int2 *tab,
                                         const int lane)
{
    constexpr bool STATS = KIND == PHIP_STATS, TRACE = KIND == PHIP_TRACE;
    constexpr bool SW = ALG == PHIP_SW;
    constexpr int alg = ALG;
    const int T = wd.T;
    const int lt = T == 64 ? 6 : (T == 32 ? 5 : 4);
    const int jl = lane >> lt, t = lane & (T - 1);
    const bool jv = jl < wd.njobs;
    const Dp32Job J = c.jobs[wd.job0 + (jv ? jl : 0)];
    const int pr = J.pair;
    const int m = c.qlen[pr], nreal = c.tlen[pr], n = jv ? nreal : 0;
    const uint8_t *s1 = c.q + c.qoff[pr];
    const uint8_t *s2 = c.t + c.toff[pr];
    const int open = c.open, gap = c.gap, msize = c.msize;
    const int fl = alg == PHIP_SG ? c.sg_flags : 0;
    const bool s1b = fl & PHIP_S1_BEG, s1e = fl & PHIP_S1_END;
    const bool s2b = fl & PHIP_S2_BEG, s2e = fl & PHIP_S2_END;
    const bool zero_left = SW || s1b, zero_top = SW || s2b;
    int8_t *const trb = TRACE ? c.trace + J.troff : nullptr;

    // rows of this lane: table row offsets (pad row past m), H - open of
    // the left boundary column, E = -inf, statistics 0
    int qo[R], HL[R], E[R];
    int HS[STATS ? R : 1], ES[STATS ? R : 1];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = t * R + r;
        qo[r] = (i < m ? (int)s1[i] : msize) * msize;
        HL[r] = (zero_left ? 0 : -open - i * gap) - open;
        E[r] = kNeg;
        if (STATS) { HS[r] = 0; ES[r] = 0; }
    }
    // diagonal of row t*R at column 0: the left boundary of the row above
    // (the corner H[0][0] = 0 for row 0)
    const int i0 = t * R - 1;
    int prevRH = (i0 < 0 || zero_left ? 0 : -open - i0 * gap) - open;
    int prevRS = 0;
    int botH = kNeg, botF = kNeg, botS = 0, botFS = 0;

    // end tracking: SW best (value, column, row, stats); NW/SG the cell
    // (m-1, n-1); SG the last column over rows < m-1 and the last row
    int best = kNeg, bj = 0, bi = 0, bS = 0;
    int vmin = 0, vmax = 0, vminh = 0, vmaxl = 0;
    const int nvalid = min(max(m - t * R, 0), R);  // rows of this lane inside the query
    int colv = kNeg, colr = 0, colS = 0;
    int rowv = kNeg, rowj = 0, rowS = 0;
    int lastv = kNeg, lastS = 0;
    const int tm = (m - 1) / R, rm = (m - 1) - tm * R;

    auto ld_t = [&](int j) -> int { return (int)s2[j < 0 ? 0 : (j >= nreal ? nreal - 1 : j)]; };
    int tq0 = ld_t(-t), tq1 = ld_t(1 - t);

    // FAST: every lane's column lies in [0, n) of its job (the steady state)
    auto step = [&](auto fastc, const int s, int &tq) __attribute__((always_inline)) {
        constexpr bool FAST = decltype(fastc)::value;
        const int j = s - t;
        const int xH = shr1(botH), xF = shr1(botF);
        const int xS = STATS ? shr1(botS) : 0, xFS = STATS ? shr1(botFS) : 0;
        if (FAST || (j >= 0 && j < n)) {
            const int tc = tq;
            int2 Wv[R];
#pragma unroll
            for (int r = 0; r < R; ++r) Wv[r] = tab[qo[r] + tc];
            int up, F, upS = 0, FS = 0;
            if (t == 0) {
                up = (zero_top ? 0 : -open - j * gap) - open;
                F = kNeg;
            } else {
                up = xH;
                F = xF;
                if (STATS) { upS = xS; FS = xFS; }
            }
            // diagonal terms of the lane's first row; each next row's are
            // taken before its registers are rewritten (no copies)
            int t1 = prevRH + Wv[0].x, dS = STATS ? prevRS + Wv[0].y : 0;
            prevRH = up;
            prevRS = upS;
            int hv[R];
            uint32_t code8 = 0;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int Fext = F - gap;
                const bool cF = up > Fext;
                F = max(up, Fext);
                const int Eext = E[r] - gap;
                const bool cE = HL[r] > Eext;
                const int En = max(HL[r], Eext);
                const int t1n = r + 1 < R ? HL[r] + Wv[r + 1].x : 0;
                int mEF = max(En, F);
                if (SW) mEF = max(mEF, 0);
                const int h = max(t1, mEF);
                const bool cD = t1 >= mEF;
                const bool cFE = F >= En;
                if (STATS) {
                    const int dSn = r + 1 < R ? HS[r] + Wv[r + 1].y : 0;
                    FS = (cF ? upS : FS) + 1;
                    ES[r] = (cE ? HS[r] : ES[r]) + 1;
                    int hs = cD ? dS : (cFE ? FS : ES[r]);
                    if (SW) hs = h > 0 ? hs : 0;
                    HS[r] = hs;
                    upS = hs;
                    dS = dSn;
                }
                if (TRACE) {
                    // the cell's direction as a 4-bit field (dp32.h)
                    int nib = (cD ? 3 : (cFE ? 2 : 1)) | (cE ? 4 : 0) | (cF ? 8 : 0);
                    if (SW && h == 0) nib &= 12;  // ZERO_MASK, sw_trace.c:145-147
                    code8 |= (uint32_t)nib << ((r & 7) * 4);
                    if ((r & 7) == 7 || r == R - 1) {
                        constexpr int RP = (R + 7) & ~7;
                        *(uint32_t *)(trb + (((int64_t)(s * T + t) * RP + (r & ~7)) >> 1)) = code8;
                        code8 = 0;
                    }
                }
                E[r] = En;
                HL[r] = h - open;
                up = HL[r];
                hv[r] = h;
                if (EXT && r < nvalid) {
                    vmin = min(vmin, min(h, min(En, F)));
                    vmax = max(vmax, h);
                    vminh = min(vminh, h);
                    if (STATS) vmaxl = max(vmaxl, HS[r] & 0xfff);
                }
                if (TAB && r < nvalid) {
                    const int i = t * R + r;
                    int64_t at = -1;
                    if (TAB == 1) at = (int64_t)i * n + j;
                    else if (i == m - 1) at = j;
                    for (int rep = 0; rep < (TAB == 2 ? 2 : 1); ++rep) {
                        if (at >= 0) {
                            c.dtab[0][at] = h;
                            if (STATS) {
                                const uint32_t w = (uint32_t)HS[r];
                                c.dtab[1][at] = (int)(w >> 22);
                                c.dtab[2][at] = (int)((w >> 12) & 0x3ff);
                                c.dtab[3][at] = (int)(w & 0xfff);
                            }
                        }
                        at = (TAB == 2 && j == n - 1) ? (int64_t)n + i : -1;
                    }
                }
                t1 = t1n;
            }
            botH = HL[R - 1];
            botF = F;
            if (STATS) { botS = HS[R - 1]; botFS = FS; }
            if (SW) {
                int sm = hv[0];
#pragma unroll
                for (int r = 1; r < R; ++r) sm = max(sm, hv[r]);
                if (sm > best) {
                    // first row of the column with the new maximum
                    int rr = 0, ss = 0;
#pragma unroll
                    for (int r = R - 1; r >= 0; --r)
                        if (hv[r] == sm) { rr = r; if (STATS) ss = HS[r]; }
                    best = sm;
                    bj = j;
                    bi = t * R + rr;
                    bS = ss;
                }
            } else {
                if (alg == PHIP_SG && t == tm) {
                    int v = 0, vs = 0;
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        if (r == rm) { v = hv[r]; if (STATS) vs = HS[r]; }
                    if (v > rowv) { rowv = v; rowj = j; rowS = vs; }
                }
                if (j == n - 1) {
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const int i = t * R + r;
                        if (i < m - 1 && hv[r] > colv) {
                            colv = hv[r];
                            colr = i;
                            if (STATS) colS = HS[r];
                        }
                        if (i == m - 1) { lastv = hv[r]; if (STATS) lastS = HS[r]; }
                    }
                }
            }
        }
        tq = ld_t(j + 2);
    };
    const int steps = wd.steps;
    // the wave's shortest job (its lane-0 values; an empty slot has n = 0)
    int nmin = n;
    for (int k = T; k < 64; k += T) nmin = min(nmin, __builtin_amdgcn_readlane(n, k));
    nmin = min(__builtin_amdgcn_readfirstlane(n), __builtin_amdgcn_readfirstlane(nmin));
    const std::integral_constant<bool, true> fastc;
    const std::integral_constant<bool, false> slowc;
    // pairs of steps whose every lane is inside its job's columns: the even
    // s with s >= T - 1 and s + 1 < nmin
    const int lo = min(T, steps), hi = max(lo, nmin - 1);
    int s = 0;
    for (; s < lo; s += 2) {
        step(slowc, s, tq0);
        step(slowc, s + 1, tq1);
    }
    for (; s < hi; s += 2) {
        step(fastc, s, tq0);
        step(fastc, s + 1, tq1);
    }
    for (; s < steps; s += 2) {
        step(slowc, s, tq0);
        step(slowc, s + 1, tq1);
    }

    // reduce over the job's lanes, then lane 0 applies the reference's end
    // rules (sw.c:150-158; sg_stats.c:252-309; sg.c:171-235)
    int score, eq, er, w = 0;
    if (SW) {
        for (int k = 1; k < T; k <<= 1) {
            const int ob = __shfl_xor(best, k, T), oj = __shfl_xor(bj, k, T);
            const int oi = __shfl_xor(bi, k, T), os = __shfl_xor(bS, k, T);
            const bool take = ob > best || (ob == best && (oj < bj || (oj == bj && oi < bi)));
            if (take) { best = ob; bj = oj; bi = oi; bS = os; }
        }
        score = best;
        eq = bi;
        er = bj;
        w = bS;
    } else {
        lastv = __shfl(lastv, tm, T);
        lastS = __shfl(lastS, tm, T);
        if (alg == PHIP_NW) {
            score = lastv;
            eq = m - 1;
            er = n - 1;
            w = lastS;
        } else {
            for (int k = 1; k < T; k <<= 1) {
                const int ov = __shfl_xor(colv, k, T), orr = __shfl_xor(colr, k, T);
                const int os = __shfl_xor(colS, k, T);
                if (ov > colv || (ov == colv && orr < colr)) { colv = ov; colr = orr; colS = os; }
            }
            rowv = __shfl(rowv, tm, T);
            rowj = __shfl(rowj, tm, T);
            rowS = __shfl(rowS, tm, T);
            if (STATS) {
                score = kNeg;
                eq = m - 1;
                er = n - 1;
                if (s1e) {
                    if (colv > score) { score = colv; eq = colr; er = n - 1; w = colS; }
                    if (lastv > score) { score = lastv; eq = m - 1; er = n - 1; w = lastS; }
                }
                if (s1e && s2e) {
                    if (rowv > score || (rowv == score && rowj < er)) {
                        score = rowv; eq = m - 1; er = rowj; w = rowS;
                    }
                } else if (s2e) {
                    if (rowv > score) { score = rowv; eq = m - 1; er = rowj; w = rowS; }
                }
                if (!s1e && !s2e) { score = lastv; eq = m - 1; er = n - 1; w = lastS; }
            } else {
                score = kNeg;
                eq = m;
                er = n;
                if (s1e && colv > score) { score = colv; eq = colr; er = n - 1; }
                if (s1e && s2e) {
                    if (rowv > score || (rowv == score && rowj < er)) { score = rowv; eq = m - 1; er = rowj; }
                } else if (s2e) {
                    if (rowv > score) { score = rowv; eq = m - 1; er = rowj; }
                }
                if ((s1e && lastv > score) || (!s1e && !s2e)) { score = lastv; eq = m - 1; er = n - 1; }
            }
        }
    }
    if (EXT) {
        for (int k = 1; k < T; k <<= 1) {
            vmin = min(vmin, __shfl_xor(vmin, k, T));
            vmax = max(vmax, __shfl_xor(vmax, k, T));
            vminh = min(vminh, __shfl_xor(vminh, k, T));
            vmaxl = max(vmaxl, __shfl_xor(vmaxl, k, T));
        }
        // the DP's boundary row and column (k_simple's first row and WH)
        if (!zero_top) vmin = min(vmin, -open - (nreal - 1) * gap);
        if (!zero_left) vmin = min(vmin, -open - (m - 1) * gap);
    }
    if (jv && t == 0) {
        phip_out_t o;
        memset(&o, 0, sizeof(o));
        o.score = score;
        o.end_query = eq;
        o.end_ref = er;
        if (STATS) {
            o.matches = (int)((uint32_t)w >> 22);
            o.similar = (w >> 12) & 0x3ff;
            o.length = w & 0xfff;
        }
        o.status = J.status;
        o.min_val = vmin;
        o.max_val = vmax;
        o.min_h = vminh;
        o.max_len = vmaxl;
        c.out[pr] = o;
    }
}

// LDS: (size+1) x size entries {substitution + open, statistics increment},
// the last row for query rows past the end.
template <int KIND, int ALG, bool EXT, int TAB>
__global__ void __launch_bounds__(256) k_dp32(Dp32Cfg c)
{
    extern __shared__ int2 tab[];
    const int msize = c.msize, open = c.open;
    for (int x = threadIdx.x; x < (msize + 1) * msize; x += blockDim.x) {
        const int qa = x / msize, tb = x - qa * msize;
        if (qa < msize) {
            const int sub = c.mat[qa * msize + tb];
            tab[x] = make_int2(sub + open, ((qa == tb) << 22) | ((sub > 0) << 12) | 1);
        } else {
            tab[x] = make_int2(kPadSub + open, 0);
        }
    }
    __syncthreads();
    // the wave's queue entry (a scalar: its loop bounds are the wave's own)
    const int w = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (w >= c.nwaves) return;  // whole wave
    const Dp32Wave wd = c.waves[w];
    const int lane = threadIdx.x & 63;
    switch (wd.R) {
    case 4: dp32_run<4, KIND, ALG, EXT, TAB>(c, wd, tab, lane); break;
    case 8: dp32_run<8, KIND, ALG, EXT, TAB>(c, wd, tab, lane); break;
    case 12: dp32_run<12, KIND, ALG, EXT, TAB>(c, wd, tab, lane); break;
    case 16: dp32_run<16, KIND, ALG, EXT, TAB>(c, wd, tab, lane); break;
    case 20: dp32_run<20, KIND, ALG, EXT, TAB>(c, wd, tab, lane); break;
    default:
        // the finer shapes of per-call alignments (no traceback)
        if constexpr (KIND != PHIP_TRACE) {
            switch (wd.R) {
            case 2: dp32_run<2, KIND, ALG, EXT, TAB>(c, wd, tab, lane); break;
            case 3: dp32_run<3, KIND, ALG, EXT, TAB>(c, wd, tab, lane); break;
            case 5: dp32_run<5, KIND, ALG, EXT, TAB>(c, wd, tab, lane); break;
            case 6: dp32_run<6, KIND, ALG, EXT, TAB>(c, wd, tab, lane); break;
            default: dp32_run<10, KIND, ALG, EXT, TAB>(c, wd, tab, lane); break;
            }
        }
        break;
    }
}

// ------------------------------------------------------------------ CIGAR

struct CigarCfg {
    const uint8_t *qraw;   // residues as given, upper-cased
    const uint8_t *traw;
    const int64_t *qoff, *toff;
    const int32_t *qlen, *tlen;
    const int8_t *trace;
    const int64_t *troff;
    const int32_t *tlay;   // 0: row-major; else T | R << 8 (wavefront order)
    const phip_out_t *out;
    int is_sg, npairs;
    uint32_t *ops;         // per pair: qlen + tlen + 2 entries at opoff
    const int64_t *opoff;
    int32_t *nops, *beg;   // ops per pair (-1: no CIGAR); beg_query, beg_ref
};

__global__ void __launch_bounds__(256) k_cigar(CigarCfg c)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= c.npairs) return;
    const int lena = c.qlen[p], lenb = c.tlen[p];
    const uint8_t *A = c.qraw + c.qoff[p], *B = c.traw + c.toff[p];
    const int8_t *tr = c.trace + c.troff[p];
    const int lay = c.tlay[p];
    uint32_t *ops = c.ops + c.opoff[p];
    const int cap = lena + lenb + 2;
    int nops = 0;
    uint32_t mat = 0, mis = 0, del = 0, ins = 0;
    auto push = [&](uint32_t cnt, uint32_t op) { if (nops < cap) ops[nops++] = (cnt << 4) | op; };
    // '=' 7, 'X' 8, 'I' 1, 'D' 2; I and D are emitted swapped, as the
    // reference does (cigar_template.c:38-56)
    auto flush = [&]() {
        if (mat) push(mat, 7);
        else if (mis) push(mis, 8);
        else if (del) push(del, 1);
        else if (ins) push(ins, 2);
        mat = mis = del = ins = 0;
    };
    auto code_at = [&](int i, int j) -> int {
        if (!lay) return tr[(int64_t)i * lenb + j];
        return dp32_trace_code((const uint8_t *)tr, lay, i, j);
    };
    const int eq = c.out[p].end_query, er = c.out[p].end_ref;
    int i = eq, j = er;
    int where = 4;  // DIAG
    bool ok = true;
    if (c.is_sg) {
        if (eq + 1 == lena) ins += (uint32_t)max(lenb - 1 - j, 0);
        else if (er + 1 == lenb) del += (uint32_t)max(lena - 1 - i, 0);
        else ok = false;
    }
    while (ok && (i >= 0 || j >= 0)) {
        if (i < 0) {
            if (ins == 0) flush();
            ins += (uint32_t)(j + 1);
            j = -1;
            break;
        }
        if (j < 0) {
            if (del == 0) flush();
            del += (uint32_t)(i + 1);
            i = -1;
            break;
        }
        const int code = code_at(i, j);
        if (where == 4) {
            if (code & 4) {
                if (A[i] == B[j]) {
                    if (mat == 0) flush();
                    mat += 1;
                } else {
                    if (mis == 0) flush();
                    mis += 1;
                }
                --i;
                --j;
            } else if (code & 1) {
                where = 1;
            } else if (code & 2) {
                where = 2;
            } else {
                break;
            }
        } else if (where == 1) {
            if (ins == 0) flush();
            ins += 1;
            --j;
            if (code & 8) where = 4;
            else if (!(code & 16)) ok = false;
        } else {
            if (del == 0) flush();
            del += 1;
            --i;
            if (code & 32) where = 4;
            else if (!(code & 64)) ok = false;
        }
    }
    if (!ok) {
        c.nops[p] = -1;
        return;
    }
    flush();
    for (int a = 0, z = nops - 1; a < z; ++a, --z) {
        const uint32_t x = ops[a];
        ops[a] = ops[z];
        ops[z] = x;
    }
    c.nops[p] = nops;
    c.beg[2 * p] = i + 1;
    c.beg[2 * p + 1] = j + 1;
}

// ------------------------------------------------------------------ host

namespace {

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

// lanes and rows per lane of a job: the fewest lanes whose band of at most
// 20 rows per lane covers the query, rows rounded up to a multiple of 4
// fine: any R of kFineRs (no traceback, whose stores take 4 rows at a time)
const int kFineRs[] = {2, 3, 4, 5, 6, 8, 10, 12, 16, 20};

bool shape_calc(int m, int &T, int &R, int min_lanes, bool fine)
{
    for (int tt : {16, 32, 64}) {
        if (tt < min_lanes) continue;
        if (m <= tt * 20) {
            T = tt;
            const int need = (m + tt - 1) / tt;
            R = std::max(4, (need + 3) & ~3);
            if (fine)
                for (int r : kFineRs)
                    if (r >= need) { R = r; break; }
            return true;
        }
    }
    return false;
}

// shape_calc, with the default lanes and rows looked up per query length
// (the planners ask once per pair of large batches)
bool shape(int m, int &T, int &R, int min_lanes = 16, bool fine = false)
{
    constexpr int kMaxM = 64 * 20;
    struct Table {
        int16_t t[kMaxM + 1], r[kMaxM + 1];
        Table()
        {
            for (int mm = 0; mm <= kMaxM; ++mm) {
                int tt = 0, rr = 0;
                if (mm < 1 || !shape_calc(mm, tt, rr, 16, false)) tt = rr = 0;
                t[mm] = (int16_t)tt;
                r[mm] = (int16_t)rr;
            }
        }
    };
    if (min_lanes != 16 || fine || m < 1 || m > kMaxM) return shape_calc(m, T, R, min_lanes, fine);
    static const Table tab;
    if (!tab.t[m]) return false;
    T = tab.t[m];
    R = tab.r[m];
    return true;
}

}  // namespace

bool dp32_eligible(const phip_matrix_t &mx, int alg, int kind, int check, int open, int gap, int m,
                   int n)
{
    if (kind == PHIP_SCORE || mx.is_pssm || mx.size > 32) return false;
    int T, R;
    if (!shape(m, T, R)) return false;
    if (kind == PHIP_STATS && (std::min(m, n) > 1023 || m + n > 4095)) return false;
    // with a check the kernel computes the extrema (EXT)
    return true;
}

bool dp32_window_holds(const phip_matrix_t &mx, int alg, int kind, int check, int open, int gap,
                       int m, int n)
{
    if (check == 0) return true;
    const int64_t wmax = (1 << (check - 1)) - 1, wmin = -wmax - 1;
    const int64_t neg = (-open < mx.min ? wmin + open : wmin - (int64_t)mx.min) + 1;
    const int64_t pos = wmax - (int64_t)mx.max - 1;
    // a cell's length count is at most m + n, its matches and similar no more
    if (kind == PHIP_STATS && (int64_t)m + n > pos) return false;
    // SW's windows on the score are tested on the exact score
    if (alg == PHIP_SW) return true;
    // every H/E/F value lies in [lo, hi]: a cell is at least its all-gap
    // path, E/F at most `open` below an H; H at most min(m, n) best matches
    const int64_t lo = -(int64_t)4 * open - (int64_t)(m + n) * gap - std::abs(mx.min) - 64;
    const int64_t hi = (int64_t)std::min(m, n) * std::max(mx.max, 0) + mx.max + open + 64;
    return lo >= neg && hi <= pos;
}

bool dp32_eligible_one(const phip_matrix_t &mx, int alg, int kind, int check, int open, int gap,
                       int m, int n)
{
    if (mx.is_pssm || mx.size > 32) return false;
    int T, R;
    if (!shape(m, T, R, 64, kind != PHIP_TRACE)) return false;
    if (kind == PHIP_SCORE) return true;  // extrema computed when check is set
    return dp32_eligible(mx, alg, kind, check, open, gap, m, n);
}

int32_t dp32_layout(int m, int min_lanes)
{
    int T = 0, R = 0;
    if (!shape(m, T, R, min_lanes)) return 0;
    return T | (R << 8);
}

int64_t dp32_trace_bytes(int m, int n, int min_lanes)
{
    int T = 0, R = 0;
    if (!shape(m, T, R, min_lanes)) return (int64_t)m * n;
    return (((int64_t)n + T + 1) * T * dp32_rpad(R) + 1) / 2;
}

bool dp32_plan(const HostBatch &hb, const std::vector<int> &pairs, const int32_t *status,
               const int64_t *troff, Dp32Plan &P, int min_lanes, bool fine)
{
    P = Dp32Plan();
    if (pairs.empty()) return true;
    struct Key { int T, R, n, pair; };
    std::vector<Key> keys;
    keys.reserve(pairs.size());
    for (int p : pairs) {
        Key k;
        if (!shape(hb.qlen[p], k.T, k.R, min_lanes, fine)) return false;
        k.n = hb.tlen[p];
        k.pair = p;
        keys.push_back(k);
    }
    // jobs of one shape share waves; longest first so a wave's jobs end
    // together and the longest waves start first: (T, R, n) descending
    {
        const size_t N = keys.size();
        std::vector<uint32_t> k(N);
        for (size_t i = 0; i < N; ++i) {
            const int tr = keys[i].T == 64 ? 0 : keys[i].T == 32 ? 1 : 2;
            const uint32_t rank = (uint32_t)(tr * 64 + (63 - std::min(keys[i].R, 63)));
            const uint32_t nn = (uint32_t)std::min(keys[i].n, 0xFFFFFF);
            k[i] = (rank << 24) | (0xFFFFFFu - nn);
        }
        const std::vector<int> ix = radix_order(k);
        std::vector<Key> sorted(N);
        for (size_t i = 0; i < N; ++i) sorted[i] = keys[ix[i]];
        keys.swap(sorted);
    }
    std::vector<Dp32Job> jobs;
    std::vector<Dp32Wave> waves;
    jobs.reserve(keys.size());
    waves.reserve(keys.size() / 4 + 1);
    for (size_t x = 0; x < keys.size();) {
        const int per = 64 / keys[x].T;
        Dp32Wave w;
        memset(&w, 0, sizeof(w));
        w.job0 = (int)jobs.size();
        w.T = keys[x].T;
        w.R = keys[x].R;
        int maxn = 0;
        while (x < keys.size() && w.njobs < per && keys[x].T == w.T && keys[x].R == w.R) {
            Dp32Job j;
            j.pair = keys[x].pair;
            j.status = status ? status[keys[x].pair] : 0;
            j.troff = troff ? troff[keys[x].pair] : 0;
            jobs.push_back(j);
            maxn = std::max(maxn, keys[x].n);
            ++w.njobs;
            ++x;
        }
        w.steps = maxn + w.T - 1;
        waves.push_back(w);
    }
    // longest waves first
    std::stable_sort(waves.begin(), waves.end(), [](const Dp32Wave &a, const Dp32Wave &b) {
        return (int64_t)a.steps * a.R > (int64_t)b.steps * b.R;
    });
    P.njobs = (int)jobs.size();
    P.nwaves = (int)waves.size();
    P.o_jobs = 0;
    P.o_waves = al256(jobs.size() * sizeof(Dp32Job));
    P.bytes = P.o_waves + al256(waves.size() * sizeof(Dp32Wave));
    P.image.assign(P.bytes, 0);
    memcpy(P.image.data() + P.o_jobs, jobs.data(), jobs.size() * sizeof(Dp32Job));
    memcpy(P.image.data() + P.o_waves, waves.data(), waves.size() * sizeof(Dp32Wave));
    return true;
}

int dp32_launch(const DPParams &dp, const Dp32Plan &P, const char *region, hipStream_t st)
{
    if (P.nwaves <= 0) return 0;
    Dp32Cfg c;
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
    c.alg = dp.alg;
    c.sg_flags = dp.sg_flags;
    c.jobs = (const Dp32Job *)(region + P.o_jobs);
    c.waves = (const Dp32Wave *)(region + P.o_waves);
    c.nwaves = P.nwaves;
    c.trace = dp.trace;
    c.out = dp.out;
    const size_t lds = (size_t)(dp.msize + 1) * dp.msize * 8;
    const dim3 grid((P.nwaves + 3) / 4), block(256);
    typedef void (*Fn)(Dp32Cfg);
    static const Fn fns[6][3] = {
        {k_dp32<PHIP_SCORE, PHIP_SW, false, 0>, k_dp32<PHIP_SCORE, PHIP_NW, false, 0>,
         k_dp32<PHIP_SCORE, PHIP_SG, false, 0>},
        {k_dp32<PHIP_SCORE, PHIP_SW, true, 0>, k_dp32<PHIP_SCORE, PHIP_NW, true, 0>,
         k_dp32<PHIP_SCORE, PHIP_SG, true, 0>},
        {k_dp32<PHIP_STATS, PHIP_SW, false, 0>, k_dp32<PHIP_STATS, PHIP_NW, false, 0>,
         k_dp32<PHIP_STATS, PHIP_SG, false, 0>},
        {k_dp32<PHIP_TRACE, PHIP_SW, false, 0>, k_dp32<PHIP_TRACE, PHIP_NW, false, 0>,
         k_dp32<PHIP_TRACE, PHIP_SG, false, 0>},
        {k_dp32<PHIP_STATS, PHIP_SW, true, 0>, k_dp32<PHIP_STATS, PHIP_NW, true, 0>,
         k_dp32<PHIP_STATS, PHIP_SG, true, 0>},
        {k_dp32<PHIP_TRACE, PHIP_SW, true, 0>, k_dp32<PHIP_TRACE, PHIP_NW, true, 0>,
         k_dp32<PHIP_TRACE, PHIP_SG, true, 0>}};
    // tables and last row/column (one pair; extrema always computed)
    static const Fn tabs[2][2][3] = {
        {{k_dp32<PHIP_SCORE, PHIP_SW, true, 1>, k_dp32<PHIP_SCORE, PHIP_NW, true, 1>,
          k_dp32<PHIP_SCORE, PHIP_SG, true, 1>},
         {k_dp32<PHIP_STATS, PHIP_SW, true, 1>, k_dp32<PHIP_STATS, PHIP_NW, true, 1>,
          k_dp32<PHIP_STATS, PHIP_SG, true, 1>}},
        {{k_dp32<PHIP_SCORE, PHIP_SW, true, 2>, k_dp32<PHIP_SCORE, PHIP_NW, true, 2>,
          k_dp32<PHIP_SCORE, PHIP_SG, true, 2>},
         {k_dp32<PHIP_STATS, PHIP_SW, true, 2>, k_dp32<PHIP_STATS, PHIP_NW, true, 2>,
          k_dp32<PHIP_STATS, PHIP_SG, true, 2>}}};
    if (dp.alg < 0 || dp.alg > 2 || dp.kind < 0 || dp.kind > 2) return -1;
    if (dp.tab) {
        if (dp.tab < 1 || dp.tab > 2 || dp.kind == PHIP_TRACE) return -1;
        for (int k = 0; k < 4; ++k) c.dtab[k] = dp.tabs[k];
        hipLaunchKernelGGL(tabs[dp.tab - 1][dp.kind][dp.alg], grid, block, lds, st, c);
        HIP_OK(hipGetLastError());
        return 0;
    }
    const int row = dp.kind == PHIP_SCORE ? (dp.check ? 1 : 0)
                    : dp.kind == PHIP_STATS ? (dp.check ? 4 : 2) : (dp.check ? 5 : 3);
    hipLaunchKernelGGL(fns[row][dp.alg], grid, block, lds, st, c);
    HIP_OK(hipGetLastError());
    return 0;
}

// Gather each pair's CIGAR ops (nops[p] of them at ops + opoff[p]; none if
// nops[p] < 0) to dst + coff[p]: one wave per pair.
__global__ void __launch_bounds__(256) k_cigar_compact(const uint32_t *ops, const int64_t *opoff,
                                                       const int32_t *nops, const int64_t *coff,
                                                       int np, uint32_t *dst)
{
    const int p = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (p >= np) return;
    const int n = nops[p];
    const uint32_t *src = ops + opoff[p];
    uint32_t *out = dst + coff[p];
    for (int i = lane; i < n; i += 64) out[i] = src[i];
}

int cigar_compact_launch(const uint32_t *ops, const int64_t *opoff, const int32_t *nops,
                         const int64_t *coff, int np, uint32_t *dst, hipStream_t st)
{
    if (np <= 0) return 0;
    hipLaunchKernelGGL(k_cigar_compact, dim3((np + 3) / 4), dim3(256), 0, st, ops, opoff, nops, coff,
                       np, dst);
    HIP_OK(hipGetLastError());
    return 0;
}

int cigar_launch(const CigarArgs &a, hipStream_t st)
{
    if (a.npairs <= 0) return 0;
    CigarCfg c;
    c.qraw = a.qraw;
    c.traw = a.traw;
    c.qoff = a.qoff;
    c.toff = a.toff;
    c.qlen = a.qlen;
    c.tlen = a.tlen;
    c.trace = a.trace;
    c.troff = a.troff;
    c.tlay = a.tlay;
    c.out = a.out;
    c.is_sg = a.is_sg;
    c.npairs = a.npairs;
    c.ops = a.ops;
    c.opoff = a.opoff;
    c.nops = a.nops;
    c.beg = a.beg;
    hipLaunchKernelGGL(k_cigar, dim3((a.npairs + 255) / 256), dim3(256), 0, st, c);
    HIP_OK(hipGetLastError());
    return 0;
}
