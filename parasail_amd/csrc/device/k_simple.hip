// k_simple.hip -- exact int32 DP, one lane per pair.
//
// The general-purpose tier: any algorithm (SW/NW/SG with all four end
// flags), any result kind (score, _stats, _trace), square matrix or PSSM,
// any length.  Row state lives in HBM scratch, interleaved across the lanes
// of the launch so that lane k of a wave touches address base + j*stride + k
// (coalesced).  It is the int32 retry tier behind the packed-int16 wavefront
// kernels and the tier for _stats/_trace; it restates the recurrences of
// src/sw.c:131-158, src/sg.c:147-235, src/sw_stats.c:204-261,
// src/sg_stats.c:227-309 and src/sg_trace.c:144-150 cell for cell.
#include <string.h>

#include "common.h"

namespace {

struct Rows {
    int *H, *F, *HM, *HS, *HL, *FM, *FS, *FL;
};

__device__ __forceinline__ int &at(int *base, int j, int stride, int slot)
{
    return base[(size_t)j * stride + slot];
}

}  // namespace

// Pair selection: pairs first .. first+count-1, or, when `list` is given,
// entries slot, slot+stride, ... of list[0 .. min(*dcount, cap)) (the
// pairs another tier handed over).
__device__ void simple_pair(const DPParams &p, int pr, int *scratch, int stride, int slot, int maxn);

extern "C" __global__ void __launch_bounds__(256)
k_simple(DPParams p, int first, int count, const int32_t *list, const int *dcount, int *scratch,
         int stride, int maxn)
{
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= stride) return;
    if (!list) {
        if (slot < count) simple_pair(p, first + slot, scratch, stride, slot, maxn);
        return;
    }
    const int n = dcount ? imin(*dcount, count) : count;
    for (int k = slot; k < n; k += stride) simple_pair(p, list[k], scratch, stride, slot, maxn);
}

__device__ void simple_pair(const DPParams &p, int pr, int *scratch, int stride, int slot, int maxn)
{
    const int alg = p.alg, kind = p.kind;
    const int s1_beg = (alg == PHIP_SG) && (p.sg_flags & PHIP_S1_BEG);
    const int s1_end = (alg == PHIP_SG) && (p.sg_flags & PHIP_S1_END);
    const int s2_beg = (alg == PHIP_SG) && (p.sg_flags & PHIP_S2_BEG);
    const int s2_end = (alg == PHIP_SG) && (p.sg_flags & PHIP_S2_END);
    const int open = p.open, gap = p.gap;
    const int m = p.qlen ? p.qlen[pr] : p.qlen1, n = p.tlen[pr];
    const uint8_t *s1 = p.q ? p.q + (p.qoff ? p.qoff[pr] : 0) : nullptr;
    const uint8_t *s2 = p.t + p.toff[pr];
    const int *mat = p.matrix;
    const int msize = p.msize;
    const bool stats = kind == PHIP_STATS, trace = kind == PHIP_TRACE;
    int8_t *HT = trace ? p.trace + p.trace_off[pr] : nullptr;
    const int *mapper = p.mapper;
    auto code = [mapper](const uint8_t *s, int k) -> int { return mapper ? mapper[s[k]] : s[k]; };
    if (n > maxn || n <= 0 || m <= 0) {
        // outside the row scratch this launch was sized for
        phip_out_t o;
        memset(&o, 0, sizeof(o));
        o.status = PHIP_ST_ERR;
        p.out[pr] = o;
        return;
    }

    const size_t plane = (size_t)(maxn + 1) * stride;
    Rows r;
    r.H = scratch;
    r.F = scratch + plane;
    r.HM = scratch + 2 * plane;
    r.HS = scratch + 3 * plane;
    r.HL = scratch + 4 * plane;
    r.FM = scratch + 5 * plane;
    r.FS = scratch + 6 * plane;
    r.FL = scratch + 7 * plane;

    int score = PHIP_NEG_INF32, end_query = m, end_ref = n;
    int matches = 0, similar = 0, length = 0;
    if (alg == PHIP_SG && stats) { end_query = m - 1; end_ref = n - 1; }
    int vmin = 0, vmax = 0;  // exact extrema over H/E/F (check mode)
    int vminh = 0, vmaxl = 0;  // smallest cell H, largest cell H length

    // first row
    at(r.H, 0, stride, slot) = 0;
    at(r.F, 0, stride, slot) = PHIP_NEG_INF32;
    for (int j = 1; j <= n; ++j) {
        int h = (alg == PHIP_SW || s2_beg) ? 0 : -open - (j - 1) * gap;
        at(r.H, j, stride, slot) = h;
        at(r.F, j, stride, slot) = PHIP_NEG_INF32;
        vmin = imin(vmin, h);
        if (stats) {
            at(r.HM, j, stride, slot) = 0; at(r.HS, j, stride, slot) = 0;
            at(r.HL, j, stride, slot) = 0; at(r.FM, j, stride, slot) = 0;
            at(r.FS, j, stride, slot) = 0; at(r.FL, j, stride, slot) = 0;
        }
    }
    if (stats) {
        at(r.HM, 0, stride, slot) = 0; at(r.HS, 0, stride, slot) = 0;
        at(r.HL, 0, stride, slot) = 0;
    }

    for (int i = 1; i <= m; ++i) {
        const int qi = s1 ? code(s1, i - 1) : 0;
        const int *matrow = p.is_pssm ? mat + (size_t)msize * (i - 1)
                                      : mat + (size_t)msize * qi;
        int NH = at(r.H, 0, stride, slot);
        int NHM = 0, NHS = 0, NHL = 0;
        int WH = (alg == PHIP_SW || s1_beg) ? 0 : -open - (i - 1) * gap;
        vmin = imin(vmin, WH);
        int WHM = 0, WHS = 0, WHL = 0;
        int E = PHIP_NEG_INF32, EM = 0, ES = 0, EL = 0;
        at(r.H, 0, stride, slot) = WH;
        for (int j = 1; j <= n; ++j) {
            const int NWH = NH, NWM = NHM, NWS = NHS, NWL = NHL;
            NH = at(r.H, j, stride, slot);
            if (stats) {
                NHM = at(r.HM, j, stride, slot);
                NHS = at(r.HS, j, stride, slot);
                NHL = at(r.HL, j, stride, slot);
            }
            const int Fj = at(r.F, j, stride, slot);
            const int F_opn = NH - open, F_ext = Fj - gap;
            const int F = imax(F_opn, F_ext);
            const int E_opn = WH - open, E_ext = E - gap;
            E = imax(E_opn, E_ext);
            const int tj = code(s2, j - 1);
            const int sub = matrow[tj];
            const int H_dag = NWH + sub;
            int H = imax(H_dag, E);
            H = imax(H, F);
            if (alg == PHIP_SW) H = imax(H, 0);
            if (stats) {
                int fm = at(r.FM, j, stride, slot), fs = at(r.FS, j, stride, slot),
                    fl = at(r.FL, j, stride, slot);
                if (F_opn > F_ext) { fm = NHM; fs = NHS; fl = NHL; }
                fl += 1;
                at(r.FM, j, stride, slot) = fm;
                at(r.FS, j, stride, slot) = fs;
                at(r.FL, j, stride, slot) = fl;
                if (E_opn > E_ext) { EM = WHM; ES = WHS; EL = WHL; }
                EL += 1;
                if (H == H_dag) {
                    WHM = NWM + (qi == tj);
                    WHS = NWS + (sub > 0);
                    WHL = NWL + 1;
                } else if (H == F) {
                    WHM = fm; WHS = fs; WHL = fl;
                } else {
                    WHM = EM; WHS = ES; WHL = EL;
                }
                if (alg == PHIP_SW && H <= 0) { WHM = WHS = WHL = 0; }
                at(r.HM, j, stride, slot) = WHM;
                at(r.HS, j, stride, slot) = WHS;
                at(r.HL, j, stride, slot) = WHL;
            } else if (trace) {
                int code = (F_opn > F_ext) ? 32 : 64;
                code |= (E_opn > E_ext) ? 8 : 16;
                code |= (H == H_dag) ? 4 : (H == F) ? 2 : 1;
                if (alg == PHIP_SW && H == 0) code &= 120;
                HT[(size_t)(i - 1) * n + (j - 1)] = (int8_t)code;
            }
            if (p.tab) {
                // one-pair launch: the DP table (src/sw.c:159-170) or its
                // last row, then its last column
                int64_t at = -1;
                if (p.tab == 1) at = (int64_t)(i - 1) * n + (j - 1);
                else if (i == m) at = j - 1;
                for (int rep = 0; rep < (p.tab == 2 ? 2 : 1); ++rep) {
                    if (at >= 0) {
                        p.tabs[0][at] = H;
                        if (stats) {
                            p.tabs[1][at] = WHM;
                            p.tabs[2][at] = WHS;
                            p.tabs[3][at] = WHL;
                        }
                    }
                    at = (p.tab == 2 && j == n) ? (int64_t)n + (i - 1) : -1;
                }
            }
            if (p.check) {
                vmin = imin(vmin, imin(H, imin(E, F)));
                vmax = imax(vmax, H);
                vminh = imin(vminh, H);
                if (stats) vmaxl = imax(vmaxl, WHL);
            }
            WH = H;
            at(r.H, j, stride, slot) = H;
            at(r.F, j, stride, slot) = F;

            if (alg == PHIP_SW) {
                if (H > score || (H == score && j - 1 < end_ref)) {
                    score = H; end_query = i - 1; end_ref = j - 1;
                    matches = WHM; similar = WHS; length = WHL;
                }
            } else if (alg == PHIP_SG && i == m && !stats) {
                if (s1_end && s2_end) {
                    if (H > score || (H == score && j - 1 < end_ref)) {
                        score = H; end_query = m - 1; end_ref = j - 1;
                    }
                } else if (s2_end) {
                    if (H > score) { score = H; end_query = m - 1; end_ref = j - 1; }
                }
            }
        }
        if (alg == PHIP_SG) {
            if (stats) {
                if (s1_end && WH > score) {
                    score = WH; end_query = i - 1; end_ref = n - 1;
                    matches = WHM; similar = WHS; length = WHL;
                }
            } else if (i < m) {
                if (s1_end && WH > score) { score = WH; end_query = i - 1; end_ref = n - 1; }
            } else if ((s1_end && WH > score) || (!s1_end && !s2_end)) {
                score = WH; end_query = m - 1; end_ref = n - 1;
            }
        }
    }
    if (alg == PHIP_NW) {
        score = at(r.H, n, stride, slot);
        end_query = m - 1; end_ref = n - 1;
        if (stats) {
            matches = at(r.HM, n, stride, slot);
            similar = at(r.HS, n, stride, slot);
            length = at(r.HL, n, stride, slot);
        }
    } else if (alg == PHIP_SG && stats) {
        if (s1_end && s2_end) {
            for (int j = 1; j <= n; ++j) {
                int h = at(r.H, j, stride, slot);
                if (h > score || (h == score && j - 1 < end_ref)) {
                    score = h; end_query = m - 1; end_ref = j - 1;
                    matches = at(r.HM, j, stride, slot);
                    similar = at(r.HS, j, stride, slot);
                    length = at(r.HL, j, stride, slot);
                }
            }
        } else if (s2_end) {
            for (int j = 1; j <= n; ++j) {
                int h = at(r.H, j, stride, slot);
                if (h > score) {
                    score = h; end_query = m - 1; end_ref = j - 1;
                    matches = at(r.HM, j, stride, slot);
                    similar = at(r.HS, j, stride, slot);
                    length = at(r.HL, j, stride, slot);
                }
            }
        }
        if (!s1_end && !s2_end) {
            score = at(r.H, n, stride, slot);
            matches = at(r.HM, n, stride, slot);
            similar = at(r.HS, n, stride, slot);
            length = at(r.HL, n, stride, slot);
            end_query = m - 1; end_ref = n - 1;
        }
    }
    phip_out_t o;
    o.score = score;
    o.end_query = end_query;
    o.end_ref = end_ref;
    o.matches = stats ? matches : 0;
    o.similar = stats ? similar : 0;
    o.length = stats ? length : 0;
    o.status = PHIP_ST_W32;
    o.min_val = vmin;
    o.max_val = vmax;
    o.min_h = vminh;
    o.max_len = vmaxl;
    p.out[pr] = o;
}
