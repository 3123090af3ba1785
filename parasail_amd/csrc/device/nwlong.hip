// nwlong.hip -- global alignment score (parasail_nw / nw_scan_32 semantics,
// src/nw.c:29-167) of long pairs in int32.
//
// The pair's query rows are cut into bands of 64*R rows and the bands into
// groups of LW.  A workgroup of LW waves computes one group: wave w works on
// band w of the group, its 64 lanes forming the wavefront chain of the int16
// kernels (lane t holds R rows and computes column s - t at step s; the row
// above comes from lane t-1 by DPP), here in int32 because long global scores
// leave int16.  The LW bands run as a pipeline: band w starts LL = LK + 66 steps
// after band w-1 and takes the bottom row of band w-1 from an LDS ring; all
// waves meet at a barrier every LK steps, and every ring value a wave reads
// was written at least one barrier earlier.
//
// Groups are work items of a queue that persistent workgroups drain in
// order: the groups of one pair are consecutive items, so groups of a long
// pair run on several CUs at once, each LK steps behind the one above.  The
// last band of a group leaves its bottom row in HBM and publishes how many
// of its columns are final (a release store after the row's stores); the
// first band of the next group waits, before each LK steps, until the
// columns it will read are published (an acquire load).  A workgroup only
// waits on an item taken from the queue before its own, by a workgroup that
// is already running, so every wait ends; each wait is also bounded, and a
// wait that runs out marks the launch failed instead of hanging it.
#include <string.h>

#include <algorithm>
#include <vector>

#include "search.h"
#include "wave.h"

namespace {

constexpr int LW = 16;     // waves (bands in flight) per workgroup: 4 per SIMD
constexpr int LK = 32;     // steps between barriers
constexpr int LL = LK + 66; // lag between consecutive bands
constexpr int LRING = 512; // ring slots per band boundary (power of two)
// In the steps [P, P+LK) between two barriers band w+1's lane 0 reads (two
// steps ahead) columns up to P + LK + 1 - LL*(w+1), while band w's lane 63
// had written, before the barrier at P, the columns up to P - 64 - LL*w:
// every read is of a value written at least one barrier earlier iff
// LL >= LK + 65.  The writer runs at most LK + LL - 64 columns ahead of the
// oldest column the reader still holds, which the ring must cover.
static_assert(LL >= LK + 65, "band lag too short for the barrier spacing");
static_assert(LK + LL - 64 < LRING, "ring too small for the band lag");
static_assert(LK % 2 == 0, "steps run in pairs");
constexpr int LNEG = -(1 << 29);
constexpr int LWNEG = -(1 << 24);  // substitution value of the pad rows
constexpr int kLongR = 16;
constexpr int kGroupRows = LW * 64 * kLongR;
// bound on the polls of one wait (each sleeps ~0.2 us); legitimate waits
// are a few thousand steps of the group above, well under a second
constexpr int kSpinLimit = 1 << 23;
constexpr int kPubPhases = 4;  // barriers between progress publications (power of two)

}  // namespace

struct LongItem {
    int pair;      // index into the batch
    int g;         // group of bands
    int64_t gin;   // offset (int2) in gbuf of the row above the group, or -1
    int64_t gout;  // offset (int2) in gbuf of the group's bottom row, or -1
};

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
    const LongItem *items;
    int nitems;
    int *prog;          // per item: columns of its bottom row final in gbuf
    int *ctl;           // [0] queue head, [1] failure flag
    int2 *gbuf;
    phip_out_t *out;
};

// LDS: table of W+open ((size+1) rows: codes, then a pad row), then the
// LW-1 rings of (H-open, F).  RP: an alphabet of at most 8 codes whose
// W+open all lie in [0, 15] (DNA): each row's scores sit in one register as
// 4-bit fields, so a lookup is one bit-field extract instead of an LDS read
// (rows past the query read 0: their values are never used).
template <int R, bool RP>
__global__ void __launch_bounds__(LW * 64) k_nwlong(LongCfg c)
{
    extern __shared__ int lds[];
    __shared__ int s_item;
    const int msize = c.msize, open = c.open, gap = c.gap;
    const int tabn = (msize + 1) * msize;
    for (int x = threadIdx.x; x < tabn; x += blockDim.x)
        lds[x] = x < msize * msize ? c.mat[x] + open : LWNEG;
    int2 *const ring = (int2 *)(lds + ((tabn + 1) & ~1));
    // the wave's band (a scalar: every band-level branch and index below is
    // the wave's own)
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), t = threadIdx.x & 63;
    const int2 *ring_in = ring + (w > 0 ? w - 1 : 0) * LRING;
    int2 *ring_out = ring + (w < LW - 1 ? w : 0) * LRING;
    int *const failed = c.ctl + 1;

    for (;;) {
        if (threadIdx.x == 0) {
            int it = atomicAdd(c.ctl, 1);
            if (__hip_atomic_load(failed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) it = c.nitems;
            s_item = it;
        }
        __syncthreads();
        const int item = s_item;
        __syncthreads();
        if (item >= c.nitems) break;  // whole block
        const LongItem li = c.items[item];
        const int pr = li.pair, g = li.g;
        const int m = c.qlen[pr], n = c.tlen[pr];
        const uint8_t *s1 = c.q + c.qoff[pr];
        const uint8_t *s2 = c.t + c.toff[pr];
        const int2 *gin = li.gin >= 0 ? c.gbuf + li.gin : nullptr;
        int2 *gout = li.gout >= 0 ? c.gbuf + li.gout : nullptr;
        const int BR = 64 * R;
        const int nb = (m + BR - 1) / BR;
        const int b = g * LW + w;
        const bool live = b < nb;
        const int bm = (m - 1) / BR, rem = (m - 1) % BR, tm = rem / R, rm = rem % R;
        auto left = [&](int r) -> int { return r < 0 ? -open : -open - r * gap - open; };
        const int G = n + 63 + LL * (LW - 1);
        const int nph = (G + LK - 1) / LK;
        int score = 0;

        // band 0 of a later group: before steps [ph*LK, ph*LK+LK), the rows
        // it reads there and two steps beyond (prefetch) must be final.  The
        // count last acquired is kept: an acquire (which invalidates this
        // CU's caches) is made only when the rows needed pass it
        int seen = 0;
        auto wait_rows = [&](int ph) {
            if (gin && w == 0 && t == 0) {
                const int need = min(n, ph * LK + LK + 2);
                if (need <= seen) return;
                const int *pin = c.prog + item - 1;
                int spins = 0;
                while ((seen = __hip_atomic_load(pin, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)) < need) {
                    if (__hip_atomic_load(failed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                        ++spins > kSpinLimit) {
                        __hip_atomic_store(failed, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(8);
                }
            }
        };

        const int first_row = b * BR + t * R;
        int qo[R], HL[R], E[R];
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int row = first_row + i;
            qo[i] = (row < m ? (int)s1[row] : msize) * msize;
            if (RP) {
                int pf = 0;
                if (row < m)
                    for (int k = 0; k < msize; ++k) pf |= lds[qo[i] + k] << (4 * k);
                qo[i] = pf;
            }
            HL[i] = left(row);
            E[i] = LNEG;
        }
        int prevRH = left(first_row - 1);
        int botH = LNEG, botF = LNEG;
        // lane 0's input for column jj: this group's band above (MODE 0, the
        // LDS ring), the previous group's last band (1, HBM) or the DP's row
        // above the query (2); a template, so that the loads of a step never
        // meet at a branch
        auto ld_in = [&](auto modec, int jj) -> int2 {
            constexpr int MODE = decltype(modec)::value;
            const int jc = jj < 0 ? 0 : (jj >= n ? n - 1 : jj);
            if (MODE == 2) return make_int2(-open - jc * gap - open, LNEG);
            if (MODE == 1) return gin[jc];
            return ring_in[jc & (LRING - 1)];
        };
        auto ld_t = [&](int jj) -> int { return (int)s2[jj < 0 ? 0 : (jj >= n ? n - 1 : jj)]; };
        const int s0 = -LL * w;  // this band's step at group step 0
        const bool pub = gout && w == LW - 1 && t == 63;

        // FAST: every lane's column lies in [0, n) (the band's steady state)
        auto step = [&](auto modec, auto fastc, const int s, int2 &bx, int &tq)
                        __attribute__((always_inline)) {
            constexpr bool FAST = decltype(fastc)::value;
            const int j = s - t;
            // lane 0 takes the row above from bx, every other lane from its
            // left neighbour
            const int rH = __builtin_amdgcn_update_dpp(bx.x, botH, 0x138, 0xF, 0xF, false);
            const int rF = __builtin_amdgcn_update_dpp(bx.y, botF, 0x138, 0xF, 0xF, false);
            if (FAST || (j >= 0 && j < n)) {
                const int code = tq;
                int W[R];
                if (RP) {
                    const int sh = code * 4;
#pragma unroll
                    for (int i = 0; i < R; ++i) W[i] = (int)__builtin_amdgcn_ubfe((unsigned)qo[i], sh, 4);
                } else {
#pragma unroll
                    for (int i = 0; i < R; ++i) W[i] = lds[qo[i] + code];
                }
                int t1 = prevRH + W[0];
                prevRH = rH;
                int up = rH, F = rF;
                // the next row's diagonal term is taken before HL[i] is
                // rewritten, so every HL[i] stays in its register
#pragma unroll
                for (int i = 0; i < R; ++i) {
                    E[i] = max(E[i] - gap, HL[i]);
                    const int t1n = i + 1 < R ? HL[i] + W[i + 1] : 0;
                    F = max(F - gap, up);
                    const int h = max(max(t1, E[i]), F);
                    HL[i] = h - open;
                    up = HL[i];
                    t1 = t1n;
                }
                botH = HL[R - 1];
                botF = F;
                if (t == 63 && w < LW - 1) ring_out[j & (LRING - 1)] = make_int2(botH, botF);
                if (pub) gout[j] = make_int2(botH, botF);
            }
            bx = ld_in(modec, s + 2);
            tq = ld_t(j + 2);
        };
        const std::integral_constant<bool, true> fastc;
        const std::integral_constant<bool, false> slowc;
        auto run = [&](auto modec) {
            wait_rows(0);
            int2 in0 = ld_in(modec, s0), in1 = ld_in(modec, s0 + 1);
            int tc0 = ld_t(s0 - t), tc1 = ld_t(s0 + 1 - t);
            for (int ph = 0; ph < nph; ++ph) {
                if (live) {
                    const int sb = ph * LK + s0;
                    if (sb >= 63 && sb + LK <= n) {
                        for (int k = 0; k < LK; k += 2) {
                            step(modec, fastc, sb + k, in0, tc0);
                            step(modec, fastc, sb + k + 1, in1, tc1);
                        }
                    } else {
                        for (int k = 0; k < LK; k += 2) {
                            step(modec, slowc, sb + k, in0, tc0);
                            step(modec, slowc, sb + k + 1, in1, tc1);
                        }
                    }
                    if (pub && ((ph & (kPubPhases - 1)) == kPubPhases - 1 || ph + 1 == nph)) {
                        // columns of the bottom row stored so far, then
                        // published (every kPubPhases barriers: each release
                        // writes back this CU's L2)
                        const int done = min(n, max(0, (ph + 1) * LK - LL * (LW - 1) - 63));
                        __hip_atomic_store(c.prog + item, done, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
                if (ph + 1 < nph) wait_rows(ph + 1);
                lds_barrier();
            }
        };
        if (w > 0)
            run(std::integral_constant<int, 0>());
        else if (gin)
            run(std::integral_constant<int, 1>());
        else
            run(std::integral_constant<int, 2>());
        // a lane stops updating after its last column, so its rows now hold
        // column n-1: the score is H[m-1][n-1] (nw.c:155-157)
        if (b == bm && t == tm) {
#pragma unroll
            for (int i = 0; i < R; ++i)
                if (i == rm) score = HL[i] + open;
            phip_out_t o;
            memset(&o, 0, sizeof(o));
            o.score = score;
            o.end_query = m - 1;
            o.end_ref = n - 1;
            o.status = __hip_atomic_load(failed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                           ? PHIP_ST_ERR : PHIP_ST_W32;
            c.out[pr] = o;
        }
    }
}

__global__ void k_nwlong_zero(int *p, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// A failed launch (a wait ran out) leaves every long pair marked.
__global__ void k_nwlong_check(const int *ctl, const LongItem *items, int nitems, phip_out_t *out)
{
    if (!ctl[1]) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nitems; i += gridDim.x * blockDim.x)
        out[items[i].pair].status = PHIP_ST_ERR;
}

// ------------------------------------------------------------------ host

namespace {

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

bool nwlong_eligible(const phip_matrix_t &mx, int alg, int kind, int check, int m, int n)
{
    // the table and the LW-1 rings must fit the 64 KiB of LDS a workgroup gets
    return alg == PHIP_NW && kind == PHIP_SCORE && !check && !mx.is_pssm && mx.size <= 24 &&
           m >= 1024 && n >= 64 && (int64_t)m * n >= ((int64_t)1 << 24);
}

bool nwlong_plan(const HostBatch &hb, const std::vector<int> &pairs, int open, LongPlan &lp)
{
    lp = LongPlan();
    if (pairs.empty()) return true;
    // scores in registers when every W + open fits a 4-bit field
    lp.rp = hb.mx->size <= 8;
    for (int x = 0; lp.rp && x < hb.mx->size * hb.mx->size; ++x) {
        const int v = hb.mx->values[x] + open;
        if (v < 0 || v > 15) lp.rp = false;
    }
    const int fill = LL * (LW - 1) + 64;
    // longest chain first: a pair's last group ends about n + groups * fill
    // steps after its first starts
    std::vector<int> ord(pairs);
    auto groups = [&](int i) { return (hb.qlen[i] + kGroupRows - 1) / kGroupRows; };
    auto chain = [&](int i) { return (int64_t)hb.tlen[i] + (int64_t)groups(i) * fill; };
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return chain(x) > chain(y); });
    std::vector<LongItem> items;
    int64_t goff = 0;
    for (int i : ord) {
        const int ng = groups(i);
        int64_t prev = -1;
        for (int g = 0; g < ng; ++g) {
            LongItem it;
            it.pair = i;
            it.g = g;
            it.gin = prev;
            it.gout = -1;
            if (g + 1 < ng) {
                it.gout = goff;
                goff += (int64_t)hb.tlen[i] + 64;
            }
            prev = it.gout;
            items.push_back(it);
        }
    }
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    const int tabn = (hb.mx->size + 1) * hb.mx->size;
    lp.lds = (size_t)((tabn + 1) & ~1) * 4 + (size_t)(LW - 1) * LRING * 8;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, lp.rp ? k_nwlong<kLongR, true>
                                                                     : k_nwlong<kLongR, false>,
                                                     LW * 64, lp.lds) !=
            hipSuccess || per_cu <= 0)
        per_cu = 1;
    lp.nitems = (int)items.size();
    lp.grid = std::min(lp.nitems, cus * per_cu);
    size_t off = 0;
    auto slot = [&](size_t bytes) { size_t o = off; off += al256(bytes ? bytes : 1); return o; };
    lp.o_items = slot(items.size() * sizeof(LongItem));
    lp.o_prog = slot((size_t)lp.nitems * 4 + 64);
    lp.o_gbuf = slot((size_t)goff * 8);
    lp.bytes = off;
    lp.items.resize(items.size() * sizeof(LongItem));
    memcpy(lp.items.data(), items.data(), lp.items.size());
    return true;
}

int nwlong_launch(const DPParams &dp, const LongPlan &lp, char *region, hipStream_t st)
{
    if (lp.nitems <= 0) return 0;
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
    c.items = (const LongItem *)(region + lp.o_items);
    c.nitems = lp.nitems;
    c.prog = (int *)(region + lp.o_prog);
    c.ctl = c.prog + lp.nitems;
    c.gbuf = (int2 *)(region + lp.o_gbuf);
    c.out = dp.out;
    hipLaunchKernelGGL(k_nwlong_zero, dim3(std::min((lp.nitems + 2 + 255) / 256, 1024)), dim3(256), 0,
                       st, c.prog, lp.nitems + 2);
    if (lp.rp) hipLaunchKernelGGL((k_nwlong<kLongR, true>), dim3(lp.grid), dim3(LW * 64), lp.lds, st, c);
    else hipLaunchKernelGGL((k_nwlong<kLongR, false>), dim3(lp.grid), dim3(LW * 64), lp.lds, st, c);
    hipLaunchKernelGGL(k_nwlong_check, dim3(std::min((lp.nitems + 255) / 256, 1024)), dim3(256), 0, st,
                       c.ctl, c.items, lp.nitems, c.out);
    HIP_OK(hipGetLastError());
    return 0;
}
