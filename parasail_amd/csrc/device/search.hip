// search.hip -- one query against an HBM-resident target database (SW,
// top-k hits by score, then target index), and the serving loop that keeps
// many such requests in flight on one GPU.
//
// The reference searches a database by looping over targets with OpenMP
// around parasail_sw_striped_profile_* (tests/test_openmp.c:353-369); here
// the database is resident in HBM, packed once into two-target jobs ordered
// by length, and each request is one query profile plus kernel launches.
//
// Kernel layout (k_search): as the pair wavefront kernel (wave.hip) -- two
// targets per job in the two int16 halves of every VGPR, a chain of T lanes
// per job, R query rows per lane -- but all jobs share the query, so the
// substitution scores come from a query profile (code-major, W per query
// row) in LDS or L2: a lane reads its R rows for the column's residue with
// R/4 8-byte loads per half and interleaves the halves with one v_perm_b32
// per row.  Values are kept clamped at 0 (H, max(H - open, 0), E, F), so
// the local clamp costs no instruction of its own.
#include <pthread.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include <algorithm>
#include <deque>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "pk.h"
#include "search.h"

namespace {

constexpr int kSnapRows = 20;  // largest R; LDS snapshot rows per half
constexpr int kThreads = 512;  // threads per search block (8 waves)

}  // namespace

// Two database targets sharing one chain (A low half, B high half).
struct DbJob {
    int tgt[2];      // target index in the caller's order, -1 for an empty half
    int n[2];
    int64_t ccoff;   // column stream (uint16 cA | cB << 8): T pads, width, pads
    int64_t bboff;   // band row (int2), bbw entries, rewritten in place per band
};

// One queue entry of the database: 64/T jobs of similar length.
struct DbWave {
    int job0, njobs, T, width;
    int bbw, pad0, pad1, pad2;
};

struct SearchCfg {
    const DbJob *jobs;
    const DbWave *waves;
    int w0, nw;               // this launch's queue entries [w0, w0 + nw)
    int *next_wave;
    const uint16_t *colcode;
    const short *prof;        // (size+1) code rows of mp int16 (W; W_NEG pads)
    int mp, profn;            // rows per code; profile size in dwords
    int nb16, nb64;           // bands for T = 16 and T = 64 chains
    int2 *bandbuf;
    const int2 *toprow;       // (H, F) of DP row 0 for SW, clamped: (0, 0)
    phip_out_t *out;          // per target
    int32_t *redo;
    int *nredo;
    int ovf_limit;
    int m;                    // query rows
    int open, gap;
};

namespace {

// All bands of one queue entry, R rows per lane (R a multiple of 4).
template <int R, bool LDSP>
__device__ __forceinline__ void search_wave(const SearchCfg &c, const DbWave &wd, const short *prof,
                                            int lane, short *snp)
{
    const int T = wd.T;
    const int t = lane & (T - 1);
    const int jslot = T == 64 ? 0 : lane >> 4;
    const bool live = jslot < wd.njobs;
    const DbJob J = c.jobs[wd.job0 + (live ? jslot : 0)];
    const int open = c.open, gap = c.gap;
    const int openv = pk::pack(open, open), gapv = pk::pack(gap, gap);
    const int width = wd.width;
    const int steps = width + T - 1;
    const int nb = T == 64 ? c.nb64 : c.nb16;
    const int mp = c.mp;

    int cS[2] = {-1, -1}, cC[2] = {0, 0}, cR[2] = {0, 0};
    for (int b = 0; b < nb; ++b) {
        const int r0 = b * T * R + t * R;
        // every value is clamped at 0 (values below 0 never reach a local
        // score): Hs = H, HL = max(H - open, 0), E and F >= 0, so
        // H = max(diag + W, E, F) needs no clamp of its own
        int Hs[R], HL[R], E[R];
#pragma unroll
        for (int i = 0; i < R; ++i) {
            Hs[i] = 0;  // column 0 of SW: H = 0
            HL[i] = 0;
            E[i] = 0;
        }
        int prevRH = 0;
        int botH = 0, botF = 0;
        int best = (int)0xFFFFFFFF;
        int bcolA = 0, bcolB = 0;
        const uint16_t *ccl = c.colcode + J.ccoff + (T - t);
        const int2 *bbs = (b == 0 || !live) ? c.toprow : c.bandbuf + J.bboff;
        const bool wr_lane = t == T - 1 && live && b < nb - 1;
        int2 *const bbr = c.bandbuf + J.bboff;
        const short *prow = prof + r0;

        // FAST: every lane's column lies in [0, width) (the steady state)
        auto step = [&](auto fastc, const int sv, uint32_t &cx, int2 &bx) __attribute__((always_inline)) {
            constexpr bool FAST = decltype(fastc)::value;
            const int j = sv - t;
            const int xH = pk::shr1(botH);
            const int xF = pk::shr1(botF);
            if (FAST || (j >= 0 && j < width)) {
                const uint2 *pa = (const uint2 *)(prow + (cx & 0xff) * mp);
                const uint2 *pb = (const uint2 *)(prow + (cx >> 8) * mp);
                int W[R];
#pragma unroll
                for (int k = 0; k < R / 4; ++k) {
                    const uint2 a = pa[k], bq = pb[k];
                    W[4 * k + 0] = (int)((a.x & 0xffffu) | (bq.x << 16));
                    W[4 * k + 1] = (int)((a.x >> 16) | (bq.x & 0xffff0000u));
                    W[4 * k + 2] = (int)((a.y & 0xffffu) | (bq.y << 16));
                    W[4 * k + 3] = (int)((a.y >> 16) | (bq.y & 0xffff0000u));
                }
                const int rH = t == 0 ? bx.x : xH;
                const int rF = t == 0 ? bx.y : xF;
                int t1 = pk::add(prevRH, W[0]);
                prevRH = rH;
                int up = pk::subu(rH, openv), F = rF;
                int cm = 0;
#pragma unroll
                for (int i = 0; i < R; ++i) {
                    E[i] = pk::max(pk::subu(E[i], gapv), HL[i]);
                    const int t1n = i + 1 < R ? pk::add(Hs[i], W[i + 1]) : 0;
                    F = pk::max(pk::subu(F, gapv), up);
                    const int h = pk::max(pk::max(t1, E[i]), F);
                    Hs[i] = h;
                    HL[i] = pk::subu(h, openv);
                    up = HL[i];
                    t1 = t1n;
                    cm = i == 0 ? h : pk::max(cm, h);
                }
                botH = Hs[R - 1];
                botF = F;
                if (wr_lane) bbr[j] = make_int2(botH, botF);
                const int nbst = pk::max(best, cm);
                if (nbst != best) {
                    const int d = nbst ^ best;
                    if (d & 0xffff) {
#pragma unroll
                        for (int i = 0; i < R; ++i) snp[i * kThreads] = (short)Hs[i];
                        bcolA = j;
                    }
                    if (d >> 16) {
#pragma unroll
                        for (int i = 0; i < R; ++i) snp[(kSnapRows + i) * kThreads] = (short)(Hs[i] >> 16);
                        bcolB = j;
                    }
                    best = nbst;
                }
            }
            cx = ccl[sv + 2];
            bx = bbs[sv + 2];
        };
        uint32_t c0 = ccl[0], c1 = ccl[1];
        int2 b0 = bbs[0], b1 = bbs[1];
        const std::integral_constant<bool, true> fastc;
        const std::integral_constant<bool, false> slowc;
        // pairs of steps whose every lane is inside the columns: the even s
        // with s >= T - 1 and s + 1 < width
        const int lo = min(T, steps), hi = max(lo, width - 1);
        int s = 0;
        for (; s < lo; s += 2) {
            step(slowc, s, c0, b0);
            step(slowc, s + 1, c1, b1);
        }
        for (; s < hi; s += 2) {
            step(fastc, s, c0, b0);
            step(fastc, s + 1, c1, b1);
        }
        for (; s < steps; s += 2) {
            step(slowc, s, c0, b0);
            step(slowc, s + 1, c1, b1);
        }
        if (live) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int S = h ? pk::hi(best) : pk::lo(best);
                const int col = h ? bcolB : bcolA;
                int row = -1;
#pragma unroll
                for (int i = R - 1; i >= 0; --i) {
                    const int v = (int)snp[(h * kSnapRows + i) * kThreads];
                    if (r0 + i < c.m && v == S) row = r0 + i;
                }
                if (row >= 0 && (S > cS[h] || (S == cS[h] && col < cC[h]))) {
                    cS[h] = S; cC[h] = col; cR[h] = row;
                }
            }
        }
        // lane T-1's row is read by lane 0 of the same wave in the next band
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    // chain reduction: max score, then min column, then min row
    // (src/sw.c:150-158)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        int S = cS[h], C = cC[h], Rw = cR[h];
        for (int d = T >> 1; d >= 1; d >>= 1) {
            const int S2 = __shfl_xor(S, d, T), C2 = __shfl_xor(C, d, T), R2 = __shfl_xor(Rw, d, T);
            if (S2 > S || (S2 == S && (C2 < C || (C2 == C && R2 < Rw)))) { S = S2; C = C2; Rw = R2; }
        }
        if (t == 0 && live && J.tgt[h] >= 0) {
            phip_out_t o;
            memset(&o, 0, sizeof(o));
            o.score = S;
            o.end_query = Rw;
            o.end_ref = C;
            c.out[J.tgt[h]] = o;
            if (S > c.ovf_limit) c.redo[atomicAdd(c.nredo, 1)] = J.tgt[h];
        }
    }
}

}  // namespace

// LDS: [profile (LDSP)] [snapshots: 2 x kSnapRows rows of kThreads int16]
template <int R, bool LDSP>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2))) k_search(SearchCfg c)
{
    extern __shared__ int lds[];
    const short *prof = c.prof;
    int off = 0;
    if (LDSP) {
        for (int k = threadIdx.x; k < c.profn; k += blockDim.x) lds[k] = ((const int *)c.prof)[k];
        __syncthreads();
        prof = (const short *)lds;
        off = (c.profn + 3) & ~3;
    }
    short *const snp = (short *)(lds + off) + threadIdx.x;
    const int lane = threadIdx.x & 63;
    for (;;) {
        int wv = 0;
        if (lane == 0) wv = atomicAdd(c.next_wave, 1);
        wv = __builtin_amdgcn_readfirstlane(wv);
        if (wv >= c.nw) break;
        const DbWave wd = c.waves[c.w0 + wv];
        search_wave<R, LDSP>(c, wd, prof, lane, snp);
    }
}

// Column streams of the database jobs: T pad codes, the wave's width, pads.
__global__ void k_dbcode(const DbJob *jobs, const int *clen, const int *cT, int njobs,
                         const uint8_t *codes, const int64_t *toff, int PADT, uint16_t *colcode)
{
    const int jb = blockIdx.x;
    if (jb >= njobs) return;
    const DbJob J = jobs[jb];
    const int L = clen[jb], T = cT[jb];
    const uint8_t *ta = J.tgt[0] >= 0 ? codes + toff[J.tgt[0]] : nullptr;
    const uint8_t *tb = J.tgt[1] >= 0 ? codes + toff[J.tgt[1]] : nullptr;
    for (int k = threadIdx.x; k < L; k += blockDim.x) {
        const int j = k - T;
        const uint32_t a = (ta && j >= 0 && j < J.n[0]) ? ta[j] : PADT;
        const uint32_t b = (tb && j >= 0 && j < J.n[1]) ? tb[j] : PADT;
        colcode[J.ccoff + k] = (uint16_t)(a | (b << 8));
    }
}

__global__ void k_fill_int(int *p, int64_t n, int v)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x)
        p[k] = v;
}

// Sort keys: score (biased) high, complemented target index low, so that a
// descending sort gives score desc, then target index asc.
__global__ void k_keys(const phip_out_t *out, int n, uint64_t *keys)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t s = (uint32_t)(out[k].score + 0x40000000);
    keys[k] = ((uint64_t)s << 32) | (uint32_t)(0xFFFFFFFFu - (uint32_t)k);
}

__global__ void k_gather(const uint64_t *sorted, int k, const phip_out_t *out, phip_hit_t *hits)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k) return;
    const uint32_t tgt = 0xFFFFFFFFu - (uint32_t)(sorted[i] & 0xFFFFFFFFu);
    const phip_out_t o = out[tgt];
    hits[i].score = o.score;
    hits[i].end_query = o.end_query;
    hits[i].end_ref = o.end_ref;
    hits[i].target = (int)tgt;
}

// ------------------------------------------------------------------ host

namespace {

constexpr int kLongN = 2000;   // targets longer than this get a whole wave
constexpr int kSlots = 4;      // requests in flight (one stream each)
const int kRsSearch[] = {4, 8, 12, 16, 20};

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

double now_s()
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

}  // namespace

struct phip_db {
    int n = 0, size = 0;
    int64_t residues = 0;
    int maxw = 0, nwaves = 0;
    int64_t bbtot = 0;           // band-row entries over all jobs
    std::vector<int32_t> tlen;   // caller order
    std::vector<int> wT, wwidth;  // per queue entry: chain width and columns
    char *base = nullptr;
    uint8_t *dcodes = nullptr;
    int64_t *dtoff = nullptr;
    int32_t *dtlen = nullptr;
    DbJob *djobs = nullptr;
    DbWave *dwaves = nullptr;
    uint16_t *dcc = nullptr;
    int2 *dtop = nullptr;         // SW row 0 per open: rebuilt per request (tiny)
    int ntop = 0;
};

extern "C" phip_db_t *phip_db_create(const uint8_t *codes, const int64_t *off, const int32_t *len,
                                     int n, int size)
{
    if (phip_init() != 0) return nullptr;
    if (n <= 0 || size <= 0 || size > 255) {
        phip_set_error("phip_db_create: bad arguments");
        return nullptr;
    }
    phip_db_t *db = new phip_db_t();
    db->n = n;
    db->size = size;
    db->tlen.assign(len, len + n);
    std::vector<int> order(n);
    int64_t tot = 0;
    for (int i = 0; i < n; ++i) {
        order[i] = i;
        tot += len[i];
        if (len[i] <= 0 || len[i] > 60000) {
            phip_set_error("phip_db_create: target %d length %d out of range", i, len[i]);
            delete db;
            return nullptr;
        }
    }
    db->residues = tot;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return len[a] > len[b]; });
    std::vector<DbJob> jobs;
    std::vector<DbWave> waves;
    std::vector<int> clen, cT;
    int64_t ccoff = 0, bboff = 0;
    size_t k = 0;
    while (k < order.size()) {
        const int T = len[order[k]] > kLongN ? 64 : 16;
        DbWave w;
        memset(&w, 0, sizeof(w));
        w.job0 = (int)jobs.size();
        w.T = T;
        w.width = len[order[k]];
        const size_t per = (size_t)(64 / T) * 2;
        const size_t k1 = std::min(order.size(), k + per);
        w.njobs = (int)((k1 - k + 1) / 2);
        w.bbw = (w.width + T + 8 + 15) & ~15;
        const int L = (w.width + 2 * T + 8 + 15) & ~15;
        for (int q = 0; q < w.njobs; ++q) {
            DbJob J;
            memset(&J, 0, sizeof(J));
            for (int h = 0; h < 2; ++h) {
                const size_t idx = k + 2 * q + h;
                J.tgt[h] = idx < k1 ? order[idx] : -1;
                J.n[h] = idx < k1 ? len[order[idx]] : 0;
            }
            J.ccoff = ccoff;
            ccoff += L;
            J.bboff = bboff;
            bboff += w.bbw;
            jobs.push_back(J);
            clen.push_back(L);
            cT.push_back(T);
        }
        db->wT.push_back(T);
        db->wwidth.push_back(w.width);
        db->maxw = std::max(db->maxw, w.width);
        waves.push_back(w);
        k = k1;
    }
    db->nwaves = (int)waves.size();
    db->bbtot = bboff;
    db->ntop = db->maxw + 64 + 16;
    const size_t nj = jobs.size();
    size_t o = 0;
    auto slot = [&](size_t bytes) { size_t r = o; o += al256(bytes ? bytes : 1); return r; };
    const size_t o_codes = slot((size_t)tot), o_toff = slot((size_t)n * 8), o_tlen = slot((size_t)n * 4),
                 o_jobs = slot(nj * sizeof(DbJob)), o_waves = slot(waves.size() * sizeof(DbWave)),
                 o_cc = slot((size_t)ccoff * 2), o_top = slot((size_t)db->ntop * 8),
                 o_clen = slot(nj * 4), o_cT = slot(nj * 4);
    phip_lock();
    db->base = (char *)dev_alloc(o);
    if (!db->base) {
        phip_unlock();
        delete db;
        return nullptr;
    }
    char *b = db->base;
    db->dcodes = (uint8_t *)(b + o_codes);
    db->dtoff = (int64_t *)(b + o_toff);
    db->dtlen = (int32_t *)(b + o_tlen);
    db->djobs = (DbJob *)(b + o_jobs);
    db->dwaves = (DbWave *)(b + o_waves);
    db->dcc = (uint16_t *)(b + o_cc);
    db->dtop = (int2 *)(b + o_top);
    int *dclen = (int *)(b + o_clen), *dcT = (int *)(b + o_cT);
    hipStream_t st = phip_stream();
    std::vector<int64_t> toff(off, off + n);
    bool ok = true;
    // residues: one contiguous upload of the caller's concatenation
    int64_t span = 0;
    for (int i = 0; i < n; ++i) span = std::max<int64_t>(span, off[i] + len[i]);
    if (span != tot) {
        // compact into a dense copy so device offsets are 0..tot
        std::vector<uint8_t> dense((size_t)tot);
        int64_t p = 0;
        for (int i = 0; i < n; ++i) {
            memcpy(dense.data() + p, codes + off[i], (size_t)len[i]);
            toff[i] = p;
            p += len[i];
        }
        ok &= hipMemcpy(db->dcodes, dense.data(), (size_t)tot, hipMemcpyHostToDevice) == hipSuccess;
    } else {
        ok &= hipMemcpy(db->dcodes, codes, (size_t)tot, hipMemcpyHostToDevice) == hipSuccess;
    }
    ok &= hipMemcpy(db->dtoff, toff.data(), (size_t)n * 8, hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(db->dtlen, len, (size_t)n * 4, hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(db->djobs, jobs.data(), nj * sizeof(DbJob), hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(db->dwaves, waves.data(), waves.size() * sizeof(DbWave), hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(dclen, clen.data(), nj * 4, hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(dcT, cT.data(), nj * 4, hipMemcpyHostToDevice) == hipSuccess;
    if (ok && nj) {
        hipLaunchKernelGGL(k_dbcode, dim3((unsigned)nj), dim3(256), 0, st, db->djobs, dclen, dcT,
                           (int)nj, db->dcodes, db->dtoff, size, db->dcc);
        ok &= hipGetLastError() == hipSuccess;
    }
    ok &= hipStreamSynchronize(st) == hipSuccess;
    phip_unlock();
    if (!ok) {
        phip_set_error("phip_db_create: upload failed");
        phip_db_free(db);
        return nullptr;
    }
    return db;
}

extern "C" void phip_db_free(phip_db_t *db)
{
    if (!db) return;
    phip_lock();
    dev_free(db->base);
    phip_unlock();
    delete db;
}

extern "C" int64_t phip_db_residues(const phip_db_t *db) { return db ? db->residues : 0; }

namespace {

// Per-request device state: profile, results, band rows, sort buffers.
struct Req {
    int64_t id = 0;
    double arrival = 0;
    int m = 0, k = 0, R = 0, nb16 = 0, nb64 = 0, mp = 0, profn = 0;
    bool ldsp = false, needbb = false;
    int open = 0, gap = 0, ovf_limit = 0;
    std::vector<short> hprof;
    // exact int32 tier for overflowing targets
    std::vector<int> pssm;      // PSSM values (length x size) or square matrix
    std::vector<uint8_t> qcodes;
    int is_pssm = 0, msize = 0, mlen = 0;
    // device
    char *base = nullptr;
    size_t bytes = 0;
    short *dprof = nullptr;
    phip_out_t *dout = nullptr;
    int32_t *dredo = nullptr;
    int *dcnt = nullptr;         // [0] nredo, [1..] chunk queue counters
    int2 *dbb = nullptr, *dtop = nullptr;
    uint8_t *dq = nullptr;
    int *dmat = nullptr;
    int *dscratch = nullptr;
    uint64_t *dkeys = nullptr, *dsorted = nullptr;
    void *dtemp = nullptr;
    size_t temp_bytes = 0;
    phip_hit_t *dhits = nullptr;
    // chunks of the database queue
    std::vector<int> chunk_w0, chunk_nw;
    int next_chunk = 0, inflight = 0;
    std::deque<hipEvent_t> chunk_ev;
    bool finalized = false;
    hipEvent_t done{};
    phip_hit_t *hhits = nullptr;  // pinned
    int slot = -1;
    bool deferred = false;  // waited for memory at admission
};

int choose_R(int m, int *nb16, int *nb64)
{
    int bestR = 20;
    double bc = 1e300;
    for (int R : kRsSearch) {
        const int n16 = (m + 16 * R - 1) / (16 * R);
        const double c = (double)n16 * (11.0 * R + 30.0);
        if (c < bc) { bc = c; bestR = R; }
    }
    *nb16 = (m + 16 * bestR - 1) / (16 * bestR);
    *nb64 = (m + 64 * bestR - 1) / (64 * bestR);
    return bestR;
}

typedef void (*SFn)(SearchCfg);

SFn search_kernel(int R, bool ldsp)
{
    switch (R) {
    case 4: return ldsp ? k_search<4, true> : k_search<4, false>;
    case 8: return ldsp ? k_search<8, true> : k_search<8, false>;
    case 12: return ldsp ? k_search<12, true> : k_search<12, false>;
    case 16: return ldsp ? k_search<16, true> : k_search<16, false>;
    default: return ldsp ? k_search<20, true> : k_search<20, false>;
    }
}

int search_blocks()
{
    static int blocks = 0;
    if (!blocks) {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0)
            cus = 256;
        blocks = cus;  // one 8-wave block per CU (2 waves per SIMD)
    }
    return blocks;
}

// Builds the host side of a request (profile, tier choice, chunks).
int req_prepare(Req *r, const phip_db_t *db, const phip_query_t *q, int k)
{
    const int m = q->m;
    if (m <= 0 || m > 60000 || k <= 0 || q->open < 0 || q->gap < 0) {
        phip_set_error("search: bad query arguments");
        return -1;
    }
    if (q->mx.size != db->size) {
        phip_set_error("search: matrix alphabet differs from the database's");
        return -1;
    }
    r->m = m;
    r->k = std::min(k, db->n);
    r->open = q->open;
    r->gap = q->gap;
    r->R = choose_R(m, &r->nb16, &r->nb64);
    r->mp = std::max(r->nb16 * 16 * r->R, r->nb64 * 64 * r->R);
    r->needbb = r->nb16 > 1 || r->nb64 > 1;
    const int size = db->size;
    const int rows = size + 1;  // codes 0..size-1 and the pad code
    r->profn = rows * r->mp / 2;
    // the profile goes to LDS when it fits beside the snapshots in 64 KiB
    r->ldsp = (size_t)((r->profn + 3) & ~3) * 4 + 2 * kSnapRows * kThreads * 2 <= (64u << 10);
    r->hprof.assign((size_t)rows * r->mp, (short)W_NEG);
    int vmax = 0;
    bool fits = q->open <= 8000 && q->gap <= 8000;
    for (int c = 0; c < size; ++c)
        for (int i = 0; i < m; ++i) {
            const int v = q->mx.is_pssm ? q->mx.values[(size_t)i * size + c]
                                        : q->mx.values[(size_t)q->codes[i] * size + c];
            vmax = std::max(vmax, v);
            if (v + q->open > 16000 || v < -16000) fits = false;
            r->hprof[(size_t)c * r->mp + i] = (short)std::max(-16000, std::min(16000, v));
        }
    // int16 tier: targets whose running maximum passes this limit are re-run
    // exactly in int32 (all of them when the scores do not fit int16)
    r->ovf_limit = fits ? 32767 - (vmax + q->open + 1) : -1;
    r->is_pssm = q->mx.is_pssm;
    r->msize = size;
    r->mlen = q->mx.is_pssm ? m : size;
    r->pssm.assign(q->mx.values, q->mx.values + (size_t)r->mlen * size);
    if (q->codes) r->qcodes.assign(q->codes, q->codes + m);
    // chunks: consecutive queue entries (longest first).  An entry's cost is
    // its bands times its steps; a chunk holds at least 1/12 of the work and
    // at least its first (longest) entry's cost times the entries in flight,
    // so a chunk's duration is set by its work, not by one long target
    std::vector<double> cost(db->nwaves);
    double total = 0;
    for (int w = 0; w < db->nwaves; ++w) {
        const int nbT = db->wT[w] == 64 ? r->nb64 : r->nb16;
        cost[w] = (double)nbT * (db->wwidth[w] + db->wT[w]);
        total += cost[w];
    }
    const double inflight = (double)search_blocks() * (kThreads / 64);
    int w = 0;
    while (w < db->nwaves) {
        const int w0 = w;
        const double need = std::max(total / 12.0, cost[w0] * inflight);
        double acc = 0;
        while (w < db->nwaves && (acc < need || w == w0)) acc += cost[w++];
        r->chunk_w0.push_back(w0);
        r->chunk_nw.push_back(w - w0);
    }
    return 0;
}

// Lays out the request's device memory (profile, per-target results, band
// rows, top-k sort scratch...) into r->bytes and, if allocate, allocates it.
int req_alloc(Req *r, const phip_db_t *db, bool allocate)
{
    const int n = db->n;
    const int nch = (int)r->chunk_w0.size();
    size_t temp = 0;
    if (hipcub::DeviceRadixSort::SortKeysDescending(nullptr, temp, (uint64_t *)nullptr,
                                                    (uint64_t *)nullptr, n) != hipSuccess) {
        phip_set_error("search: sort sizing failed");
        return -1;
    }
    r->temp_bytes = temp;
    const size_t per_lane = (size_t)(60000 + 1) * 2 * sizeof(int);
    size_t o = 0;
    auto slot = [&](size_t bytes) { size_t rr = o; o += al256(bytes ? bytes : 1); return rr; };
    const size_t o_prof = slot(r->hprof.size() * 2), o_out = slot((size_t)n * sizeof(phip_out_t)),
                 o_redo = slot((size_t)n * 4), o_cnt = slot((size_t)(nch + 2) * 4),
                 o_bb = slot(r->needbb ? (size_t)db->bbtot * 8 : 0),
                 o_top = slot((size_t)db->ntop * 8),
                 o_q = slot(r->qcodes.size()), o_mat = slot(r->pssm.size() * 4),
                 o_scr = slot(per_lane * 64), o_keys = slot((size_t)n * 8),
                 o_sorted = slot((size_t)n * 8), o_temp = slot(temp), o_hits = slot((size_t)r->k * sizeof(phip_hit_t));
    r->bytes = o;
    if (!allocate) return 0;
    r->base = (char *)dev_alloc(o);
    if (!r->base) return -1;
    char *b = r->base;
    r->dprof = (short *)(b + o_prof);
    r->dout = (phip_out_t *)(b + o_out);
    r->dredo = (int32_t *)(b + o_redo);
    r->dcnt = (int *)(b + o_cnt);
    r->dbb = (int2 *)(b + o_bb);
    r->dtop = (int2 *)(b + o_top);
    r->dq = r->qcodes.empty() ? nullptr : (uint8_t *)(b + o_q);
    r->dmat = (int *)(b + o_mat);
    r->dscratch = (int *)(b + o_scr);
    r->dkeys = (uint64_t *)(b + o_keys);
    r->dsorted = (uint64_t *)(b + o_sorted);
    r->dtemp = b + o_temp;
    r->dhits = (phip_hit_t *)(b + o_hits);
    if (hipHostMalloc((void **)&r->hhits, (size_t)r->k * sizeof(phip_hit_t), 0) != hipSuccess) {
        r->hhits = nullptr;
        phip_set_error("search: pinned allocation failed");
        return -1;
    }
    return 0;
}

void req_release(Req *r)
{
    if (r->base) dev_free(r->base);
    r->base = nullptr;
    if (r->hhits) (void)hipHostFree(r->hhits);
    r->hhits = nullptr;
    for (hipEvent_t e : r->chunk_ev) (void)hipEventDestroy(e);
    r->chunk_ev.clear();
    if (r->finalized) (void)hipEventDestroy(r->done);
    r->finalized = false;
}

// Uploads and zeroes the request's state on stream st.
int req_start(Req *r, const phip_db_t *db, hipStream_t st)
{
    const int nch = (int)r->chunk_w0.size();
    HIP_OK(hipMemcpyAsync(r->dprof, r->hprof.data(), r->hprof.size() * 2, hipMemcpyHostToDevice, st));
    if (r->dq) HIP_OK(hipMemcpyAsync(r->dq, r->qcodes.data(), r->qcodes.size(), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(r->dmat, r->pssm.data(), r->pssm.size() * 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(256), 0, st, r->dcnt, (int64_t)(nch + 2), 0);
    const int2 top = make_int2(0, 0);  // H = 0, F clamped at 0
    std::vector<int2> htop((size_t)db->ntop, top);
    HIP_OK(hipMemcpyAsync(r->dtop, htop.data(), htop.size() * 8, hipMemcpyHostToDevice, st));
    // the pageable sources above must outlive the copies
    HIP_OK(hipStreamSynchronize(st));
    return 0;
}

int req_launch_chunk(Req *r, const phip_db_t *db, int ch, int blocks, hipStream_t st)
{
    SearchCfg c;
    memset(&c, 0, sizeof(c));
    c.jobs = db->djobs;
    c.waves = db->dwaves;
    c.w0 = r->chunk_w0[ch];
    c.nw = r->chunk_nw[ch];
    c.next_wave = r->dcnt + 2 + ch;
    c.colcode = db->dcc;
    c.prof = r->dprof;
    c.mp = r->mp;
    c.profn = r->profn;
    c.nb16 = r->nb16;
    c.nb64 = r->nb64;
    c.bandbuf = r->dbb;
    c.toprow = r->dtop;
    c.out = r->dout;
    c.redo = r->dredo;
    c.nredo = r->dcnt;
    c.ovf_limit = r->ovf_limit;
    c.m = r->m;
    c.open = r->open;
    c.gap = r->gap;
    const size_t lds = (r->ldsp ? (size_t)((r->profn + 3) & ~3) * 4 : 0) + 2 * kSnapRows * kThreads * 2;
    const int nbk = std::max(1, std::min(blocks, (c.nw + 7) / 8));
    hipLaunchKernelGGL(search_kernel(r->R, r->ldsp), dim3(nbk), dim3(kThreads), lds, st, c);
    HIP_OK(hipGetLastError());
    return 0;
}

// Int32 tier for overflowing targets, then the top-k, then the copy back.
int req_finish(Req *r, const phip_db_t *db, hipStream_t st)
{
    DPParams dp;
    memset(&dp, 0, sizeof(dp));
    dp.alg = PHIP_SW;
    dp.kind = PHIP_SCORE;
    dp.open = r->open;
    dp.gap = r->gap;
    dp.matrix = r->dmat;
    dp.msize = r->msize;
    dp.is_pssm = r->is_pssm;
    dp.q = r->dq;
    dp.qlen1 = r->m;
    dp.t = db->dcodes;
    dp.toff = db->dtoff;
    dp.tlen = db->dtlen;
    dp.npairs = db->n;
    dp.out = r->dout;
    hipLaunchKernelGGL(k_simple, dim3(1), dim3(64), 0, st, dp, 0, db->n, r->dredo, r->dcnt,
                       r->dscratch, 64, 60000);
    HIP_OK(hipGetLastError());
    const int n = db->n;
    hipLaunchKernelGGL(k_keys, dim3((n + 255) / 256), dim3(256), 0, st, r->dout, n, r->dkeys);
    HIP_OK(hipGetLastError());
    size_t temp = r->temp_bytes;
    HIP_OK(hipcub::DeviceRadixSort::SortKeysDescending(r->dtemp, temp, r->dkeys, r->dsorted, n, 0, 64, st));
    hipLaunchKernelGGL(k_gather, dim3((r->k + 63) / 64), dim3(64), 0, st, r->dsorted, r->k, r->dout,
                       r->dhits);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(r->hhits, r->dhits, (size_t)r->k * sizeof(phip_hit_t), hipMemcpyDeviceToHost, st));
    HIP_OK(hipEventCreateWithFlags(&r->done, hipEventDisableTiming));
    HIP_OK(hipEventRecord(r->done, st));
    r->finalized = true;
    return 0;
}

}  // namespace

extern "C" int phip_db_search(phip_db_t *db, const phip_query_t *q, int k, phip_hit_t *hits,
                              int *nhits)
{
    if (!db || !q || !hits || !nhits) {
        phip_set_error("phip_db_search: missing argument");
        return -1;
    }
    Req r;
    if (req_prepare(&r, db, q, k) != 0) return -1;
    phip_lock();
    int rc = -1;
    hipStream_t st = phip_stream();
    do {
        if (req_alloc(&r, db, true) != 0) break;
        if (req_start(&r, db, st) != 0) break;
        bool ok = true;
        for (int ch = 0; ch < (int)r.chunk_w0.size() && ok; ++ch)
            ok = req_launch_chunk(&r, db, ch, search_blocks(), st) == 0;
        if (!ok || req_finish(&r, db, st) != 0) break;
        if (hipStreamSynchronize(st) != hipSuccess) {
            phip_set_error("phip_db_search: device error");
            break;
        }
        memcpy(hits, r.hhits, (size_t)r.k * sizeof(phip_hit_t));
        *nhits = r.k;
        rc = 0;
    } while (0);
    req_release(&r);
    phip_unlock();
    return rc;
}

// ------------------------------------------------------------- serving loop

struct Pending {
    int64_t id;
    double arrival;
    Req *r;
};

struct Done {
    int64_t id;
    double latency_ms;
    int rc;
    std::vector<phip_hit_t> hits;
};

struct phip_server {
    phip_db_t *db = nullptr;
    size_t budget = 0, in_use = 0;
    int slots = kSlots;
    hipStream_t streams[kSlots]{};
    Req *active[kSlots]{};
    std::deque<Pending> pending;
    std::deque<Done> done;
    pthread_t thread{};
    pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
    pthread_cond_t cv = PTHREAD_COND_INITIALIZER;
    bool stop = false, started = false;
    int64_t completed = 0;
    int64_t deferred = 0;  // requests that waited for memory at admission
    int max_active = 0;    // most requests on the device at once
    double busy_cells = 0;
};

namespace {

int count_active(const phip_server_t *s)
{
    int a = 0;
    for (int i = 0; i < s->slots; ++i) a += s->active[i] != nullptr;
    return a;
}

void *server_main(void *arg)
{
    phip_server_t *s = (phip_server_t *)arg;
    for (;;) {
        bool progressed = false;
        // admission: a free slot and enough device memory
        pthread_mutex_lock(&s->mu);
        if (s->stop && s->pending.empty() && count_active(s) == 0) {
            pthread_mutex_unlock(&s->mu);
            break;
        }
        while (!s->pending.empty()) {
            int free_slot = -1;
            for (int i = 0; i < s->slots; ++i)
                if (!s->active[i]) { free_slot = i; break; }
            if (free_slot < 0) break;
            Pending p = s->pending.front();
            // its footprint against the budget first: with others running,
            // a request that does not fit waits (oldest first) for one of
            // them to finish; alone, it always runs
            int rc = p.r->bytes || req_alloc(p.r, s->db, false) == 0 ? 0 : -1;
            if (rc == 0 && s->in_use + p.r->bytes > s->budget && count_active(s) > 0) {
                if (!p.r->deferred) {
                    p.r->deferred = true;
                    s->deferred++;
                }
                break;
            }
            pthread_mutex_unlock(&s->mu);
            phip_lock();
            if (rc == 0) rc = req_alloc(p.r, s->db, true);
            if (rc == 0) rc = req_start(p.r, s->db, s->streams[free_slot]);
            phip_unlock();
            pthread_mutex_lock(&s->mu);
            s->pending.pop_front();
            if (rc != 0) {
                Done d;
                d.id = p.id;
                d.latency_ms = (now_s() - p.arrival) * 1e3;
                d.rc = -1;
                s->done.push_back(d);
                pthread_cond_broadcast(&s->cv);
                phip_lock();
                req_release(p.r);
                phip_unlock();
                delete p.r;
                continue;
            }
            p.r->id = p.id;
            p.r->arrival = p.arrival;
            p.r->slot = free_slot;
            s->in_use += p.r->bytes;
            s->active[free_slot] = p.r;
            s->max_active = std::max(s->max_active, count_active(s));
            progressed = true;
        }
        pthread_mutex_unlock(&s->mu);
        // keep every active request's stream fed with chunks, each launched
        // over the whole grid: the device runs the streams' kernels side by
        // side as CUs free up
        const int blocks = search_blocks();
        for (int i = 0; i < s->slots; ++i) {
            Req *r = s->active[i];
            if (!r) continue;
            while (!r->chunk_ev.empty() && hipEventQuery(r->chunk_ev.front()) == hipSuccess) {
                (void)hipEventDestroy(r->chunk_ev.front());
                r->chunk_ev.pop_front();
                progressed = true;
            }
            phip_lock();
            while (r->next_chunk < (int)r->chunk_w0.size() && r->chunk_ev.size() < 2) {
                hipEvent_t e;
                if (req_launch_chunk(r, s->db, r->next_chunk, blocks, s->streams[i]) != 0 ||
                    hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess ||
                    hipEventRecord(e, s->streams[i]) != hipSuccess)
                    break;
                r->chunk_ev.push_back(e);
                ++r->next_chunk;
                progressed = true;
            }
            if (r->next_chunk == (int)r->chunk_w0.size() && !r->finalized) {
                (void)req_finish(r, s->db, s->streams[i]);
                progressed = true;
            }
            phip_unlock();
            if (r->finalized && hipEventQuery(r->done) == hipSuccess) {
                Done d;
                d.id = r->id;
                d.latency_ms = (now_s() - r->arrival) * 1e3;
                d.rc = 0;
                d.hits.assign(r->hhits, r->hhits + r->k);
                phip_lock();
                req_release(r);
                phip_unlock();
                pthread_mutex_lock(&s->mu);
                s->in_use -= r->bytes;
                s->active[i] = nullptr;
                s->done.push_back(d);
                s->completed++;
                pthread_cond_broadcast(&s->cv);
                pthread_mutex_unlock(&s->mu);
                delete r;
                progressed = true;
            }
        }
        if (!progressed) {
            pthread_mutex_lock(&s->mu);
            const bool idle = s->pending.empty() && count_active(s) == 0 && !s->stop;
            if (idle) pthread_cond_wait(&s->cv, &s->mu);
            pthread_mutex_unlock(&s->mu);
            if (!idle) usleep(50);
        }
    }
    return nullptr;
}

}  // namespace

extern "C" phip_server_t *phip_server_create(phip_db_t *db, size_t mem_budget)
{
    if (!db) {
        phip_set_error("phip_server_create: missing database");
        return nullptr;
    }
    phip_server_t *s = new phip_server_t();
    s->db = db;
    s->budget = mem_budget ? mem_budget : ((size_t)32 << 30);
    for (int i = 0; i < s->slots; ++i)
        if (hipStreamCreateWithFlags(&s->streams[i], hipStreamNonBlocking) != hipSuccess) {
            phip_set_error("phip_server_create: stream creation failed");
            for (int k = 0; k < i; ++k) (void)hipStreamDestroy(s->streams[k]);
            delete s;
            return nullptr;
        }
    if (pthread_create(&s->thread, nullptr, server_main, s) != 0) {
        phip_set_error("phip_server_create: thread creation failed");
        for (int i = 0; i < s->slots; ++i) (void)hipStreamDestroy(s->streams[i]);
        delete s;
        return nullptr;
    }
    s->started = true;
    return s;
}

extern "C" int phip_server_submit(phip_server_t *s, int64_t id, const phip_query_t *q, int k,
                                  double arrival)
{
    if (!s || !q) return -1;
    Req *r = new Req();
    if (req_prepare(r, s->db, q, k) != 0) {
        delete r;
        return -1;
    }
    pthread_mutex_lock(&s->mu);
    s->pending.push_back(Pending{id, arrival > 0 ? arrival : now_s(), r});
    pthread_cond_broadcast(&s->cv);
    pthread_mutex_unlock(&s->mu);
    return 0;
}

extern "C" int phip_server_poll(phip_server_t *s, int64_t *id, phip_hit_t *hits, int *nhits,
                                double *latency_ms, int wait)
{
    if (!s) return -1;
    pthread_mutex_lock(&s->mu);
    while (wait && s->done.empty()) pthread_cond_wait(&s->cv, &s->mu);
    if (s->done.empty()) {
        pthread_mutex_unlock(&s->mu);
        return 0;
    }
    Done d = s->done.front();
    s->done.pop_front();
    pthread_mutex_unlock(&s->mu);
    if (id) *id = d.id;
    if (latency_ms) *latency_ms = d.latency_ms;
    if (nhits) *nhits = (int)d.hits.size();
    if (hits && !d.hits.empty()) memcpy(hits, d.hits.data(), d.hits.size() * sizeof(phip_hit_t));
    return d.rc == 0 ? 1 : -1;
}

extern "C" double phip_now(void) { return now_s(); }

extern "C" int phip_server_stats(phip_server_t *s, int64_t *completed, int64_t *deferred, int *max_active)
{
    if (!s) return -1;
    pthread_mutex_lock(&s->mu);
    if (completed) *completed = s->completed;
    if (deferred) *deferred = s->deferred;
    if (max_active) *max_active = s->max_active;
    pthread_mutex_unlock(&s->mu);
    return 0;
}

extern "C" void phip_server_free(phip_server_t *s)
{
    if (!s) return;
    pthread_mutex_lock(&s->mu);
    s->stop = true;
    pthread_cond_broadcast(&s->cv);
    pthread_mutex_unlock(&s->mu);
    if (s->started) pthread_join(s->thread, nullptr);
    for (int i = 0; i < s->slots; ++i) (void)hipStreamSynchronize(s->streams[i]);
    for (int i = 0; i < s->slots; ++i) (void)hipStreamDestroy(s->streams[i]);
    while (!s->pending.empty()) {
        delete s->pending.front().r;
        s->pending.pop_front();
    }
    delete s;
}
