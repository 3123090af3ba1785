// runtime.hip -- device context, memory pool, resident batches and the
// phip_* C ABI the host layer calls.
//
// One process-wide context (once-guard): device 0, one stream, a pool of
// device blocks reused across calls.  A resident batch keeps a batch's
// residues, lengths and results in HBM, and caches one wavefront plan per
// scoring so that repeated alignments of the same batch launch kernels
// only.  phip_align (one-shot) is a resident batch created, run and freed.
// Calls are serialised by one lock: every entry point may be called from
// any host thread.
#include <errno.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <vector>

#include "common.h"
#include "dp32.h"
#include "wave.h"

static char g_err[512];
static pthread_once_t g_once = PTHREAD_ONCE_INIT;
static int g_init_rc = -1;
static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;
static hipStream_t g_stream;
static hipEvent_t g_ev0, g_ev1;
static double g_last_ms, g_last_main_ms;

extern "C" void phip_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

extern "C" const char *phip_last_error(void) { return g_err; }

hipEvent_t phip_event(int i) { return i ? g_ev1 : g_ev0; }
void phip_note_kernel_time()
{
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, g_ev0, g_ev1) == hipSuccess) g_last_ms = g_last_main_ms = ms;
}
void phip_add_kernel_ms(double ms)
{
    g_last_ms += ms;
    g_last_main_ms += ms;
}
void phip_lock() { pthread_mutex_lock(&g_lock); }
void phip_unlock() { pthread_mutex_unlock(&g_lock); }
hipStream_t phip_stream() { return g_stream; }
extern "C" double phip_last_kernel_ms(void) { return g_last_ms; }
extern "C" double phip_last_main_kernel_ms(void) { return g_last_main_ms; }

static void init_once(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        phip_set_error("no HIP device available");
        g_init_rc = -1;
        return;
    }
    if (hipSetDevice(0) != hipSuccess || hipStreamCreate(&g_stream) != hipSuccess ||
        hipEventCreate(&g_ev0) != hipSuccess || hipEventCreate(&g_ev1) != hipSuccess) {
        phip_set_error("HIP context initialisation failed");
        g_init_rc = -1;
        return;
    }
    g_init_rc = 0;
}

extern "C" int phip_init(void)
{
    pthread_once(&g_once, init_once);
    if (g_init_rc != 0) errno = ENOSYS;
    return g_init_rc;
}

// ---------------------------------------------------------------- pool

namespace {

struct Block {
    char *p;
    size_t size;
    bool used;
};
std::vector<Block> g_blocks;

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

void *dev_alloc(size_t bytes)
{
    bytes = al256(bytes ? bytes : 1);
    Block *best = nullptr;
    for (Block &b : g_blocks)
        if (!b.used && b.size >= bytes && (!best || b.size < best->size)) best = &b;
    if (best && best->size <= 4 * bytes + ((size_t)16 << 20)) {
        best->used = true;
        return best->p;
    }
    size_t sz = std::max(bytes, (size_t)1 << 20);
    char *p = nullptr;
    if (hipMalloc(&p, sz) != hipSuccess) {
        // release idle blocks and retry once
        (void)hipStreamSynchronize(g_stream);
        for (Block &b : g_blocks)
            if (!b.used) { (void)hipFree(b.p); b.p = nullptr; }
        g_blocks.erase(std::remove_if(g_blocks.begin(), g_blocks.end(),
                                      [](const Block &b) { return b.p == nullptr; }),
                       g_blocks.end());
        if (hipMalloc(&p, sz) != hipSuccess) {
            phip_set_error("hipMalloc(%zu) failed", sz);
            return nullptr;
        }
    }
    g_blocks.push_back(Block{p, sz, true});
    return p;
}

void dev_free(void *p)
{
    for (Block &b : g_blocks)
        if (b.p == p) { b.used = false; return; }
}

// ---------------------------------------------------------------- resident

struct phip_resident {
    int np;
    int kind;
    int has_q;
    std::vector<int> mvals;
    phip_matrix_t mx;
    std::vector<int32_t> qlen, tlen;
    std::vector<int64_t> qoff, toff, troff;
    int maxn;
    char *base;
    int *dmat;
    uint8_t *dq, *dt;
    int64_t *dqoff, *dtoff, *dtroff;
    int32_t *dqlen, *dtlen;
    phip_out_t *dout;
    int8_t *dtr;
    int64_t trtot;
    std::vector<WavePlan *> plans;
    bool last_wave = false;  // the last pass ran the wavefront tier
    // a saturation check on NW/SG runs on the packed kernel when every pair
    // fits it (cover, for cover_key's alg, open, gap) and min_h is not tested
    bool want_min_h = false;
    int cover_key[3] = {-1, -1, -1};
    bool cover = false;
    // _stats / _trace: the wavefront plan (built for one check, alg, open,
    // gap), its device image, and per pair the trace layout of the last
    // pass (0 row-major, else dp32_layout)
    Dp32Plan dplan;
    int dplan_key[4] = {-1, -1, -1, -1};
    bool dplan_ext = false;  // some pair of the plan needs its extrema
    char *dplan_mem = nullptr;
    std::vector<int32_t> slow, tlay;
    int32_t *dslow = nullptr, *dtlay = nullptr;
    phip_out_t *hout = nullptr;  // pinned host copy of the results (NULL: none)
    // score-only results packed to 16 bytes a pair (device, then pinned host)
    phip_out4_t *dout4 = nullptr, *hout4 = nullptr;
    bool hout4_pinned = false;
    // several passes in one call (phip_resident_launch4): per pass a device
    // and a pinned copy of the packed results and its events
    static constexpr int kPasses = 4;
    phip_out4_t *pd4[kPasses] = {}, *ph4[kPasses] = {};
    hipEvent_t pev[kPasses][5] = {};  // start, end, main start, main end, downloaded
    int npass = 0;
    bool pass_wave[kPasses] = {};
    // the multi-pass call as a HIP graph: a call repeating the previous
    // call's passes (whose plans that call built) is captured once, then
    // replayed by every later such call
    bool graph_on = true;
    int gkey[3 + 2 * kPasses] = {};
    bool gkey_set = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool gwave[kPasses] = {};
    int graph_captures = 0, graph_replays = 0;
};

typedef struct phip_resident phip_resident_t;

static void resident_free(phip_resident_t *r);
static int resident_enqueue4(phip_resident_t *r, int npass, const int *alg, const int *sg_flags,
                             int open, int gap);

// wait: return once the uploads have ended.  A one-call batch passes false:
// its host sources live until the call returns, every later use of the
// uploads is ordered after them on the stream, and the host plans the pass
// meanwhile; resident_free waits for the stream before it frees.
static phip_resident_t *resident_create(const phip_batch_t *b, bool wait = true)
{
    phip_resident_t *r = new phip_resident_t();
    const int np = b->npairs;
    r->np = np;
    r->kind = b->kind;
    r->want_min_h = b->want_min_h != 0;
    r->has_q = b->q != nullptr;
    r->mx = b->matrix;
    r->mvals.assign(b->matrix.values, b->matrix.values + (size_t)b->matrix.size * b->matrix.length);
    r->mx.values = r->mvals.data();
    r->qlen.assign(b->qlen, b->qlen + np);
    r->tlen.assign(b->tlen, b->tlen + np);
    r->qoff.assign(b->qoff, b->qoff + np);
    r->toff.assign(b->toff, b->toff + np);
    int64_t qtot = 0, ttot = 0;
    r->maxn = 0;
    r->trtot = 0;
    r->troff.assign(b->kind == PHIP_TRACE ? np : 0, 0);
    for (int i = 0; i < np; ++i) {
        qtot = std::max<int64_t>(qtot, b->qoff[i] + b->qlen[i]);
        ttot = std::max<int64_t>(ttot, b->toff[i] + b->tlen[i]);
        r->maxn = std::max(r->maxn, (int)b->tlen[i]);
        if (b->kind == PHIP_TRACE) {
            // room for either layout of the pair's trace
            r->troff[i] = r->trtot;
            const int64_t rm = (int64_t)b->qlen[i] * b->tlen[i];
            r->trtot += (std::max(rm, dp32_trace_bytes(b->qlen[i], b->tlen[i])) + 15) & ~(int64_t)15;
        }
    }
    size_t off = 0;
    auto slot = [&](size_t bytes) { size_t o = off; off += al256(bytes ? bytes : 1); return o; };
    size_t o_mat = slot(r->mvals.size() * 4), o_q = slot(r->has_q ? qtot : 0), o_t = slot(ttot),
           o_qo = slot(np * 8), o_to = slot(np * 8), o_ql = slot(np * 4), o_tl = slot(np * 4),
           o_out = slot(np * sizeof(phip_out_t)), o_tro = slot(r->troff.size() * 8),
           o_tr = slot(r->trtot), o_slow = slot(np * 4), o_tlay = slot(np * 4),
           o_out4 = slot(np * sizeof(phip_out4_t));
    r->base = (char *)dev_alloc(off);
    if (!r->base) { delete r; return nullptr; }
    r->dmat = (int *)(r->base + o_mat);
    r->dq = r->has_q ? (uint8_t *)(r->base + o_q) : nullptr;
    r->dt = (uint8_t *)(r->base + o_t);
    r->dqoff = (int64_t *)(r->base + o_qo);
    r->dtoff = (int64_t *)(r->base + o_to);
    r->dqlen = (int32_t *)(r->base + o_ql);
    r->dtlen = (int32_t *)(r->base + o_tl);
    r->dout = (phip_out_t *)(r->base + o_out);
    r->dtroff = (int64_t *)(r->base + o_tro);
    r->dtr = b->kind == PHIP_TRACE ? (int8_t *)(r->base + o_tr) : nullptr;
    r->dslow = (int32_t *)(r->base + o_slow);
    r->dtlay = (int32_t *)(r->base + o_tlay);
    r->dout4 = (phip_out4_t *)(r->base + o_out4);
    bool ok = true;
    ok &= hipMemcpyAsync(r->dmat, r->mvals.data(), r->mvals.size() * 4, hipMemcpyHostToDevice, g_stream) == hipSuccess;
    if (r->has_q) ok &= hipMemcpyAsync(r->dq, b->q, qtot, hipMemcpyHostToDevice, g_stream) == hipSuccess;
    ok &= hipMemcpyAsync(r->dt, b->t, ttot, hipMemcpyHostToDevice, g_stream) == hipSuccess;
    ok &= hipMemcpyAsync(r->dqoff, b->qoff, np * 8, hipMemcpyHostToDevice, g_stream) == hipSuccess;
    ok &= hipMemcpyAsync(r->dtoff, b->toff, np * 8, hipMemcpyHostToDevice, g_stream) == hipSuccess;
    ok &= hipMemcpyAsync(r->dqlen, b->qlen, np * 4, hipMemcpyHostToDevice, g_stream) == hipSuccess;
    ok &= hipMemcpyAsync(r->dtlen, b->tlen, np * 4, hipMemcpyHostToDevice, g_stream) == hipSuccess;
    if (!r->troff.empty())
        ok &= hipMemcpyAsync(r->dtroff, r->troff.data(), np * 8, hipMemcpyHostToDevice, g_stream) == hipSuccess;
    if (wait || !ok) ok &= hipStreamSynchronize(g_stream) == hipSuccess;
    if (!ok) {
        phip_set_error("resident batch upload failed");
        dev_free(r->base);
        delete r;
        return nullptr;
    }
    return r;
}

// Pinned host copies of the results, for the batches the API keeps resident
// (a pageable copy of a large batch's results costs more than some of its
// kernels); a one-call batch copies straight into the caller's array.
static int resident_host_buffers(phip_resident_t *r)
{
    const int np = r->np;
    if (!r->hout &&
        hipHostMalloc((void **)&r->hout, (size_t)np * sizeof(phip_out_t), hipHostMallocDefault) !=
            hipSuccess) {
        (void)hipGetLastError();
        r->hout = nullptr;
    }
    if (!r->hout4) {
        r->hout4_pinned = hipHostMalloc((void **)&r->hout4, (size_t)np * sizeof(phip_out4_t),
                                        hipHostMallocDefault) == hipSuccess;
        if (!r->hout4_pinned) {
            (void)hipGetLastError();
            r->hout4 = (phip_out4_t *)malloc((size_t)np * sizeof(phip_out4_t));
        }
    }
    if (!r->hout4) {
        phip_set_error("resident batch: out of host memory");
        return -1;
    }
    return 0;
}

static void resident_free(phip_resident_t *r)
{
    if (!r) return;
    // copies from r's host vectors, or kernels on its blocks, may be in flight
    (void)hipStreamSynchronize(g_stream);
    if (r->gexec) (void)hipGraphExecDestroy(r->gexec);
    if (r->graph) (void)hipGraphDestroy(r->graph);
    for (int k = 0; k < phip_resident_t::kPasses; ++k) {
        if (r->pd4[k]) dev_free(r->pd4[k]);
        if (r->ph4[k]) (void)hipHostFree(r->ph4[k]);
        for (int e = 0; e < 5; ++e)
            if (r->pev[k][e]) (void)hipEventDestroy(r->pev[k][e]);
    }
    for (WavePlan *p : r->plans) wave_plan_free(p);
    if (r->dplan_mem) dev_free(r->dplan_mem);
    dev_free(r->base);
    if (r->hout) (void)hipHostFree(r->hout);
    if (r->hout4) {
        if (r->hout4_pinned)
            (void)hipHostFree(r->hout4);
        else
            free(r->hout4);
    }
    delete r;
}

static DPParams make_params(const phip_resident_t *r, int alg, int kind, int sg_flags, int open,
                            int gap, int check)
{
    DPParams dp;
    memset(&dp, 0, sizeof(dp));
    dp.alg = alg;
    dp.kind = kind;
    dp.sg_flags = alg == PHIP_SG ? sg_flags : 0;
    dp.open = open;
    dp.gap = gap;
    dp.check = check;
    dp.matrix = r->dmat;
    dp.msize = r->mx.size;
    dp.is_pssm = r->mx.is_pssm;
    dp.q = r->dq;
    dp.qoff = r->dqoff;
    dp.qlen = r->dqlen;
    dp.t = r->dt;
    dp.toff = r->dtoff;
    dp.tlen = r->dtlen;
    dp.npairs = r->np;
    dp.out = r->dout;
    dp.trace = r->dtr;
    dp.trace_off = r->dtroff;
    return dp;
}

static int simple_run(phip_resident_t *r, const DPParams &dp, const int32_t *list, int count);

// Pinned bounce buffer for the planner's uploads (held under g_lock; each
// use waits for its copies before returning): a pageable copy is staged by
// the runtime in pieces and costs ~1 ms for a plan of 100,000 pairs.
static char *g_up = nullptr;
static size_t g_up_cap = 0;

static int upload_via_pinned(void *dst, const void *src, size_t bytes, size_t at)
{
    if (!bytes) return 0;
    if (at + bytes > g_up_cap) return -1;
    memcpy(g_up + at, src, bytes);
    return hipMemcpyAsync(dst, g_up + at, bytes, hipMemcpyHostToDevice, g_stream) == hipSuccess ? 0 : -1;
}

static int reserve_up(size_t bytes)
{
    if (bytes <= g_up_cap) return 0;
    if (g_up) (void)hipHostFree(g_up);
    g_up = nullptr;
    g_up_cap = 0;
    const size_t cap = std::max(bytes + (bytes >> 2), (size_t)4 << 20);
    if (hipHostMalloc((void **)&g_up, cap, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        g_up = nullptr;
        return -1;
    }
    g_up_cap = cap;
    return 0;
}

// Runs one alignment pass over the resident batch; results stay in r->dout.
static int resident_run(phip_resident_t *r, int alg, int kind, int sg_flags, int open, int gap,
                        int check, hipEvent_t me0 = nullptr, hipEvent_t me1 = nullptr)
{
    DPParams dp = make_params(r, alg, kind, sg_flags, open, gap, check);
    HostBatch hb;
    hb.np = r->np;
    hb.qlen = r->qlen.data();
    hb.tlen = r->tlen.data();
    hb.qoff = r->qoff.data();
    hb.toff = r->toff.data();
    hb.mx = &r->mx;
    bool wave = kind == PHIP_SCORE && !r->mx.is_pssm;
    if (wave && check && alg != PHIP_SW) {
        // the packed kernel returns the extrema of H/E/F (not min_h); its
        // int32 tiers inside the plan would leave long pairs to the general
        // kernel, so a batch with such pairs takes the int32 lane chains
        const int key[3] = {alg, open, gap};
        if (memcmp(key, r->cover_key, sizeof(key)) != 0) {
            r->cover = wave_covers(hb, dp);
            memcpy(r->cover_key, key, sizeof(key));
        }
        wave = r->cover && !r->want_min_h;
    }
    r->last_wave = wave;
    if (wave) {
        WavePlan *plan = nullptr;
        for (WavePlan *p : r->plans)
            if (wave_plan_matches(p, alg, dp.sg_flags, open, gap, check)) plan = p;
        if (!plan) {
            plan = wave_plan_create(hb, dp, g_stream);
            if (!plan) return -1;
            r->plans.push_back(plan);
        }
        HIP_OK(hipEventRecord(g_ev0, g_stream));
        if (wave_plan_run(plan, dp, g_stream, me0, me1) != 0) return -1;
    } else {
        // the int32 wavefront kernel for the pairs it covers, the general
        // tier for the rest
        const int key[4] = {check, alg, open, gap};
        if (memcmp(key, r->dplan_key, sizeof(key)) != 0) {
            std::vector<int> fast;
            std::vector<int32_t> status(r->np, PHIP_ST_W32);
            r->slow.clear();
            r->tlay.assign(r->np, 0);
            r->dplan_ext = false;
            for (int i = 0; i < r->np; ++i) {
                if (dp32_eligible(r->mx, alg, kind, check, open, gap, r->qlen[i], r->tlen[i])) {
                    fast.push_back(i);
                    r->tlay[i] = dp32_layout(r->qlen[i]);
                    if (!dp32_window_holds(r->mx, alg, kind, check, open, gap, r->qlen[i],
                                           r->tlen[i]))
                        r->dplan_ext = true;
                } else {
                    r->slow.push_back(i);
                }
            }
            if (!dp32_plan(hb, fast, status.data(), r->troff.empty() ? nullptr : r->troff.data(),
                           r->dplan)) {
                phip_set_error("dp32 plan failed");
                return -1;
            }
            if (r->dplan_mem) dev_free(r->dplan_mem);
            r->dplan_mem = nullptr;
            if (r->dplan.bytes) {
                r->dplan_mem = (char *)dev_alloc(r->dplan.bytes);
                if (!r->dplan_mem) return -1;
            }
            const size_t a_slow = al256(r->dplan.bytes), a_tlay = a_slow + al256(r->slow.size() * 4);
            if (reserve_up(a_tlay + (size_t)r->np * 4) == 0) {
                if (upload_via_pinned(r->dplan_mem, r->dplan.image.data(), r->dplan.bytes, 0) != 0 ||
                    upload_via_pinned(r->dslow, r->slow.data(), r->slow.size() * 4, a_slow) != 0 ||
                    upload_via_pinned(r->dtlay, r->tlay.data(), (size_t)r->np * 4, a_tlay) != 0) {
                    phip_set_error("dp32 plan upload failed");
                    return -1;
                }
            } else {
                if (r->dplan.bytes)
                    HIP_OK(hipMemcpyAsync(r->dplan_mem, r->dplan.image.data(), r->dplan.bytes,
                                          hipMemcpyHostToDevice, g_stream));
                if (!r->slow.empty())
                    HIP_OK(hipMemcpyAsync(r->dslow, r->slow.data(), r->slow.size() * 4,
                                          hipMemcpyHostToDevice, g_stream));
                HIP_OK(hipMemcpyAsync(r->dtlay, r->tlay.data(), (size_t)r->np * 4,
                                      hipMemcpyHostToDevice, g_stream));
            }
            // the pageable host vectors stay alive in r; the copies must end
            // before they change
            HIP_OK(hipStreamSynchronize(g_stream));
            memcpy(r->dplan_key, key, sizeof(key));
        }
        HIP_OK(hipEventRecord(g_ev0, g_stream));
        DPParams dpw = dp;
        if (!r->dplan_ext) dpw.check = 0;  // every window provably holds
        if (dp32_launch(dpw, r->dplan, r->dplan_mem, g_stream) != 0) return -1;
        if (!r->slow.empty() && simple_run(r, dp, r->dslow, (int)r->slow.size()) != 0) return -1;
    }
    HIP_OK(hipEventRecord(g_ev1, g_stream));
    return 0;
}

// The general int32 tier over `count` pairs: all of the batch (list NULL)
// or the device list.
static int simple_run(phip_resident_t *r, const DPParams &dp, const int32_t *list, int count)
{
    const int narr = dp.kind == PHIP_STATS ? 8 : 2;
    const size_t per_lane = (size_t)(r->maxn + 1) * narr * sizeof(int);
    const int lanes = (int)std::max<size_t>(
        64, std::min<size_t>((size_t)count, ((size_t)1 << 30) / per_lane));
    int *scratch = (int *)dev_alloc(per_lane * lanes);
    if (!scratch) return -1;
    bool ok = true;
    if (list) {
        hipLaunchKernelGGL(k_simple, dim3((lanes + 255) / 256), dim3(256), 0, g_stream, dp, 0, count,
                           list, (const int *)nullptr, scratch, lanes, r->maxn);
        ok = hipGetLastError() == hipSuccess;
    } else {
        for (int first = 0; ok && first < count; first += lanes) {
            const int cnt = std::min(lanes, count - first);
            hipLaunchKernelGGL(k_simple, dim3((cnt + 255) / 256), dim3(256), 0, g_stream, dp, first,
                               cnt, (const int32_t *)nullptr, (const int *)nullptr, scratch, cnt,
                               r->maxn);
            ok = hipGetLastError() == hipSuccess;
        }
    }
    // the block goes back to the pool; later users queue behind these
    // kernels on the same stream
    dev_free(scratch);
    if (!ok) {
        phip_set_error("k_simple launch failed");
        return -1;
    }
    return 0;
}

static int resident_fetch(phip_resident_t *r, phip_out_t *out, int8_t **traces)
{
    if (out)
        HIP_OK(hipMemcpyAsync(out, r->dout, (size_t)r->np * sizeof(phip_out_t),
                              hipMemcpyDeviceToHost, g_stream));
    HIP_OK(hipStreamSynchronize(g_stream));
    if (out)
        for (int i = 0; i < r->np; ++i)
            if (out[i].status & PHIP_ST_ERR) {
                phip_set_error("pair %d could not be computed (long-pair pipeline failed)", i);
                return -1;
            }
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, g_ev0, g_ev1));
    g_last_ms = ms;
    g_last_main_ms = r->last_wave ? wave_last_kernel_ms() : ms;
    if (traces && r->dtr) {
        std::vector<int8_t> wbuf;
        for (int i = 0; i < r->np; ++i) {
            if (!traces[i]) continue;
            const int m = r->qlen[i], n = r->tlen[i];
            const int lay = r->tlay.empty() ? 0 : r->tlay[i];
            if (!lay) {
                HIP_OK(hipMemcpy(traces[i], r->dtr + r->troff[i], (size_t)m * n,
                                 hipMemcpyDeviceToHost));
                continue;
            }
            // wavefront-order 4-bit fields to row-major codes
            wbuf.resize((size_t)dp32_trace_bytes(m, n));
            HIP_OK(hipMemcpy(wbuf.data(), r->dtr + r->troff[i], wbuf.size(), hipMemcpyDeviceToHost));
            for (int row = 0; row < m; ++row)
                for (int j = 0; j < n; ++j)
                    traces[i][(size_t)row * n + j] =
                        (int8_t)dp32_trace_code((const uint8_t *)wbuf.data(), lay, row, j);
        }
    }
    return 0;
}

// ---------------------------------------------------------------- C ABI

extern "C" int phip_align(const phip_batch_t *b, phip_out_t *out, int8_t **traces)
{
    if (phip_init() != 0) return -1;
    if (!b || !out || b->npairs < 0) { phip_set_error("phip_align: bad arguments"); return -1; }
    if (b->npairs == 0) return 0;
    pthread_mutex_lock(&g_lock);
    int rc = -1;
    phip_resident_t *r = resident_create(b, false);
    if (r) {
        rc = resident_run(r, b->alg, b->kind, b->sg_flags, b->open, b->gap, b->want_sat_check);
        if (rc == 0) rc = resident_fetch(r, out, traces);
        resident_free(r);
    }
    pthread_mutex_unlock(&g_lock);
    return rc;
}

// ---------------------------------------------------------------- one pair

// A profile's device copy: one allocation holding the query codes, then the
// matrix values at a 256-byte boundary.  It is made with hipMalloc, not the
// pool, since it lives as long as the caller's profile.
struct phip_dprofile {
    uint8_t *q;
    int *mat;
    int m;
};

namespace {

// pinned staging and its device mirror for per-call alignments (kept for
// the process: every call reuses them)
char *g_one_h = nullptr, *g_one_d = nullptr;
size_t g_one_cap = 0;

int one_reserve(size_t bytes)
{
    if (bytes <= g_one_cap) return 0;
    const size_t cap = std::max(bytes + (bytes >> 1), (size_t)1 << 20);
    if (g_one_h) (void)hipHostFree(g_one_h);
    if (g_one_d) (void)hipFree(g_one_d);
    g_one_h = g_one_d = nullptr;
    g_one_cap = 0;
    HIP_OK(hipHostMalloc((void **)&g_one_h, cap, hipHostMallocDefault));
    HIP_OK(hipMalloc((void **)&g_one_d, cap));
    g_one_cap = cap;
    return 0;
}

}  // namespace

// Per-call alignments.  One pair runs on a whole wave (64 lanes, the
// shortest chain for one pair) of the int32 lane-chain kernel.  Calls made
// at the same time from several host threads -- how parasail's callers use
// the per-pair functions, OpenMP over pairs (tests/test_openmp.c:353-369)
// -- are coalesced into one launch: a caller queues its pair, the first
// caller that finds no launch under way leads, takes every queued pair
// that shares its scoring (algorithm, kind, end flags, gaps, matrix and
// saturation check; at most kOneMax) and runs them as one batch of 64-lane
// jobs with one upload, one launch and one download; the others wait for
// their results.  A lone caller is a batch of one.  Callers that were in
// one launch come back together, so a leader first waits, at most
// kGatherUs, until as many calls are queued as its predecessor launched
// (the first caller back would otherwise launch alone).  Pairs outside the
// kernel's range take phip_align.
namespace {

struct OneReq {
    const phip_batch_t *b;
    phip_out_t *out;
    int8_t **traces;
    int rc;
    bool done;
};

pthread_mutex_t g_oq_lock = PTHREAD_MUTEX_INITIALIZER;
pthread_cond_t g_oq_cv = PTHREAD_COND_INITIALIZER;
pthread_cond_t g_oq_gather_cv = PTHREAD_COND_INITIALIZER;
std::vector<OneReq *> g_oq;
bool g_oq_busy = false, g_oq_gathering = false;
int g_oq_last = 1;  // calls in the previous launch
constexpr int kOneMax = 512;
constexpr long kGatherUs = 200;

bool same_scoring(const phip_batch_t *a, const phip_batch_t *b)
{
    return a->alg == b->alg && a->kind == b->kind && a->sg_flags == b->sg_flags &&
           a->open == b->open && a->gap == b->gap && a->want_sat_check == b->want_sat_check &&
           a->matrix.values == b->matrix.values && a->matrix.size == b->matrix.size &&
           a->matrix.is_pssm == b->matrix.is_pssm && a->dprof == b->dprof;
}

// one launch for the requests of one scoring; each request's rc is set
void one_batch(const std::vector<OneReq *> &reqs)
{
    const int np = (int)reqs.size();
    const phip_batch_t *b0 = reqs[0]->b;
    const int kind = b0->kind, msize = b0->matrix.size;
    // a profile's calls share its device query and matrix (same_scoring)
    const phip_dprofile_t *dpf = b0->dprof;
    std::vector<int32_t> ql(np), tl(np), status(np, PHIP_ST_W32);
    std::vector<int64_t> qo(np), to(np), tro(np, 0);
    int64_t qtot = 0, ttot = 0, trtot = 0;
    bool ext = false;
    for (int k = 0; k < np; ++k) {
        const phip_batch_t *b = reqs[k]->b;
        ql[k] = b->qlen[0];
        tl[k] = b->tlen[0];
        qo[k] = dpf ? 0 : qtot;
        to[k] = ttot;
        if (!dpf) qtot += ql[k];
        ttot += tl[k];
        if (kind == PHIP_TRACE) {
            tro[k] = trtot;
            trtot += (dp32_trace_bytes(ql[k], tl[k], 64) + 15) & ~(int64_t)15;
        }
        if (!dp32_window_holds(b->matrix, b->alg, kind, b->want_sat_check, b->open, b->gap, ql[k],
                               tl[k]))
            ext = true;
    }
    HostBatch hb;
    hb.np = np;
    hb.qlen = ql.data();
    hb.tlen = tl.data();
    hb.qoff = qo.data();
    hb.toff = to.data();
    hb.mx = &b0->matrix;
    std::vector<int> pairs(np);
    for (int k = 0; k < np; ++k) pairs[k] = k;
    Dp32Plan P;
    int rc = -1;
    if (!dp32_plan(hb, pairs, status.data(), tro.data(), P, 64, kind != PHIP_TRACE)) {
        for (OneReq *r : reqs) r->rc = phip_align(r->b, r->out, r->traces);
        return;
    }
    {
        size_t off = 0;
        auto slot = [&](size_t bytes) { size_t o = off; off += al256(bytes ? bytes : 1); return o; };
        const size_t o_plan = slot(P.bytes), o_out = slot(sizeof(phip_out_t) * np),
                     o_qo = slot(8 * np), o_to = slot(8 * np), o_ql = slot(4 * np),
                     o_tl = slot(4 * np), o_mat = slot(dpf ? 0 : (size_t)msize * msize * 4),
                     o_q = slot((size_t)qtot), o_t = slot((size_t)ttot);
        const size_t up_end = off;
        const size_t o_tr = slot((size_t)trtot);
        pthread_mutex_lock(&g_lock);
        do {
            if (one_reserve(off) != 0) break;
            char *h = g_one_h, *d = g_one_d;
            memcpy(h + o_plan, P.image.data(), P.bytes);
            memcpy(h + o_qo, qo.data(), 8 * np);
            memcpy(h + o_to, to.data(), 8 * np);
            memcpy(h + o_ql, ql.data(), 4 * np);
            memcpy(h + o_tl, tl.data(), 4 * np);
            if (!dpf) memcpy(h + o_mat, b0->matrix.values, (size_t)msize * msize * 4);
            for (int k = 0; k < np; ++k) {
                if (!dpf) memcpy(h + o_q + qo[k], reqs[k]->b->q, (size_t)ql[k]);
                memcpy(h + o_t + to[k], reqs[k]->b->t, (size_t)tl[k]);
            }
            if (hipMemcpyAsync(d, h, up_end, hipMemcpyHostToDevice, g_stream) != hipSuccess) {
                phip_set_error("phip_align_one: upload failed");
                break;
            }
            DPParams dp;
            memset(&dp, 0, sizeof(dp));
            dp.alg = b0->alg;
            dp.kind = kind;
            dp.sg_flags = b0->alg == PHIP_SG ? b0->sg_flags : 0;
            dp.open = b0->open;
            dp.gap = b0->gap;
            dp.check = ext ? b0->want_sat_check : 0;
            dp.matrix = dpf ? dpf->mat : (const int *)(d + o_mat);
            dp.msize = msize;
            dp.q = dpf ? dpf->q : (const uint8_t *)(d + o_q);
            dp.qoff = (const int64_t *)(d + o_qo);
            dp.qlen = (const int32_t *)(d + o_ql);
            dp.t = (const uint8_t *)(d + o_t);
            dp.toff = (const int64_t *)(d + o_to);
            dp.tlen = (const int32_t *)(d + o_tl);
            dp.npairs = np;
            dp.out = (phip_out_t *)(d + o_out);
            dp.trace = (int8_t *)(d + o_tr);
            if (hipEventRecord(g_ev0, g_stream) != hipSuccess) break;
            if (dp32_launch(dp, P, d + o_plan, g_stream) != 0) break;
            if (hipEventRecord(g_ev1, g_stream) != hipSuccess) break;
            bool ok = hipMemcpyAsync(h + o_out, d + o_out, sizeof(phip_out_t) * np,
                                     hipMemcpyDeviceToHost, g_stream) == hipSuccess;
            if (trtot)
                ok &= hipMemcpyAsync(h + o_tr, d + o_tr, (size_t)trtot, hipMemcpyDeviceToHost,
                                     g_stream) == hipSuccess;
            ok &= hipStreamSynchronize(g_stream) == hipSuccess;
            if (!ok) {
                phip_set_error("phip_align_one: download failed");
                break;
            }
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, g_ev0, g_ev1) == hipSuccess) g_last_ms = g_last_main_ms = ms;
            for (int k = 0; k < np; ++k) {
                memcpy(reqs[k]->out, h + o_out + sizeof(phip_out_t) * k, sizeof(phip_out_t));
                if (kind == PHIP_TRACE) {
                    // wavefront-order 4-bit fields to row-major codes
                    const int m = ql[k], n = tl[k], lay = dp32_layout(m, 64);
                    const uint8_t *w = (const uint8_t *)(h + o_tr + tro[k]);
                    int8_t *dst = reqs[k]->traces[0];
                    for (int row = 0; row < m; ++row)
                        for (int j = 0; j < n; ++j)
                            dst[(size_t)row * n + j] = (int8_t)dp32_trace_code(w, lay, row, j);
                }
            }
            rc = 0;
        } while (0);
        pthread_mutex_unlock(&g_lock);
    }
    for (OneReq *r : reqs) r->rc = rc;
}

}  // namespace

extern "C" int phip_align_one(const phip_batch_t *b, phip_out_t *out, int8_t **traces)
{
    if (phip_init() != 0) return -1;
    if (!b || !out || b->npairs != 1 || !b->q || (b->kind == PHIP_TRACE && !traces))
        return phip_align(b, out, traces);
    const int m = b->qlen[0], n = b->tlen[0];
    if (b->dprof && b->dprof->m != m) {
        phip_set_error("phip_align_one: the profile's device query has %d residues, not %d",
                       b->dprof->m, m);
        return -1;
    }
    if (m <= 0 || n <= 0 ||
        !dp32_eligible_one(b->matrix, b->alg, b->kind, b->want_sat_check, b->open, b->gap, m, n))
        return phip_align(b, out, traces);
    OneReq me{b, out, traces, -1, false};
    pthread_mutex_lock(&g_oq_lock);
    g_oq.push_back(&me);
    if (g_oq_gathering) pthread_cond_signal(&g_oq_gather_cv);
    while (!me.done) {
        if (g_oq_busy) {
            pthread_cond_wait(&g_oq_cv, &g_oq_lock);
            continue;
        }
        g_oq_busy = true;
        // gather: the previous launch's callers are on their way back
        if ((int)g_oq.size() < g_oq_last) {
            struct timespec dl;
            clock_gettime(CLOCK_REALTIME, &dl);
            dl.tv_nsec += kGatherUs * 1000;
            if (dl.tv_nsec >= 1000000000L) {
                dl.tv_sec += 1;
                dl.tv_nsec -= 1000000000L;
            }
            g_oq_gathering = true;
            while ((int)g_oq.size() < g_oq_last &&
                   pthread_cond_timedwait(&g_oq_gather_cv, &g_oq_lock, &dl) == 0) {
            }
            g_oq_gathering = false;
        }
        // lead: the oldest request's scoring, every queued request that
        // shares it
        std::vector<OneReq *> batch, rest;
        const phip_batch_t *key = g_oq.front()->b;
        for (OneReq *r : g_oq) {
            if ((int)batch.size() < kOneMax && same_scoring(r->b, key)) batch.push_back(r);
            else rest.push_back(r);
        }
        g_oq.swap(rest);
        g_oq_last = (int)batch.size();
        pthread_mutex_unlock(&g_oq_lock);
        one_batch(batch);
        pthread_mutex_lock(&g_oq_lock);
        for (OneReq *r : batch) r->done = true;
        g_oq_busy = false;
        pthread_cond_broadcast(&g_oq_cv);
    }
    pthread_mutex_unlock(&g_oq_lock);
    return me.rc;
}

extern "C" phip_dprofile_t *phip_profile_upload(const uint8_t *q, int m, const phip_matrix_t *mx)
{
    if (phip_init() != 0) return nullptr;
    if (!q || m <= 0 || !mx || !mx->values || mx->size <= 0 || mx->length <= 0) {
        phip_set_error("phip_profile_upload: bad arguments");
        return nullptr;
    }
    const size_t qb = al256((size_t)m), mb = (size_t)mx->length * mx->size * 4;
    phip_dprofile_t *p = (phip_dprofile_t *)calloc(1, sizeof(phip_dprofile_t));
    char *d = nullptr;
    if (!p || hipMalloc((void **)&d, qb + mb) != hipSuccess) {
        phip_set_error("phip_profile_upload: out of memory (%zu bytes)", qb + mb);
        free(p);
        return nullptr;
    }
    if (hipMemcpy(d, q, (size_t)m, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d + qb, mx->values, mb, hipMemcpyHostToDevice) != hipSuccess) {
        phip_set_error("phip_profile_upload: upload failed");
        (void)hipFree(d);
        free(p);
        return nullptr;
    }
    p->q = (uint8_t *)d;
    p->mat = (int *)(d + qb);
    p->m = m;
    return p;
}

extern "C" void phip_profile_release(phip_dprofile_t *p)
{
    if (!p) return;
    // hipFree waits for the device, so no launch still reads the copy
    (void)hipFree(p->q);
    free(p);
}

extern "C" int phip_align_tables(const phip_batch_t *b, int tab, phip_out_t *out, int **dst)
{
    if (phip_init() != 0) return -1;
    if (!b || !out || !dst || b->npairs != 1 || (tab != 1 && tab != 2) ||
        (b->kind != PHIP_SCORE && b->kind != PHIP_STATS) || (b->kind == PHIP_STATS && !b->q)) {
        phip_set_error("phip_align_tables: bad arguments");
        return -1;
    }
    const int m = b->qlen[0], n = b->tlen[0];
    const int nplanes = b->kind == PHIP_STATS ? 4 : 1;
    const size_t per = tab == 1 ? (size_t)m * n : (size_t)m + n;
    const int64_t zero = 0;
    HostBatch hb;
    hb.np = 1;
    hb.qlen = &m;
    hb.tlen = &n;
    hb.qoff = &zero;
    hb.toff = &zero;
    hb.mx = &b->matrix;
    const std::vector<int> pairs(1, 0);
    const int32_t status = PHIP_ST_W32;
    Dp32Plan P;
    // the int32 lane chain when the pair fits it, else the one-lane tier
    bool wave = b->q && !b->matrix.is_pssm && b->matrix.size <= 32 &&
                (b->kind != PHIP_STATS || (std::min(m, n) <= 1023 && m + n <= 4095)) &&
                dp32_plan(hb, pairs, &status, &zero, P, 64, true);
    pthread_mutex_lock(&g_lock);
    int rc = -1;
    phip_resident_t *r = resident_create(b);
    char *tabs = nullptr, *plan = nullptr;
    int *scratch = nullptr;
    do {
        if (!r) break;
        tabs = (char *)dev_alloc(per * nplanes * sizeof(int));
        if (!tabs) break;
        DPParams dp = make_params(r, b->alg, b->kind, b->sg_flags, b->open, b->gap,
                                  b->want_sat_check);
        dp.tab = tab;
        for (int k = 0; k < 4; ++k)
            dp.tabs[k] = k < nplanes ? (int *)(tabs + (size_t)k * per * sizeof(int)) : nullptr;
        if (hipEventRecord(g_ev0, g_stream) != hipSuccess) break;
        if (wave) {
            plan = (char *)dev_alloc(P.bytes);
            if (!plan) break;
            if (hipMemcpyAsync(plan, P.image.data(), P.bytes, hipMemcpyHostToDevice, g_stream) !=
                hipSuccess) {
                phip_set_error("phip_align_tables: plan upload failed");
                break;
            }
            if (dp32_launch(dp, P, plan, g_stream) != 0) break;
        } else {
            const int narr = b->kind == PHIP_STATS ? 8 : 2;
            scratch = (int *)dev_alloc((size_t)(n + 1) * narr * sizeof(int) * 64);
            if (!scratch) break;
            hipLaunchKernelGGL(k_simple, dim3(1), dim3(64), 0, g_stream, dp, 0, 1,
                               (const int32_t *)nullptr, (const int *)nullptr, scratch, 64, n);
            if (hipGetLastError() != hipSuccess) {
                phip_set_error("phip_align_tables: launch failed");
                break;
            }
        }
        if (hipEventRecord(g_ev1, g_stream) != hipSuccess) break;
        if (resident_fetch(r, out, nullptr) != 0) break;
        bool ok = true;
        for (int k = 0; k < nplanes; ++k)
            ok &= hipMemcpy(dst[k], tabs + (size_t)k * per * sizeof(int), per * sizeof(int),
                            hipMemcpyDeviceToHost) == hipSuccess;
        if (!ok) {
            phip_set_error("phip_align_tables: download failed");
            break;
        }
        rc = 0;
    } while (0);
    if (scratch) dev_free(scratch);
    if (plan) dev_free(plan);
    if (tabs) dev_free(tabs);
    if (r) resident_free(r);
    pthread_mutex_unlock(&g_lock);
    return rc;
}

extern "C" int phip_align_cigars(const phip_batch_t *b, const uint8_t *qraw, const uint8_t *traw,
                                 phip_out_t *out, phip_cigars_t *cig)
{
    if (phip_init() != 0) return -1;
    if (!b || !out || !cig || !qraw || !traw || b->kind != PHIP_TRACE || b->npairs < 0) {
        phip_set_error("phip_align_cigars: bad arguments");
        return -1;
    }
    const int np = b->npairs;
    if (np == 0) return 0;
    pthread_mutex_lock(&g_lock);
    int rc = -1;
    phip_resident_t *r = resident_create(b);
    char *mem = nullptr;
    do {
        if (!r) break;
        // raw residues (upper-cased by the caller) and the op buffers; the
        // uploads run while the host plans the DP below
        int64_t qtot = 0, ttot = 0, otot = 0;
        std::vector<int64_t> opoff(np);
        for (int i = 0; i < np; ++i) {
            qtot = std::max<int64_t>(qtot, b->qoff[i] + b->qlen[i]);
            ttot = std::max<int64_t>(ttot, b->toff[i] + b->tlen[i]);
            opoff[i] = otot;
            otot += (int64_t)b->qlen[i] + b->tlen[i] + 2;
        }
        size_t off = 0;
        auto slot = [&](size_t bytes) { size_t o = off; off += al256(bytes ? bytes : 1); return o; };
        const size_t o_q = slot(qtot), o_t = slot(ttot), o_oo = slot((size_t)np * 8),
                     o_ops = slot((size_t)otot * 4), o_n = slot((size_t)np * 4),
                     o_b = slot((size_t)np * 8);
        mem = (char *)dev_alloc(off);
        if (!mem) break;
        bool ok = hipMemcpyAsync(mem + o_q, qraw, qtot, hipMemcpyHostToDevice, g_stream) == hipSuccess;
        ok &= hipMemcpyAsync(mem + o_t, traw, ttot, hipMemcpyHostToDevice, g_stream) == hipSuccess;
        ok &= hipMemcpyAsync(mem + o_oo, opoff.data(), (size_t)np * 8, hipMemcpyHostToDevice,
                             g_stream) == hipSuccess;
        if (!ok) { phip_set_error("phip_align_cigars: upload failed"); break; }
        if (resident_run(r, b->alg, PHIP_TRACE, b->sg_flags, b->open, b->gap, b->want_sat_check) != 0)
            break;
        CigarArgs a;
        a.qraw = (const uint8_t *)(mem + o_q);
        a.traw = (const uint8_t *)(mem + o_t);
        a.qoff = r->dqoff;
        a.toff = r->dtoff;
        a.qlen = r->dqlen;
        a.tlen = r->dtlen;
        a.trace = r->dtr;
        a.troff = r->dtroff;
        a.tlay = r->dtlay;
        a.out = r->dout;
        a.is_sg = b->alg == PHIP_SG;
        a.npairs = np;
        a.ops = (uint32_t *)(mem + o_ops);
        a.opoff = (const int64_t *)(mem + o_oo);
        a.nops = (int32_t *)(mem + o_n);
        a.beg = (int32_t *)(mem + o_b);
        if (cigar_launch(a, g_stream) != 0) break;
        // the call's device time: DP kernels and the walk
        if (hipEventRecord(g_ev1, g_stream) != hipSuccess) break;
        if (resident_fetch(r, out, nullptr) != 0) break;
        // downloads land in the pinned bounce buffer (DMA), then in the
        // caller's arrays
        const size_t a_beg = al256((size_t)np * 4);
        if (reserve_up(a_beg + (size_t)np * 8) == 0) {
            ok = hipMemcpyAsync(g_up, a.nops, (size_t)np * 4, hipMemcpyDeviceToHost, g_stream) == hipSuccess;
            ok &= hipMemcpyAsync(g_up + a_beg, a.beg, (size_t)np * 8, hipMemcpyDeviceToHost, g_stream) ==
                  hipSuccess;
            ok &= hipStreamSynchronize(g_stream) == hipSuccess;
            if (ok) {
                memcpy(cig->nops, g_up, (size_t)np * 4);
                memcpy(cig->beg, g_up + a_beg, (size_t)np * 8);
            }
        } else {
            ok = hipMemcpy(cig->nops, a.nops, (size_t)np * 4, hipMemcpyDeviceToHost) == hipSuccess;
            ok &= hipMemcpy(cig->beg, a.beg, (size_t)np * 8, hipMemcpyDeviceToHost) == hipSuccess;
        }
        if (!ok) { phip_set_error("phip_align_cigars: download failed"); break; }
        // the ops reserve m + n + 2 slots a pair and use a few: gather the
        // used ones back to back on the device and download only those
        int64_t ctot = 0;
        for (int i = 0; i < np; ++i) {
            cig->off[i] = ctot;
            ctot += cig->nops[i] > 0 ? cig->nops[i] : 0;
        }
        if (ctot > 0) {
            int64_t *dco = (int64_t *)dev_alloc((size_t)np * 8);
            uint32_t *dcomp = (uint32_t *)dev_alloc((size_t)ctot * 4);
            ok = dco && dcomp;
            ok = ok && hipMemcpyAsync(dco, cig->off, (size_t)np * 8, hipMemcpyHostToDevice, g_stream) ==
                           hipSuccess;
            ok = ok && cigar_compact_launch(a.ops, a.opoff, a.nops, dco, np, dcomp, g_stream) == 0;
            const bool bounce = ok && reserve_up((size_t)ctot * 4) == 0;
            ok = ok && hipMemcpyAsync(bounce ? (void *)g_up : (void *)cig->ops, dcomp,
                                      (size_t)ctot * 4, hipMemcpyDeviceToHost, g_stream) == hipSuccess;
            ok = ok && hipStreamSynchronize(g_stream) == hipSuccess;
            if (ok && bounce) memcpy(cig->ops, g_up, (size_t)ctot * 4);
            if (dco) dev_free(dco);
            if (dcomp) dev_free(dcomp);
            if (!ok) { phip_set_error("phip_align_cigars: download failed"); break; }
        }
        rc = 0;
    } while (0);
    if (mem) dev_free(mem);
    if (r) resident_free(r);
    pthread_mutex_unlock(&g_lock);
    return rc;
}

extern "C" void *phip_resident_create(const phip_batch_t *b)
{
    if (phip_init() != 0) return nullptr;
    if (!b || b->npairs <= 0) { phip_set_error("phip_resident_create: bad arguments"); return nullptr; }
    pthread_mutex_lock(&g_lock);
    phip_resident_t *r = resident_create(b);
    if (r && resident_host_buffers(r) != 0) {
        resident_free(r);
        r = nullptr;
    }
    pthread_mutex_unlock(&g_lock);
    return r;
}

// ---------------------------------------------------------------- staging

namespace {
// One pinned host buffer for the per-call batch entry points: residues are
// encoded straight into it and uploaded by DMA from there (a pageable
// upload is staged by the runtime in small pieces, and fresh host buffers
// cost a page fault per page).  It grows to the largest batch seen, up to
// kStageMax, and is kept for the process; one caller holds it at a time.
constexpr size_t kStageMax = (size_t)1 << 30;
pthread_mutex_t g_stage_mu = PTHREAD_MUTEX_INITIALIZER;
char *g_stage = nullptr;
size_t g_stage_cap = 0;
}  // namespace

extern "C" void *phip_stage_acquire(size_t bytes)
{
    if (bytes == 0 || bytes > kStageMax || phip_init() != 0) return nullptr;
    pthread_mutex_lock(&g_stage_mu);
    if (bytes > g_stage_cap) {
        const size_t cap = std::min(kStageMax, std::max(bytes + (bytes >> 2), (size_t)16 << 20));
        if (g_stage) (void)hipHostFree(g_stage);
        g_stage = nullptr;
        g_stage_cap = 0;
        if (hipHostMalloc((void **)&g_stage, cap, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            g_stage = nullptr;
            pthread_mutex_unlock(&g_stage_mu);
            return nullptr;
        }
        g_stage_cap = cap;
    }
    return g_stage;
}

extern "C" void phip_stage_release(void *p)
{
    if (p) pthread_mutex_unlock(&g_stage_mu);
}

__global__ void k_pack_out4(const phip_out_t *out, phip_out4_t *o4, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const phip_out_t o = out[i];
    o4[i] = phip_out4_t{o.score, o.end_query, o.end_ref, o.status};
}

extern "C" int phip_resident_align4(void *handle, int alg, int sg_flags, int open, int gap,
                                    const phip_out4_t **res)
{
    phip_resident_t *r = (phip_resident_t *)handle;
    if (!r || !res) { phip_set_error("phip_resident_align4: bad arguments"); return -1; }
    pthread_mutex_lock(&g_lock);
    int rc = resident_host_buffers(r);
    if (rc == 0) rc = resident_run(r, alg, PHIP_SCORE, sg_flags, open, gap, 0);
    do {
        if (rc != 0) break;
        rc = -1;
        hipLaunchKernelGGL(k_pack_out4, dim3((r->np + 255) / 256), dim3(256), 0, g_stream, r->dout,
                           r->dout4, r->np);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(r->hout4, r->dout4, (size_t)r->np * sizeof(phip_out4_t),
                           hipMemcpyDeviceToHost, g_stream) != hipSuccess ||
            hipStreamSynchronize(g_stream) != hipSuccess) {
            phip_set_error("phip_resident_align4: download failed");
            break;
        }
        bool bad = false;
        for (int i = 0; i < r->np && !bad; ++i)
            if (r->hout4[i].status & PHIP_ST_ERR) {
                phip_set_error("pair %d could not be computed (long-pair pipeline failed)", i);
                bad = true;
            }
        if (bad) break;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, g_ev0, g_ev1) != hipSuccess) break;
        g_last_ms = ms;
        g_last_main_ms = r->last_wave ? wave_last_kernel_ms() : ms;
        *res = r->hout4;
        rc = 0;
    } while (0);
    pthread_mutex_unlock(&g_lock);
    return rc;
}

// Several score passes over one resident batch in one call: every pass is
// launched (kernels, packing, download into its own pinned buffer) before
// the host waits for the first, so the host's work on a pass's results
// overlaps the next passes' kernels.  phip_resident_launch4 takes the device
// lock and phip_resident_end4 releases it; in between the caller collects
// each pass with phip_resident_wait4.
extern "C" int phip_resident_launch4(void *handle, int npass, const int *alg, const int *sg_flags,
                                     int open, int gap)
{
    phip_resident_t *r = (phip_resident_t *)handle;
    if (!r || npass < 1 || npass > phip_resident_t::kPasses || !alg || !sg_flags) {
        phip_set_error("phip_resident_launch4: bad arguments");
        return -1;
    }
    pthread_mutex_lock(&g_lock);
    int rc = 0;
    for (int k = 0; k < npass && rc == 0; ++k) {
        if (!r->pd4[k]) r->pd4[k] = (phip_out4_t *)dev_alloc((size_t)r->np * sizeof(phip_out4_t));
        if (!r->ph4[k] && hipHostMalloc((void **)&r->ph4[k], (size_t)r->np * sizeof(phip_out4_t),
                                        hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            r->ph4[k] = nullptr;
        }
        for (int e = 0; e < 5 && rc == 0; ++e)
            if (!r->pev[k][e] && hipEventCreate(&r->pev[k][e]) != hipSuccess) rc = -1;
        if (rc || !r->pd4[k] || !r->ph4[k]) {
            phip_set_error("phip_resident_launch4: out of memory");
            rc = -1;
        }
    }
    if (rc) {
        pthread_mutex_unlock(&g_lock);
        return -1;
    }
    int key[3 + 2 * phip_resident_t::kPasses] = {npass, open, gap};
    for (int k = 0; k < npass; ++k) {
        key[3 + 2 * k] = alg[k];
        key[4 + 2 * k] = sg_flags[k];
    }
    const bool same = r->gkey_set && memcmp(key, r->gkey, sizeof(key)) == 0;
    if (!same && r->gexec) {
        (void)hipGraphExecDestroy(r->gexec);
        (void)hipGraphDestroy(r->graph);
        r->gexec = nullptr;
        r->graph = nullptr;
    }
    memcpy(r->gkey, key, sizeof(key));
    r->gkey_set = true;
    if (r->graph_on && same && r->gexec) {
        // the same passes as the captured call: one graph launch
        if (hipGraphLaunch(r->gexec, g_stream) != hipSuccess) {
            phip_set_error("phip_resident_launch4: graph launch failed");
            pthread_mutex_unlock(&g_lock);
            return -1;
        }
        memcpy(r->pass_wave, r->gwave, sizeof(r->gwave));
        r->graph_replays++;
        r->npass = npass;
        return 0;
    }
    // the same passes as the previous call, all on the wavefront tier (their
    // plans exist and own every buffer the launches touch, so the passes are
    // launches and copies only): capture them into a graph
    bool all_wave = true;
    for (int k = 0; k < npass; ++k) all_wave = all_wave && r->pass_wave[k];
    const bool capture = r->graph_on && same && all_wave &&
                         hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
    rc = resident_enqueue4(r, npass, alg, sg_flags, open, gap);
    if (capture) {
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(g_stream, &g);
        hipGraphExec_t ge = nullptr;
        if (rc == 0 && ec == hipSuccess && g &&
            hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) == hipSuccess &&
            hipGraphLaunch(ge, g_stream) == hipSuccess) {
            r->graph = g;
            r->gexec = ge;
            memcpy(r->gwave, r->pass_wave, sizeof(r->gwave));
            r->graph_captures++;
        } else {
            // a step the graph cannot hold: run the passes as launches
            if (ge) (void)hipGraphExecDestroy(ge);
            if (g) (void)hipGraphDestroy(g);
            (void)hipGetLastError();
            r->graph_on = false;
            rc = resident_enqueue4(r, npass, alg, sg_flags, open, gap);
        }
    }
    if (rc) {
        (void)hipStreamSynchronize(g_stream);
        pthread_mutex_unlock(&g_lock);
        return -1;
    }
    r->npass = npass;
    return 0;
}

// The passes of phip_resident_launch4 on g_stream: each pass's kernels, its
// results packed and downloaded, and its events.  Launches and copies only.
static int resident_enqueue4(phip_resident_t *r, int npass, const int *alg, const int *sg_flags,
                             int open, int gap)
{
    int rc = 0;
    for (int k = 0; k < npass && rc == 0; ++k) {
        hipEvent_t *ev = r->pev[k];
        // the pass's events are recorded as nodes of a captured graph too
        // (external records), so every replay records them again
        if (hipEventRecordWithFlags(ev[0], g_stream, hipEventRecordExternal) != hipSuccess) { rc = -1; break; }
        // the main-kernel events stay unrecorded unless the wavefront runs
        rc = resident_run(r, alg[k], PHIP_SCORE, sg_flags[k], open, gap, 0, ev[2], ev[3]);
        if (rc) break;
        r->pass_wave[k] = r->last_wave;
        hipLaunchKernelGGL(k_pack_out4, dim3((r->np + 255) / 256), dim3(256), 0, g_stream, r->dout,
                           r->pd4[k], r->np);
        if (hipGetLastError() != hipSuccess ||
            hipEventRecordWithFlags(ev[1], g_stream, hipEventRecordExternal) != hipSuccess ||
            hipMemcpyAsync(r->ph4[k], r->pd4[k], (size_t)r->np * sizeof(phip_out4_t),
                           hipMemcpyDeviceToHost, g_stream) != hipSuccess ||
            hipEventRecordWithFlags(ev[4], g_stream, hipEventRecordExternal) != hipSuccess) {
            phip_set_error("phip_resident_launch4: launch failed");
            rc = -1;
        }
    }
    return rc;
}

extern "C" int phip_resident_set_graph(void *handle, int on)
{
    phip_resident_t *r = (phip_resident_t *)handle;
    if (!r) return -1;
    pthread_mutex_lock(&g_lock);
    const int prev = r->graph_on ? 1 : 0;
    r->graph_on = on != 0;
    pthread_mutex_unlock(&g_lock);
    return prev;
}

extern "C" int phip_resident_graph_stats(void *handle, int *captures, int *replays)
{
    phip_resident_t *r = (phip_resident_t *)handle;
    if (!r) return -1;
    if (captures) *captures = r->graph_captures;
    if (replays) *replays = r->graph_replays;
    return 0;
}

extern "C" int phip_resident_wait4(void *handle, int k, const phip_out4_t **res, double *kernel_ms,
                                   double *main_ms)
{
    phip_resident_t *r = (phip_resident_t *)handle;
    if (!r || k < 0 || k >= r->npass || !res) {
        phip_set_error("phip_resident_wait4: bad arguments");
        return -1;
    }
    if (hipEventSynchronize(r->pev[k][4]) != hipSuccess) {
        phip_set_error("phip_resident_wait4: download failed");
        return -1;
    }
    for (int i = 0; i < r->np; ++i)
        if (r->ph4[k][i].status & PHIP_ST_ERR) {
            phip_set_error("pair %d could not be computed (long-pair pipeline failed)", i);
            return -1;
        }
    float ms = 0.f, mm = 0.f;
    if (hipEventElapsedTime(&ms, r->pev[k][0], r->pev[k][1]) != hipSuccess) ms = 0.f;
    if (!r->pass_wave[k] || hipEventElapsedTime(&mm, r->pev[k][2], r->pev[k][3]) != hipSuccess) mm = ms;
    if (kernel_ms) *kernel_ms = ms;
    if (main_ms) *main_ms = mm;
    g_last_ms = ms;
    g_last_main_ms = mm;
    *res = r->ph4[k];
    return 0;
}

extern "C" void phip_resident_end4(void *handle)
{
    phip_resident_t *r = (phip_resident_t *)handle;
    (void)hipStreamSynchronize(g_stream);
    if (r) r->npass = 0;
    pthread_mutex_unlock(&g_lock);
}

extern "C" const phip_out_t *phip_resident_results(void *handle)
{
    return handle ? ((phip_resident_t *)handle)->hout : nullptr;
}

extern "C" int phip_resident_align(void *handle, int alg, int sg_flags, int open, int gap,
                                   phip_out_t *out)
{
    phip_resident_t *r = (phip_resident_t *)handle;
    if (!r) { phip_set_error("phip_resident_align: no batch"); return -1; }
    if (!out) out = r->hout;
    if (!out) { phip_set_error("phip_resident_align: no result buffer"); return -1; }
    pthread_mutex_lock(&g_lock);
    int rc = resident_run(r, alg, PHIP_SCORE, sg_flags, open, gap, 0);
    if (rc == 0) rc = resident_fetch(r, out, nullptr);
    pthread_mutex_unlock(&g_lock);
    return rc;
}

extern "C" void phip_resident_free(void *handle)
{
    pthread_mutex_lock(&g_lock);
    resident_free((phip_resident_t *)handle);
    pthread_mutex_unlock(&g_lock);
}

extern "C" int phip_nw_banded(const uint8_t *s1, int m, const uint8_t *s2, int n, int open, int gap,
                              int k, const phip_matrix_t *mx, phip_out_t *out)
{
    if (!s1 || !s2 || m <= 0 || n <= 0 || k <= 0 || !mx || !out) {
        phip_set_error("phip_nw_banded: bad arguments");
        return -1;
    }
    const int64_t zero = 0;
    phip_batch_t b;
    memset(&b, 0, sizeof(b));
    b.alg = PHIP_NW;
    b.open = open;
    b.gap = gap;
    b.matrix = *mx;
    b.npairs = 1;
    b.q = s1;
    b.qoff = &zero;
    b.qlen = &m;
    b.t = s2;
    b.toff = &zero;
    b.tlen = &n;
    return phip_nw_banded_batch(&b, k, out);
}

extern "C" int phip_device_info(int *cus, int *clock_khz)
{
    if (phip_init() != 0) return -1;
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, 0));
    if (cus) *cus = prop.multiProcessorCount;
    if (clock_khz) *clock_khz = prop.clockRate;
    return 0;
}
