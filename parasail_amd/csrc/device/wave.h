// wave.h -- packed-int16 wavefront tier (score-only SW / NW / SG): the
// device-resident plan that binds a batch of pairs to the kernels.
#pragma once

#include <vector>

#include "common.h"

// Host view of a batch (lengths and offsets) used for planning.
struct HostBatch {
    int np;
    const int32_t *qlen, *tlen;
    const int64_t *qoff, *toff;
    const phip_matrix_t *mx;
};

// Device pool: reuses blocks across calls so per-call entry points do not
// pay hipMalloc/hipFree.
void *dev_alloc(size_t bytes);
void dev_free(void *p);

// The order that sorts 32-bit keys ascending, ties in index order (a
// stable radix sort in two 16-bit digits: a comparison sort of a large
// batch, most of whose keys tie, costs milliseconds).
std::vector<int> radix_order(const std::vector<uint32_t> &keys);

struct WavePlan;

// Plans and uploads everything that depends only on the batch and the
// scoring (jobs, column streams, LDS table, output slots).  Pairs that do
// not provably fit int16 are listed for the int32 tier.  NULL on failure.
WavePlan *wave_plan_create(const HostBatch &hb, const DPParams &dp, hipStream_t s);
// Launches the plan asynchronously on s; results land in dp.out.
// e0/e1 (optional): also recorded around the main wavefront launch
int wave_plan_run(WavePlan *p, const DPParams &dp, hipStream_t s, hipEvent_t e0 = nullptr,
                  hipEvent_t e1 = nullptr);
void wave_plan_free(WavePlan *p);
bool wave_plan_matches(const WavePlan *p, int alg, int sg_flags, int open, int gap, int check);
// Whether every pair of the batch goes to the packed-int16 kernel (none to
// the int32 tiers).
bool wave_covers(const HostBatch &hb, const DPParams &dp);
// Device time (ms, HIP events on the launch stream) of the wavefront kernel
// of the last wave_plan_run, once that stream has been synchronised; 0 if
// the run launched none.
double wave_last_kernel_ms();

// Long-pair global alignment (nwlong.hip): whether a pair goes there; the
// plan (work items of a device queue: groups of bands, the groups of a pair
// consecutive) and the device region it needs, whose first lp.items.size()
// bytes the caller uploads from lp.items; the launch.
struct LongPlan {
    std::vector<char> items;
    size_t bytes = 0, lds = 0;
    size_t o_items = 0, o_prog = 0, o_gbuf = 0;
    int nitems = 0, grid = 0;
    bool rp = false;  // scores in registers (small alphabet)
};
bool nwlong_eligible(const phip_matrix_t &mx, int alg, int kind, int check, int m, int n);
bool nwlong_plan(const HostBatch &hb, const std::vector<int> &pairs, int open, LongPlan &lp);
int nwlong_launch(const DPParams &dp, const LongPlan &lp, char *region, hipStream_t st);
