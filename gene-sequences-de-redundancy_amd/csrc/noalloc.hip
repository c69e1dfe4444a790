// Clustering without any device allocation (gsdr_cluster_noalloc).
//
// gsdr_cluster runs in a launch workspace the context allocates and grows
// with hipMalloc: the per-tile-shape alignment lists, the shape counters,
// the statistics counters, the device job table and the running-jobs
// counter. Callers that cannot allocate (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc) get the same clustering
// with that workspace in `scratch`, a file-scope device array of fixed size
// that every GPU the library runs on holds until the process ends. Its
// address on a device is looked up once, with that device current, when a
// context opens there (scratch_setup, called by gsdr_open); host code uses
// that address, never the array's name, for kernel arguments, clears and
// copies. Every call on one GPU shares that GPU's array and leaves the last
// call's contents behind: the drivers clear the counters and statistics and
// rewrite the job table at the start of each call, and two calls that
// overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "driver.hpp"
#include "kernels_impl.hpp"

namespace gsdr {
namespace {

// Layout of the array: job table, list counters, statistics, running-jobs
// counter, then the kNumShapes alignment lists of equal capacity and the
// split list (one job's phase-A pairs).
constexpr size_t kJobsOff = 0, kCountersOff = 256, kStatsOff = 512;
constexpr size_t kLeftOff = kStatsOff + kStatWords * sizeof(unsigned long long), kItemsOff = kLeftOff + 256;
static_assert(sizeof(JobDev) <= kCountersOff - kJobsOff, "job table slot");
static_assert(kNumCounters * sizeof(int32_t) <= kStatsOff - kCountersOff, "counter slot");
static_assert(kStatWords * sizeof(unsigned long long) <= kLeftOff - kStatsOff, "statistics slot");

__device__ __align__(256) unsigned char scratch[GSDR_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices] = {};

}  // namespace

int scratch_setup(int device, void** ptr) {
  if (device < 0 || device >= kMaxDevices) return fail(GSDR_E_LIMIT, "device index above 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[device]) {
    // the caller made `device` current: each GPU has its own copy of the
    // array at its own address
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(GSDR_E_DEVICE, "scratch array has no device address");
    g_scratch_addr[device] = p;
  }
  *ptr = g_scratch_addr[device];
  return GSDR_OK;
}

int cluster_noalloc(gsdr_ctx* c, gsdr_store& s, const gsdr_params& p, int32_t* out, gsdr_stats* stats) {
  if (!c->scratch) return fail(GSDR_E_DEVICE, "the scratch address of this device was not looked up");
  if (p.timing) return fail(GSDR_E_INVALID, "timing needs HIP events, which this entry does not create");
  if ((int64_t)s.n > GSDR_NOALLOC_MAX_SEQS)
    return fail(GSDR_E_LIMIT, "gsdr_cluster_noalloc takes at most GSDR_NOALLOC_MAX_SEQS (600,000) sequences; "
                              "use gsdr_cluster for larger stores");
  unsigned char* base = static_cast<unsigned char*>(c->scratch);
  Workspace w;
  w.jobs = reinterpret_cast<JobDev*>(base + kJobsOff);
  w.jobs_cap = 1;
  w.counters = reinterpret_cast<int32_t*>(base + kCountersOff);
  w.stats = reinterpret_cast<unsigned long long*>(base + kStatsOff);
  w.jobs_left = reinterpret_cast<int32_t*>(base + kLeftOff);
  w.items = reinterpret_cast<AlignItem*>(base + kItemsOff);
  constexpr size_t kListBytes = GSDR_NOALLOC_SCRATCH_BYTES - kItemsOff - kSplitPerJob * sizeof(AlignItem);
  w.item_cap = (int64_t)(kListBytes / (sizeof(AlignItem) * kNumShapes));
  w.split_cap = kSplitPerJob;
  w.bytes = GSDR_NOALLOC_SCRATCH_BYTES;
  static_assert(kListBytes / (sizeof(AlignItem) * kNumShapes) >= (size_t)GSDR_NOALLOC_MAX_SEQS,
                "every sequence fits each shape list once");
  if (p.mode == GSDR_MODE_PLAIN) return cluster_plain(c, s, p, out, stats, &w);
  if (p.mode != GSDR_MODE_BATCHED) return fail(GSDR_E_INVALID, "unknown mode");
  gsdr_store* sp = &s;
  return cluster_batched(c, 1, &sp, &p, &out, stats, &w);
}

}  // namespace gsdr
