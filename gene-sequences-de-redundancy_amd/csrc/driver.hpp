// Host-side objects behind the opaque C ABI handles, and the drivers that
// launch the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/gsdr.h"
#include "host.hpp"
#include "kernels.hpp"
#include "status.hpp"

struct gsdr_records {
  gsdr::Records r;
};

// Launch workspace shared by every clustering call of a context: per-shape
// alignment lists, counters, statistics and the device job table.
struct Workspace {
  gsdr::AlignItem* items = nullptr;
  int64_t item_cap = 0;            // items per shape
  int64_t split_cap = 0;           // items of the split list (after the shape lists)
  int32_t* counters = nullptr;     // [kNumShapes] bin counts, [kNumShapes] work counters
  unsigned long long* stats = nullptr;
  gsdr::JobDev* jobs = nullptr;    // device job table
  int32_t jobs_cap = 0;
  int32_t* jobs_left = nullptr;
  int64_t bytes = 0;
  int64_t generation = 0;          // changes with every (re)allocation
};

// Launch knobs that change speed, never results (gsdr_tuning in the C ABI);
// gsdr_autotune measures them on a store.
struct Tuning {
  int swap_ratio = gsdr::kSwapRatio;
  int spec_max = gsdr::kMaxSpec;
  int filter_blocks = -1;  // -1: by the call (filter_blocks_for)
  int align_blocks = -1;  // -1: by the call (align_blocks_for); 0: each shape kernel's resident blocks per CU
  int fork_shapes = -1;  // tile-shape kernels concurrent on side streams (1), serial (0), by the call (-1)
  int list_factor = 64;  // alignment list capacity per shape: list_factor items per sequence
  int pigeon = 0;        // exact pigeonhole prefilter between the filter and the alignment
  int screen = 1;        // word-signature screen before the phase-B filter (k >= 6)
  gsdr::ShapeTable shapes = {};  // tile shapes per text-length bin (zero: the narrowest that holds the band)
  int win_period = 0;    // 64-lane tiles: blocks between DP-edge window rebases (0 = by width)
  int split_rows = -1;   // pairs with this many rows run on two waves (0 = never, -1 = by the call)
};

struct gsdr_ctx {
  int device = 0;
  Tuning tuned;                       // used by calls with params.tune = 1
  int64_t ws_generations = 0;         // workspace allocations so far
  bool has_tuning = false;
  hipStream_t stream = nullptr;       // kernels
  hipStream_t upload = nullptr;       // store uploads (may overlap kernels of another job)
  int cu_count = 0;
  int64_t hbm_bytes = 0;
  char name[256] = {0};
  gsdr::JobState* h_state = nullptr;  // pinned host copies of job states
  int32_t h_state_cap = 0;
  int32_t* h_left = nullptr;          // pinned copy of jobs_left
  Workspace ws;
  // side streams (default priority) that run the tile-shape kernels of a
  // step concurrently, with fork/join events
  hipStream_t side[gsdr::kNumShapes + 1] = {};  // the last one runs the split list
  hipEvent_t fork = nullptr;
  hipEvent_t join[gsdr::kNumShapes + 1] = {};
  // device address of this device's no-allocation scratch array
  // (gsdr_cluster_noalloc), looked up when the context opens
  void* scratch = nullptr;
};

struct gsdr_store {
  gsdr_ctx* ctx = nullptr;
  int32_t n = 0;
  int k = 0;
  int64_t bases = 0;
  int32_t max_m = 0;
  std::vector<int32_t> raw_len, m;  // host copies
  // HBM-resident store
  uint2* pk = nullptr;
  int64_t* pkoff = nullptr;
  int32_t* m_d = nullptr;
  uint4* comp = nullptr;
  uint32_t* kp = nullptr;
  int64_t* kpoff = nullptr;
  int32_t* bc = nullptr;
  int32_t* wc = nullptr;
  int32_t* raw_len_d = nullptr;     // raw lengths (cutoffs are computed from them on the device)
  int32_t* ksum = nullptr;          // per sequence: sum of its stored k-mer counts
  gsdr::SeqRec* rec = nullptr;      // per sequence: the filter's packed record
  uint32_t* sig = nullptr;          // word-presence signatures (k >= 6), else null
  unsigned long long* scr = nullptr;  // screen survivors per sequence (with sig)
  unsigned long long* chm = nullptr;  // screen survivors per phase-B chunk (with sig)
  int32_t* tcnt = nullptr;          // resolve: survivors per tile
  int64_t kp_total = 0;             // profile entries of all sequences
  float t_loaded = -1.0f;
  // per-job round state (allocated with the store)
  int32_t* cluster = nullptr;
  int32_t* cand[2] = {nullptr, nullptr};
  unsigned long long* acc = nullptr;
  gsdr::JobState* state = nullptr;
  int64_t footprint = 0;
  std::vector<void*> allocs;

  gsdr::StoreView view() const {
    gsdr::StoreView v;
    v.pk = pk; v.pkoff = pkoff; v.m = m_d; v.comp = comp; v.kp = kp; v.kpoff = kpoff; v.bc = bc; v.wc = wc;
    v.ksum = ksum;
    v.rec = rec;
    v.sig = sig;
    return v;
  }
  gsdr::JobDev job(const gsdr_params& p) const {
    gsdr::JobDev j;
    j.flags = p.drop_filters ? gsdr::kJobDropFilters : 0;
    j.st = view();
    j.cluster = cluster;
    j.cand[0] = cand[0];
    j.cand[1] = cand[1];
    j.acc = acc;
    j.state = state;
    j.scr = scr;
    j.chm = chm;
    j.tcnt = tcnt;
    j.n = n;
    j.k = k;
    return j;
  }
};

namespace gsdr {
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);
int host_threads();

int store_build(gsdr_ctx* c, const Records& sorted, int k, gsdr_store& s);
// Device-side build of the sequence arrays of a store (prep.hip); `alloc`
// allocates device memory owned by the store.
int store_build_device(gsdr_ctx* c, const Records& sorted, int k, gsdr_store& s,
                       int (*alloc)(gsdr_store&, void**, size_t));
void store_release(gsdr_store& s);
int store_set_threshold(gsdr_store& s, float t);
int check_params(const gsdr_store& s, const gsdr_params& p, int* k_out);
int workspace_reserve(gsdr_ctx* c, int64_t item_cap, int njobs);
void workspace_release(gsdr_ctx* c);
int max_shape_for(int n);
struct AlignArgs;
// Launches the alignment kernels of shapes 0..smax. fork = 1: concurrently
// on the side streams, forked from and joined back into `main`; 0: one after
// another on `main` (no cross-queue waits).
// split (optional): the split list's arguments, run beside the shapes.
hipError_t launch_align_shapes(gsdr_ctx* c, const AlignArgs* args, int smax, const int* grids, hipStream_t main,
                               int fork = 1, const AlignArgs* split = nullptr, int split_grid = 0);

int cluster_store(gsdr_ctx* c, gsdr_store& s, const gsdr_params& p, int32_t* out, gsdr_stats* stats);
// fixed != nullptr: run in that workspace (the no-allocation entry's
// scratch) instead of the context's, without events or graphs.
int cluster_plain(gsdr_ctx* c, gsdr_store& s, const gsdr_params& p, int32_t* out, gsdr_stats* stats,
                  const Workspace* fixed = nullptr);
int cluster_batched(gsdr_ctx* c, int njobs, gsdr_store* const* stores, const gsdr_params* params,
                    int32_t* const* out, gsdr_stats* stats, const Workspace* fixed = nullptr);
// No-allocation entry (noalloc.hip): the workspace lives in a file-scope
// device array whose per-device address scratch_setup looks up once.
int scratch_setup(int device, void** ptr);
int cluster_noalloc(gsdr_ctx* c, gsdr_store& s, const gsdr_params& p, int32_t* out, gsdr_stats* stats);
int cluster_jobs(gsdr_ctx* c, int njobs, gsdr_store* const* stores, const gsdr_params* params,
                 int32_t* const* out, gsdr_stats* stats);
int align_pairs(gsdr_ctx* c, const char* data, const int64_t* data_off, int nseq, const int32_t* pairs,
                const int32_t* bc, int npairs, int shape, int32_t* accept_out, double* cells_exec_out);
// Sums the kStatStripes copies of the device counters into ST_NUM values.
void stats_fold(const unsigned long long* all, unsigned long long* out);
void stats_from_device(const unsigned long long* d, gsdr_stats* out);
}  // namespace gsdr

#include "kernels_impl.hpp"

namespace gsdr {
// Kernel arguments of one two-phase step over the context's job table.
struct StepLaunch {
  FilterArgs fa_reps, fa_main;
  ResolveArgs ra;
  AlignArgs aa_reps[kNumShapes];
  AlignArgs aa[kNumShapes];
  AlignArgs aa_split;  // the split list (both phases)
  int split = 0;       // launch the split kernel in both phases
  int split_grid = 1;
  int smax = 0;
  unsigned kmask = 0;  // word lengths among the jobs (bit k)
  int fork = 1;  // shape kernels on side streams
  int filter_grid = 1;
  int align_grid[kNumShapes] = {};
  int reps_grid[kNumShapes] = {};
  int reps_filter_grid = 1;
  int pigeon = 0;  // pigeonhole test after each filter launch
  int pigeon_grid = 1;
  PigeonArgs pa{};
  int screen = 0;  // signature screen before the phase-B filter
  int screen_grid = 1;
  ScreenArgs sa{};
};
// How a call's steps are launched when the knobs say "by the call" (-1):
//   busy: many aligned pairs per step, i.e. at least kBusyBases in jobs of
//     threshold at most kBusyThreshold (most pairs pass the filters), and
//     no text that needs a tile wider than 64x4 (kMaxShortShape): each
//     tile-shape kernel gets its resident blocks per CU;
//   latency: otherwise, when no wide-tile kernel runs or every job has a
//     threshold of at least kLatencyThreshold (few pairs per step, many
//     steps): the tile-shape kernels run one after another with 2 blocks
//     per CU, which saves the fork/join waits across the hardware queues
//     and the launches of blocks that find no work;
//   normal: 4 blocks per CU, tile-shape kernels concurrent.
// Measured, ms per clustering (4 blocks concurrent / resident / serial with
// 2 blocks): 200,000 records of median 1,200 bases t=0.85 99.8 / 93.3 / -,
// t=0.90 550 / 493 / -, t=0.95 250 / 267 / 242, t=0.99 809 / 853 / 671;
// the 177,000-record t=0.99 job 705 / 740 / 580; 10,000 records t=0.95
// 30.3 / 32.7 / 26.2; 512 jobs of 500..10,000 records (thresholds
// 0.80..0.99) in shared launches 3,857 / 3,553 / -; 2,000,000 records
// (texts to 30,000 bases) t=0.95 375 / 377 / 571, t=0.90 259 / 263 / -,
// t=0.80 183 / 185 / -, t=0.99 9,566 / - / 9,222 (where a long-pair kernel
// ends each step, the other shapes' extra blocks share its CUs, and the
// serial form leaves the chip to that kernel alone).
enum class GridMode { kNormal, kBusy, kLatency };
constexpr int64_t kBusyBases = 32ll << 20;
constexpr float kBusyThreshold = 0.90f;
constexpr float kLatencyThreshold = 0.99f;
constexpr int kMaxShortShape = 6;  // 64x4
// low_bases: bases of the jobs of threshold <= kBusyThreshold; tmin: the
// lowest threshold of the call; smax: its widest tile shape
inline GridMode grid_mode(int64_t low_bases, float tmin, int smax) {
  if (smax <= kMaxShortShape && low_bases >= kBusyBases) return GridMode::kBusy;
  if (smax <= kMaxShortShape || tmin >= kLatencyThreshold) return GridMode::kLatency;
  return GridMode::kNormal;
}
inline int64_t low_t_bases(int64_t bases, float t) { return t <= kBusyThreshold ? bases : 0; }
// Alignment workgroups per CU of tile shape `shape`: a positive knob is the
// count, 0 each shape kernel's resident blocks per CU, -1 by the call.
inline int align_blocks_for(const Tuning& T, int shape, GridMode m) {
  if (T.align_blocks > 0) return T.align_blocks;
  if (T.align_blocks == 0 || m == GridMode::kBusy) return gsdr::align_blocks_per_cu(shape);
  return m == GridMode::kLatency ? 2 : 4;
}
// Filter workgroups per CU: a positive knob is the count; -1 is 32 in the
// normal form (long pairs: the filter's units are split statically over the
// grid, and a finer grid evens out the blocks' loads; 2,000,000 records
// t=0.95: filter 28.2 -> 23.3 ms per clustering with 32 against 8, t=0.90
// and t=0.80 unchanged) and 8 in the others (the 177,000-record t=0.99 job
// lost 14 ms with 16).
inline int filter_blocks_for(const Tuning& T, GridMode m) {
  return T.filter_blocks > 0 ? T.filter_blocks : (m == GridMode::kNormal ? 32 : 8);
}
// Tile-shape kernels concurrent on side streams (1) or one after another (0).
inline int fork_for(const Tuning& T, GridMode m) {
  return T.fork_shapes >= 0 ? T.fork_shapes : (m == GridMode::kLatency ? 0 : 1);
}
// Rows from which a pair is split over two waves: a knob of 0 or 64..32768
// is the value; -1 is 128 in the latency form, where a step ends with its
// longest pair's rows (the 177,000-record t=0.99 job 580 -> 550 ms per
// clustering, 10,000 records t=0.95 26.1 -> 24.4 ms), and never otherwise
// (2,000,000 records t=0.95 367 -> 377 ms with 2,048, 200,000 records
// t=0.90 493 -> 501 ms with 4,096: pairs past the split list's room
// contend for its counter in the filter before they go to their shapes).
constexpr int kLatencySplitRows = 128;
inline int split_rows_for(const Tuning& T, GridMode m) {
  return T.split_rows >= 0 ? T.split_rows : (m == GridMode::kLatency ? kLatencySplitRows : 0);
}
StepLaunch make_step_launch(gsdr_ctx* c, const Workspace& w, int njobs, int smax, unsigned kmask,
                            int32_t* jobs_left, GridMode mode, const Tuning& T = Tuning());
// The knobs a call runs with: the context's tuned ones when asked for and present.
inline Tuning tuning_for(const gsdr_ctx* c, bool want) { return want && c->has_tuning ? c->tuned : Tuning(); }
// The shape tables and window period of T, as the C ABI reports them.
inline void tuning_shapes_out(const Tuning& T, gsdr_tuning* out) {
  for (int b = 0; b < gsdr::kLenBins; ++b) {
    out->shape_mask[b] = T.shapes.mask[b];
    out->shape_bias[b] = T.shapes.bias[b];
  }
  out->win_period = T.win_period;
  out->split_rows = T.split_rows;
}
int autotune(gsdr_ctx* c, gsdr_store& s, const gsdr_params& p, int max_trials, gsdr_tuning* out);
hipError_t launch_step(gsdr_ctx* c, const StepLaunch& L, hipStream_t q, hipEvent_t* ev);
}  // namespace gsdr
