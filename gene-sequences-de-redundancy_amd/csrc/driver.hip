// Store construction, launch workspace and the plain per-round driver.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "driver.hpp"
#include "kernels_impl.hpp"

namespace gsdr {

#define HIP_TRY(expr, what)                          \
  do {                                               \
    hipError_t _e = (expr);                          \
    if (_e != hipSuccess) return hip_fail(_e, what); \
  } while (0)

template <typename T>
static int dev_alloc(std::vector<void*>& allocs, int64_t& footprint, T** p, size_t count) {
  size_t bytes = std::max<size_t>(count * sizeof(T), 16);
  hipError_t e = hipMalloc((void**)p, bytes);
  if (e != hipSuccess) return fail(GSDR_E_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  allocs.push_back((void*)*p);
  footprint += (int64_t)bytes;
  return GSDR_OK;
}

static int store_alloc(gsdr_store& s, void** p, size_t bytes) {
  return dev_alloc(s.allocs, s.footprint, (uint8_t**)p, bytes);
}

// The sequence arrays (pk, pkoff, m, comp, kp, kpoff) are built on the device
// from the raw bytes (prep.hip); the round state is allocated here.
int store_build(gsdr_ctx* c, const Records& sorted, int k, gsdr_store& s) {
  HIP_TRY(hipSetDevice(c->device), "hipSetDevice");
  const int64_t n = sorted.count();
  s.n = (int32_t)n;
  s.k = k;
  s.bases = (int64_t)sorted.data.size();
  int st = store_build_device(c, sorted, k, s, store_alloc);
  if (st != GSDR_OK) return st;
  auto& A = s.allocs;
  auto& F = s.footprint;
  if ((st = dev_alloc(A, F, &s.bc, (size_t)n)) || (st = dev_alloc(A, F, &s.wc, (size_t)n)) ||
      (st = dev_alloc(A, F, &s.cluster, (size_t)n)) || (st = dev_alloc(A, F, &s.cand[0], (size_t)n)) ||
      (st = dev_alloc(A, F, &s.cand[1], (size_t)n)) || (st = dev_alloc(A, F, &s.acc, (size_t)n)) ||
      (st = dev_alloc(A, F, &s.state, (size_t)1)) || (st = dev_alloc(A, F, &s.raw_len_d, (size_t)n)) ||
      (st = dev_alloc(A, F, &s.ksum, (size_t)n)) || (st = dev_alloc(A, F, &s.rec, (size_t)n)) ||
      (st = dev_alloc(A, F, &s.tcnt, (size_t)(n / kResolveTile + 2))))
    return st;
  // signature screen arrays for long words (kSigWords * 4 + 8 bytes per
  // sequence, one mask per 64 phase-B positions)
  if (k >= kScreenMinK &&
      ((st = dev_alloc(A, F, &s.sig, (size_t)n * kSigWords)) || (st = dev_alloc(A, F, &s.scr, (size_t)n)) ||
       (st = dev_alloc(A, F, &s.chm, (size_t)(n / 64 + 1)))))
    return st;
  if (s.kp_total >= (int64_t)1 << 32) return fail(GSDR_E_LIMIT, "more than 2^32 k-mer profile entries in one store");
  HIP_TRY(launch_profile_totals(s.kp, s.kpoff, (int)n, s.ksum, c->upload), "profile totals");
  HIP_TRY(launch_pack_records(s.m_d, s.comp, s.kpoff, s.ksum, (int)n, s.rec, c->upload), "pack records");
  if (s.sig) HIP_TRY(launch_signatures(s.kp, s.kpoff, (int)n, k, s.sig, c->upload), "signatures");
  // every upload goes through the upload stream: a null-stream copy here could
  // meet a graph capture that the serving loop runs on another thread
  HIP_TRY(hipMemcpyAsync(s.raw_len_d, s.raw_len.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice,
                         c->upload),
          "upload lengths");
  HIP_TRY(hipStreamSynchronize(c->upload), "profile totals");
  return GSDR_OK;
}

void store_release(gsdr_store& s) {
  if (s.ctx) (void)hipSetDevice(s.ctx->device);
  for (void* p : s.allocs) (void)hipFree(p);
  s.allocs.clear();
  s.footprint = 0;
}

int workspace_reserve(gsdr_ctx* c, int64_t item_cap, int njobs) {
  Workspace& w = c->ws;
  if (njobs > kMaxJobsPerLaunch) return fail(GSDR_E_LIMIT, "more than 1024 jobs in one launch");
  if (w.items && w.item_cap >= item_cap && w.jobs_cap >= njobs) return GSDR_OK;
  workspace_release(c);
  std::vector<void*> A;
  int64_t F = 0;
  int st;
  const int64_t cap = std::max<int64_t>(item_cap, 1024);
  // the split list follows the shape lists: every phase-A pair of every job
  const int64_t split_cap = (int64_t)kSplitPerJob * std::max(njobs, 1);
  if ((st = dev_alloc(A, F, &w.items, (size_t)cap * kNumShapes + (size_t)split_cap)) ||
      (st = dev_alloc(A, F, &w.counters, (size_t)kNumCounters)) ||
      (st = dev_alloc(A, F, &w.stats, (size_t)kStatWords)) ||
      (st = dev_alloc(A, F, &w.jobs, (size_t)std::max(njobs, 1))) || (st = dev_alloc(A, F, &w.jobs_left, (size_t)1))) {
    for (void* p : A) (void)hipFree(p);
    w = Workspace();
    return st;
  }
  w.item_cap = cap;
  w.split_cap = split_cap;
  w.jobs_cap = std::max(njobs, 1);
  w.bytes = F;
  w.generation = ++c->ws_generations;
  if (c->h_state_cap < w.jobs_cap) {
    if (c->h_state) (void)hipHostFree(c->h_state);
    HIP_TRY(hipHostMalloc((void**)&c->h_state, sizeof(JobState) * w.jobs_cap, hipHostMallocDefault), "hipHostMalloc");
    c->h_state_cap = w.jobs_cap;
  }
  return GSDR_OK;
}

void workspace_release(gsdr_ctx* c) {
  Workspace& w = c->ws;
  for (void* p : {(void*)w.items, (void*)w.counters, (void*)w.stats, (void*)w.jobs, (void*)w.jobs_left})
    if (p) (void)hipFree(p);
  w = Workspace();
}

// Cutoffs for threshold t (func.cpp:339-374), computed on the device from
// the stored raw lengths with the reference's float32 expressions.
int store_set_threshold(gsdr_store& s, float t) {
  if (s.t_loaded == t) return GSDR_OK;
  hipStream_t q = s.ctx->upload;
  HIP_TRY(launch_cutoffs(s.raw_len_d, s.n, t, s.k, s.bc, s.wc, s.rec, q), "cutoff launch");
  HIP_TRY(hipStreamSynchronize(q), "cutoff sync");
  s.t_loaded = t;
  return GSDR_OK;
}

// Option checks of func.cpp:31-51.
int check_params(const gsdr_store& s, const gsdr_params& p, int* k_out) {
  const float t = p.threshold;
  if (!(t >= 0.8f) || t >= 1.0f) return fail(GSDR_E_INVALID, "similarity out of range: 0.8<=similarity<1");
  int k = p.word_length;
  if (k == 0) k = auto_word_length(t);
  if (k < 4 || k > 8) return fail(GSDR_E_INVALID, "word length out of range: 4<=word length<=8");
  if (k != s.k) return fail(GSDR_E_INVALID, "store was built for another word length");
  if (s.max_m > GSDR_MAX_GAPFREE)
    return fail(GSDR_E_LIMIT, "a sequence exceeds 32767 gap-free bases (the representative limit)");
  *k_out = k;
  return GSDR_OK;
}

void stats_fold(const unsigned long long* all, unsigned long long* out) {
  for (int i = 0; i < ST_NUM; ++i) {
    unsigned long long v = 0;
    for (int r = 0; r < kStatStripes; ++r) v += all[r * kStatStride + i];
    out[i] = v;
  }
}

void stats_from_device(const unsigned long long* d, gsdr_stats* o) {
  o->pairs += (int64_t)d[ST_PAIRS];
  o->rej_length += (int64_t)d[ST_REJ_LEN];
  o->rej_comp += (int64_t)d[ST_REJ_COMP];
  o->rej_word += (int64_t)d[ST_REJ_WORD];
  o->aligned += (int64_t)d[ST_ALIGNED];
  o->accepted += (int64_t)d[ST_ACCEPT];
  o->rej_align += (int64_t)d[ST_REJ_ALIGN];
  o->cells_exec += (double)d[ST_CELLS_EXEC];
  o->cells_band += (double)d[ST_CELLS_BAND];
  o->cells_full += (double)d[ST_CELLS_FULL];
  o->filter_bytes += (double)d[ST_FILTER_BYTES];
  o->spec_wasted += (double)d[ST_SPEC_WASTE];
  o->cells_decided += (double)d[ST_CELLS_DECIDED];
  o->rej_pigeon += (double)d[ST_REJ_PIGEON];
  o->pigeon_checked += (double)d[ST_PIGEON_CHECKED];
  o->screened += (double)d[ST_SCREENED];
}

int cluster_store(gsdr_ctx* c, gsdr_store& s, const gsdr_params& p, int32_t* out, gsdr_stats* stats) {
  if (p.mode == GSDR_MODE_BATCHED) {
    gsdr_store* sp = &s;
    return cluster_batched(c, 1, &sp, &p, &out, stats);
  }
  if (p.mode != GSDR_MODE_PLAIN) return fail(GSDR_E_INVALID, "unknown mode");
  return cluster_plain(c, s, p, out, stats);
}

hipError_t launch_align_shapes(gsdr_ctx* c, const AlignArgs* args, int smax, const int* grids, hipStream_t main,
                               int fork, const AlignArgs* split, int split_grid) {
  hipError_t e;
  if (!fork) {
    // the split pairs are the long ones: first
    if (split && (e = launch_align_split(*split, split_grid, main)) != hipSuccess) return e;
    for (int sh = 0; sh <= smax; ++sh)
      if ((e = launch_align(sh, args[sh], grids[sh], main)) != hipSuccess) return e;
    return hipSuccess;
  }
  if ((e = hipEventRecord(c->fork, main)) != hipSuccess) return e;
  if (split) {
    hipStream_t q = c->side[kSplitList];
    if ((e = hipStreamWaitEvent(q, c->fork, 0)) != hipSuccess) return e;
    if ((e = launch_align_split(*split, split_grid, q)) != hipSuccess) return e;
    if ((e = hipEventRecord(c->join[kSplitList], q)) != hipSuccess) return e;
  }
  for (int sh = 0; sh <= smax; ++sh) {
    if ((e = hipStreamWaitEvent(c->side[sh], c->fork, 0)) != hipSuccess) return e;
    if ((e = launch_align(sh, args[sh], grids[sh], c->side[sh])) != hipSuccess) return e;
    if ((e = hipEventRecord(c->join[sh], c->side[sh])) != hipSuccess) return e;
  }
  for (int sh = 0; sh <= smax; ++sh)
    if ((e = hipStreamWaitEvent(main, c->join[sh], 0)) != hipSuccess) return e;
  if (split && (e = hipStreamWaitEvent(main, c->join[kSplitList], 0)) != hipSuccess) return e;
  return hipSuccess;
}

// Largest tile shape a representative of gap-free length n can need.
int max_shape_for(int n) {
  const int ntw = (n >> 5) + 1;
  for (int i = 0; i < kNumShapes; ++i)
    if (kShapes[i].s * kShapes[i].w >= ntw) return i;
  return kNumShapes - 1;
}

// Plain form: one representative per round (k_spec = 1). Per round the host
// launches the filter, the alignment kernels of the tile shapes this
// representative can need, and the resolve/compaction step, then reads back
// the job state (one synchronisation per round, like the reference's
// per-round waits at func.cpp:695-781 but without its host compaction).
int cluster_plain(gsdr_ctx* c, gsdr_store& s, const gsdr_params& p, int32_t* out, gsdr_stats* stats,
                  const Workspace* fixed) {
  int k = 0;
  int st = check_params(s, p, &k);
  if (st) return st;
  HIP_TRY(hipSetDevice(c->device), "hipSetDevice");
  if ((st = store_set_threshold(s, p.threshold))) return st;
  if (fixed) {
    // a round queues at most one item per candidate in each shape list
    if (p.timing) return fail(GSDR_E_INVALID, "timing needs HIP events, which this entry does not create");
    if (std::max<int64_t>(s.n, 1) > fixed->item_cap) return fail(GSDR_E_LIMIT, "too many sequences for the scratch");
  } else if ((st = workspace_reserve(c, std::max<int64_t>(s.n, 1), 1))) {
    return st;
  }
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t q = c->stream;
  const Workspace w = fixed ? *fixed : c->ws;
  const int n = s.n;
  const JobDev jd = s.job(p);
  HIP_TRY(hipMemcpyAsync(w.jobs, &jd, sizeof(JobDev), hipMemcpyHostToDevice, q), "job table");
  HIP_TRY(hipMemsetAsync(w.stats, 0, kStatWords * sizeof(unsigned long long), q), "memset");
  HIP_TRY(hipMemsetAsync(w.counters, 0, kNumCounters * sizeof(int32_t), q), "memset");
  HIP_TRY(launch_init_jobs(w.jobs, &jd, 1, 1, nullptr, q), "init");
  int64_t launches = 1;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  if (p.timing)
    for (auto& e : ev) HIP_TRY(hipEventCreate(&e), "hipEventCreate");
  double ms_filter = 0, ms_align = 0;
  int64_t filter_launches = 0, align_launches = 0, rounds = 0;
  int nc = n, first = 0;
  bool done = n == 0;
  const Tuning T = tuning_for(c, p.tune != 0);
  const GridMode mode = grid_mode(low_t_bases(s.bases, p.threshold), p.threshold, max_shape_for(s.max_m));
  while (!done) {
    ++rounds;
    const int units = (nc + kFilterChunk - 1) / kFilterChunk;
    if (p.timing) HIP_TRY(hipEventRecord(ev[0], q), "event");
    FilterArgs fa;
    fa.jobs = w.jobs;
    fa.njobs = 1;
    fa.min_shape = 0;
    fa.swap_ratio = T.swap_ratio;
    fa.shapes = T.shapes;
    fa.phase = PHASE_MAIN;
    fa.items = w.items;
    fa.item_cap = w.item_cap;
    fa.bin_count = w.counters;
    fa.stats = w.stats;
    fa.screen = T.screen && s.sig != nullptr;
    if (fa.screen) {
      const ScreenArgs sa{w.jobs, 1, w.stats};
      HIP_TRY(launch_screen(sa, std::max(1, std::min((nc + 63) / 64, c->cu_count * 8)), q), "screen launch");
      ++launches;
      ++filter_launches;
    }
    HIP_TRY(launch_filter(fa, 1u << k, std::max(1, std::min(units, c->cu_count * filter_blocks_for(T, mode))), q),
            "filter launch");
    ++launches;
    ++filter_launches;
    const int smax = max_shape_for(s.m[first]);
    if (T.pigeon) {
      PigeonArgs pa{w.jobs, w.items, w.item_cap, w.counters, smax, w.stats};
      HIP_TRY(launch_pigeon(pa, c->cu_count * 4, q), "pigeon launch");
      ++launches;
    }
    if (p.timing) HIP_TRY(hipEventRecord(ev[1], q), "event");
    AlignArgs aa[kNumShapes];
    int grids[kNumShapes];
    for (int sh = 0; sh <= smax; ++sh) {
      aa[sh].jobs = w.jobs;
      aa[sh].items = w.items + (int64_t)sh * w.item_cap;
      aa[sh].count = w.counters + sh;
      aa[sh].work = w.counters + kWorkCounter0 + sh;
      aa[sh].result = nullptr;
      aa[sh].stats = w.stats;
      aa[sh].skip_absorbed = 0;
      aa[sh].win_period = T.win_period;
      const int segs = 64 / kShapes[sh].s;
      const int waves = (nc + segs - 1) / segs;
      grids[sh] = std::max(1, std::min((waves + 3) / 4, c->cu_count * align_blocks_for(T, sh, mode)));
      ++launches;
      ++align_launches;
    }
    HIP_TRY(launch_align_shapes(c, aa, smax, grids, q, fork_for(T, mode)), "align launch");
    if (p.timing) HIP_TRY(hipEventRecord(ev[2], q), "event");
    ResolveArgs ra;
    ra.jobs = w.jobs;
    ra.njobs = 1;
    ra.counters = w.counters;
    ra.ncounters = kNumCounters;
    ra.jobs_left = nullptr;
    ra.stats = w.stats;
    ra.pair_budget = w.item_cap;
    ra.grid = std::max(1, std::min((nc + kResolveTile - 1) / kResolveTile, c->cu_count * 2));
    HIP_TRY(launch_resolve(ra, q), "resolve launch");
    launches += 3;
    HIP_TRY(hipMemcpyAsync(c->h_state, s.state, sizeof(JobState), hipMemcpyDeviceToHost, q), "state copy");
    HIP_TRY(hipStreamSynchronize(q), "round sync");
    if (p.timing) {
      float a = 0, b = 0;
      HIP_TRY(hipEventElapsedTime(&a, ev[0], ev[1]), "event time");
      HIP_TRY(hipEventElapsedTime(&b, ev[1], ev[2]), "event time");
      ms_filter += a;
      ms_align += b;
    }
    const JobState& h = c->h_state[0];
    if (!h.done && (h.nc >= nc || rounds > (int64_t)n)) return fail(GSDR_E_DEVICE, "round state did not advance");
    nc = h.nc;
    first = h.first;
    done = h.done != 0;
  }
  HIP_TRY(hipMemcpyAsync(out, s.cluster, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, q), "cluster copy");
  unsigned long long hall[kStatWords], hst[ST_NUM];
  HIP_TRY(hipMemcpyAsync(hall, w.stats, sizeof(hall), hipMemcpyDeviceToHost, q), "stats copy");
  HIP_TRY(hipStreamSynchronize(q), "final sync");
  stats_fold(hall, hst);
  if (hst[ST_OVERFLOW]) return fail(GSDR_E_LIMIT, "alignment list overflow");
  if (p.timing)
    for (auto& e : ev) (void)hipEventDestroy(e);
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats_from_device(hst, stats);
    stats->n_seqs = n;
    stats->n_bases = s.bases;
    int64_t reps = 0;
    for (int i = 0; i < n; ++i) reps += out[i] == i;
    stats->reps = reps;
    stats->rounds = rounds;
    stats->launches = launches;
    stats->ms_total = ms;
    stats->ms_filter = ms_filter;
    stats->ms_align = ms_align;
    stats->filter_launches = filter_launches;
    stats->align_launches = align_launches;
  }
  return GSDR_OK;
}

// Pair-level entry: LCS(text, query) > bc per pair, on the device.
int align_pairs(gsdr_ctx* c, const char* data, const int64_t* data_off, int nseq, const int32_t* pairs,
                const int32_t* bc, int npairs, int shape, int32_t* accept_out, double* cells_exec_out) {
  HIP_TRY(hipSetDevice(c->device), "hipSetDevice");
  Records r;
  for (int i = 0; i < nseq; ++i) r.push("", 0, data + data_off[i], data_off[i + 1] - data_off[i]);
  Prep p;
  preprocess(r, 4, host_threads(), p);
  for (int i = 0; i < npairs; ++i) {
    const int t = pairs[2 * i], qq = pairs[2 * i + 1];
    if (t < 0 || t >= nseq || qq < 0 || qq >= nseq) return fail(GSDR_E_INVALID, "pair index out of range");
    if (p.m[t] > GSDR_MAX_GAPFREE) return fail(GSDR_E_LIMIT, "text longer than 32767 gap-free bases");
    if (bc[i] < 0) return fail(GSDR_E_INVALID, "negative cutoff");
  }
  if (npairs == 0) {
    if (cells_exec_out) *cells_exec_out = 0;
    return GSDR_OK;
  }
  // shape 12 + s: split pairs starting from the window of shape s (0..11;
  // at least two words per lane)
  const bool split = shape >= kNumShapes;
  const int min_shape = shape < 0 ? 0 : std::min(split ? shape - kNumShapes : shape, kNumShapes - 1);
  // a private store: sequences, plus one query slot per pair (nseq + pair)
  // that points at the query's packed words and carries the pair's cutoff
  const int64_t nslots = (int64_t)nseq + npairs;
  std::vector<int64_t> pkoff((size_t)nslots + 1);
  std::vector<int32_t> m((size_t)nslots), bcv((size_t)nslots, 0);
  for (int i = 0; i < nseq; ++i) { pkoff[i] = p.pk_off[i]; m[i] = p.m[i]; }
  for (int i = 0; i < npairs; ++i) {
    const int qq = pairs[2 * i + 1];
    pkoff[nseq + i] = p.pk_off[qq];
    m[nseq + i] = p.m[qq];
    bcv[nseq + i] = bc[i];
  }
  std::vector<std::vector<AlignItem>> bins(kNumShapes + 1);  // the shape lists, then the split list
  std::vector<std::vector<int>> bin_pair(kNumShapes + 1);
  for (int i = 0; i < npairs; ++i) {
    const int t = pairs[2 * i];
    const int cslot = nseq + i;  // the query, carrying the pair's cutoff
    const bool sw = swap_roles(p.m[t], m[cslot]);
    const int c = bcv[cslot] + 1;
    const int sh = sw ? pick_shape(m[cslot], p.m[t], c, min_shape) : pick_shape(p.m[t], m[cslot], c, min_shape);
    AlignItem it;
    it.text = sw ? cslot : t;
    it.query = sw ? t : cslot;
    it.job = 0;
    it.slot = sw ? kSwapFlag : 0;
    int list = sh;
    if (split && (sw ? p.m[t] : m[cslot]) >= 64) {
      it.slot |= sh << kSplitShapeShift;
      list = kSplitList;
    }
    bins[list].push_back(it);
    bin_pair[list].push_back(i);
  }
  std::vector<void*> A;
  int64_t F = 0;
  uint2* d_pk = nullptr;
  int64_t* d_pkoff = nullptr;
  int32_t *d_m = nullptr, *d_bc = nullptr, *res = nullptr, *ctr = nullptr;
  AlignItem* items = nullptr;
  JobDev* d_job = nullptr;
  unsigned long long* d_stats = nullptr;
  auto release = [&]() { for (void* x : A) (void)hipFree(x); };
  int st;
  if ((st = dev_alloc(A, F, &d_pk, p.pk.size() / 2)) || (st = dev_alloc(A, F, &d_pkoff, (size_t)nslots + 1)) ||
      (st = dev_alloc(A, F, &d_m, (size_t)nslots)) || (st = dev_alloc(A, F, &d_bc, (size_t)nslots)) ||
      (st = dev_alloc(A, F, &items, (size_t)npairs)) || (st = dev_alloc(A, F, &res, (size_t)npairs)) ||
      (st = dev_alloc(A, F, &ctr, (size_t)kNumCounters)) || (st = dev_alloc(A, F, &d_job, (size_t)1)) ||
      (st = dev_alloc(A, F, &d_stats, (size_t)kStatWords))) {
    release();
    return st;
  }
  JobDev jd;
  std::memset(&jd, 0, sizeof(jd));
  jd.st.pk = d_pk;
  jd.st.pkoff = d_pkoff;
  jd.st.m = d_m;
  jd.st.bc = d_bc;
  std::vector<AlignItem> all;
  std::vector<int32_t> counts(kNumCounters, 0);
  std::vector<int> first(kNumShapes + 1, 0);
  for (int sh = 0; sh <= kNumShapes; ++sh) {
    first[sh] = (int)all.size();
    counts[sh] = (int)bins[sh].size();
    all.insert(all.end(), bins[sh].begin(), bins[sh].end());
  }
  hipStream_t q = c->stream;
  hipError_t e;
  auto bail = [&](hipError_t err, const char* what) { release(); return hip_fail(err, what); };
  if ((e = hipMemcpyAsync(d_pk, p.pk.data(), p.pk.size() * 4, hipMemcpyHostToDevice, q)) ||
      (e = hipMemcpyAsync(d_pkoff, pkoff.data(), pkoff.size() * 8, hipMemcpyHostToDevice, q)) ||
      (e = hipMemcpyAsync(d_m, m.data(), m.size() * 4, hipMemcpyHostToDevice, q)) ||
      (e = hipMemcpyAsync(d_bc, bcv.data(), bcv.size() * 4, hipMemcpyHostToDevice, q)) ||
      (e = hipMemcpyAsync(items, all.data(), all.size() * sizeof(AlignItem), hipMemcpyHostToDevice, q)) ||
      (e = hipMemcpyAsync(ctr, counts.data(), counts.size() * 4, hipMemcpyHostToDevice, q)) ||
      (e = hipMemcpyAsync(d_job, &jd, sizeof(jd), hipMemcpyHostToDevice, q)) ||
      (e = hipMemsetAsync(d_stats, 0, kStatWords * 8, q)))
    return bail(e, "align_pairs upload");
  for (int sh = 0; sh < kNumShapes; ++sh) {
    if (counts[sh] == 0) continue;
    AlignArgs aa;
    aa.jobs = d_job;
    aa.items = items + first[sh];
    aa.count = ctr + sh;
    aa.work = ctr + kWorkCounter0 + sh;
    aa.result = res + first[sh];
    aa.stats = d_stats;
    aa.skip_absorbed = 0;
    const int segs = 64 / kShapes[sh].s;
    const int waves = (counts[sh] + segs - 1) / segs;
    const int grid = std::max(1, std::min((waves + 3) / 4, c->cu_count * align_blocks_per_cu(sh)));
    if ((e = launch_align(sh, aa, grid, q)) != hipSuccess) return bail(e, "align launch");
  }
  if (counts[kSplitList] > 0) {
    AlignArgs aa;
    aa.jobs = d_job;
    aa.items = items + first[kSplitList];
    aa.count = ctr + kSplitList;
    aa.work = ctr + kWorkCounter0 + kSplitList;
    aa.result = res + first[kSplitList];
    aa.stats = d_stats;
    aa.skip_absorbed = 0;
    const int grid = std::max(1, std::min(counts[kSplitList], c->cu_count * 2));
    if ((e = launch_align_split(aa, grid, q)) != hipSuccess) return bail(e, "split launch");
  }
  std::vector<int32_t> got(all.size());
  unsigned long long hall[kStatWords], hst[ST_NUM];
  if ((e = hipMemcpyAsync(got.data(), res, got.size() * 4, hipMemcpyDeviceToHost, q)) ||
      (e = hipMemcpyAsync(hall, d_stats, sizeof(hall), hipMemcpyDeviceToHost, q)) || (e = hipStreamSynchronize(q)))
    return bail(e, "align_pairs download");
  stats_fold(hall, hst);
  for (int sh = 0; sh <= kNumShapes; ++sh)
    for (size_t j = 0; j < bin_pair[sh].size(); ++j) accept_out[bin_pair[sh][j]] = got[first[sh] + j];
  if (cells_exec_out) *cells_exec_out = (double)hst[ST_CELLS_EXEC];
  release();
  return GSDR_OK;
}

}  // namespace gsdr
