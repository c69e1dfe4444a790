// Batched form: speculative multi-representative steps for one or many jobs,
// captured as a HIP graph so the host synchronises once per graph replay
// instead of once per representative.
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

// One step for every job. Phase A: the speculative representatives against
// each other, then resolve_reps fixes the true ones. Phase B: filter,
// alignment (skipping candidates an earlier true representative already
// took), resolve. The tile-shape kernels of a phase run concurrently or one
// after another, with the split kernel first when long pairs are split.
// ev (optional) brackets filter and alignment time of both phases.
hipError_t launch_step(gsdr_ctx* c, const StepLaunch& L, hipStream_t q, hipEvent_t* ev) {
  hipError_t e;
  if (ev && (e = hipEventRecord(ev[0], q)) != hipSuccess) return e;
  if ((e = launch_filter(L.fa_reps, L.kmask, L.reps_filter_grid, q)) != hipSuccess) return e;
  if (L.pigeon && (e = launch_pigeon(L.pa, L.pigeon_grid, q)) != hipSuccess) return e;
  if (ev && (e = hipEventRecord(ev[1], q)) != hipSuccess) return e;
  if ((e = launch_align_shapes(c, L.aa_reps, L.smax, L.reps_grid, q, L.fork, L.split ? &L.aa_split : nullptr,
                               L.split_grid)) != hipSuccess)
    return e;
  if (ev && (e = hipEventRecord(ev[2], q)) != hipSuccess) return e;
  if ((e = launch_resolve_reps(L.ra, q)) != hipSuccess) return e;
  if (ev && (e = hipEventRecord(ev[3], q)) != hipSuccess) return e;
  if (L.screen && (e = launch_screen(L.sa, L.screen_grid, q)) != hipSuccess) return e;
  if ((e = launch_filter(L.fa_main, L.kmask, L.filter_grid, q)) != hipSuccess) return e;
  if (L.pigeon && (e = launch_pigeon(L.pa, L.pigeon_grid, q)) != hipSuccess) return e;
  if (ev && (e = hipEventRecord(ev[4], q)) != hipSuccess) return e;
  if ((e = launch_align_shapes(c, L.aa, L.smax, L.align_grid, q, L.fork, L.split ? &L.aa_split : nullptr,
                               L.split_grid)) != hipSuccess)
    return e;
  if (ev && (e = hipEventRecord(ev[5], q)) != hipSuccess) return e;
  if ((e = launch_resolve(L.ra, q)) != hipSuccess) return e;
  // with several jobs, bound the next step's widths by the lists' capacity
  return L.ra.njobs > 1 ? launch_clamp_spec(L.ra.jobs, L.ra.njobs, L.ra.pair_budget, q) : hipSuccess;
}

StepLaunch make_step_launch(gsdr_ctx* c, const Workspace& w, int njobs, int smax, unsigned kmask,
                            int32_t* jobs_left, GridMode mode, const Tuning& T) {
  StepLaunch L;
  L.kmask = kmask;
  L.smax = smax;
  L.fork = fork_for(T, mode);
  L.filter_grid = c->cu_count * filter_blocks_for(T, mode);
  // phase A holds at most 64 filter units per job (one representative table
  // per block for long words): one block per unit, so the units' table
  // builds run side by side (16 blocks in all took 35 us per step of a
  // t=0.99 job, four units one after another in each block)
  L.reps_filter_grid = std::max(16, std::min(njobs * 64, c->cu_count * 4));
  for (int sh = 0; sh < kNumShapes; ++sh) {
    L.align_grid[sh] = c->cu_count * align_blocks_for(T, sh, mode);
    // phase A holds at most k_spec^2/2 pairs per job
    const int segs = 64 / kShapes[sh].s;
    L.reps_grid[sh] = std::max(1, std::min(c->cu_count * 4, (njobs * kMaxSpec * kMaxSpec / 2 / segs + 3) / 4));
  }
  L.fa_main.jobs = w.jobs;
  L.fa_main.njobs = njobs;
  L.fa_main.min_shape = 0;
  L.fa_main.swap_ratio = T.swap_ratio;
  L.fa_main.shapes = T.shapes;
  L.fa_main.phase = PHASE_MAIN;
  L.fa_main.items = w.items;
  L.fa_main.item_cap = w.item_cap;
  L.fa_main.bin_count = w.counters;
  L.fa_main.stats = w.stats;
  // the screen only runs for jobs whose stores carry signatures (k >= 6)
  L.screen = T.screen && (kmask >> kScreenMinK) != 0;
  L.fa_main.screen = L.screen;
  L.screen_grid = c->cu_count * 8;
  L.sa = ScreenArgs{w.jobs, njobs, w.stats};
  // the split kernel's resident blocks: 4 per CU (two waves each, at most
  // two waves per SIMD by its registers); the split list takes one pair per
  // block, so every split pair starts at once. Both phases use the list:
  // the resolve kernels clear its counters between them.
  const int split_rows = split_rows_for(T, mode);
  L.split = split_rows > 0;
  L.split_grid = c->cu_count * 4;
  L.fa_main.split_rows = split_rows;
  L.fa_main.split_cap = std::min<int64_t>(w.split_cap, L.split_grid);
  L.fa_reps = L.fa_main;
  L.fa_reps.phase = PHASE_REPS;
  L.fa_reps.screen = 0;
  L.aa_split.jobs = w.jobs;
  L.aa_split.items = w.items + (int64_t)kSplitList * w.item_cap;
  L.aa_split.count = w.counters + kSplitList;
  L.aa_split.work = w.counters + kWorkCounter0 + kSplitList;
  L.aa_split.result = nullptr;
  L.aa_split.stats = w.stats;
  L.aa_split.skip_absorbed = 0;
  L.aa_split.win_period = T.win_period;
  for (int sh = 0; sh < kNumShapes; ++sh) {
    AlignArgs& aa = L.aa[sh];
    aa.jobs = w.jobs;
    aa.items = w.items + (int64_t)sh * w.item_cap;
    aa.count = w.counters + sh;
    aa.work = w.counters + kWorkCounter0 + sh;
    aa.result = nullptr;
    aa.stats = w.stats;
    aa.skip_absorbed = 1;
    aa.win_period = T.win_period;
    L.aa_reps[sh] = aa;
    L.aa_reps[sh].skip_absorbed = 0;
  }
  L.ra.jobs = w.jobs;
  L.ra.njobs = njobs;
  L.ra.counters = w.counters;
  L.ra.ncounters = kNumCounters;
  L.ra.jobs_left = jobs_left;
  L.ra.stats = w.stats;
  L.ra.pair_budget = w.item_cap;
  L.ra.grid = c->cu_count * 2;
  L.ra.reps_marked = 1;
  L.pigeon = T.pigeon;
  L.pigeon_grid = c->cu_count * 4;
  L.pa = PigeonArgs{w.jobs, w.items, w.item_cap, w.counters, smax, w.stats};
  return L;
}

int cluster_batched(gsdr_ctx* c, int njobs, gsdr_store* const* stores, const gsdr_params* params,
                    int32_t* const* out, gsdr_stats* stats, const Workspace* fixed) {
  if (njobs == 0) {
    if (stats) std::memset(stats, 0, sizeof(*stats));
    return GSDR_OK;
  }
  if (njobs > kMaxJobsPerLaunch) return fail(GSDR_E_LIMIT, "more than 1024 jobs in one call");
  HIP_TRY(hipSetDevice(c->device), "hipSetDevice");
  int st;
  int smax = 0, timing = 0, spec = 0, tune = 0;
  unsigned kmask = 0;
  int64_t total_n = 0, total_bases = 0, max_n = 0;
  int64_t low_bases = 0;
  float tmin = 1.f;
  // host bookkeeping: on the stack for one job (the no-allocation entry)
  JobDev job1[1];
  std::vector<JobDev> jobs_many((size_t)(njobs > 1 ? njobs : 0));
  JobDev* jobs = njobs > 1 ? jobs_many.data() : job1;
  for (int j = 0; j < njobs; ++j) {
    gsdr_store& s = *stores[j];
    if (s.ctx != c) return fail(GSDR_E_INVALID, "store belongs to another context");
    int k = 0;
    if ((st = check_params(s, params[j], &k))) return st;
    if ((st = store_set_threshold(s, params[j].threshold))) return st;
    jobs[j] = s.job(params[j]);
    kmask |= 1u << k;
    smax = std::max(smax, max_shape_for(s.max_m));
    total_n += s.n;
    total_bases += s.bases;
    max_n = std::max<int64_t>(max_n, s.n);
    timing |= params[j].timing;
    spec = std::max(spec, params[j].spec_reps);
    low_bases += low_t_bases(s.bases, params[j].threshold);
    tmin = std::min(tmin, params[j].threshold);
    tune |= params[j].tune;
  }
  const Tuning T = tuning_for(c, tune != 0);
  for (int j = 0; j + 1 < njobs; ++j)
    for (int i = j + 1; i < njobs; ++i)
      if (stores[j] == stores[i]) return fail(GSDR_E_INVALID, "the same store appears twice");
  const int k_max = std::min(spec > 0 ? std::min(spec, kMaxSpec) : kMaxSpec, T.spec_max);
  // alignment list capacity per shape: enough for several speculative
  // representatives over every job's candidates
  // (the resolve step bounds each job's k_spec so that k_spec * nc of all
  // jobs fits, so the capacity also caps the speculation width)
  int64_t item_cap =
      std::min<int64_t>(std::max<int64_t>((int64_t)T.list_factor * total_n, max_n), 64ll << 20);
  if (fixed) {
    if (timing) return fail(GSDR_E_INVALID, "timing needs HIP events, which this entry does not create");
    if (njobs > fixed->jobs_cap) return fail(GSDR_E_LIMIT, "too many jobs for the scratch");
    item_cap = fixed->item_cap;
  } else if ((st = workspace_reserve(c, item_cap, njobs))) {
    return st;
  }
  const Workspace w = fixed ? *fixed : c->ws;
  // the first step queues every job's whole candidate list once
  if (total_n > w.item_cap)
    return fail(GSDR_E_LIMIT, fixed ? "too many sequences for the scratch" : "too many sequences in one call");
  hipStream_t q = c->stream;
  const auto t0 = std::chrono::steady_clock::now();
  int32_t left = 0;
  for (int j = 0; j < njobs; ++j) left += jobs[j].n > 0;
  HIP_TRY(hipMemcpyAsync(w.jobs, jobs, sizeof(JobDev) * njobs, hipMemcpyHostToDevice, q), "job table");
  HIP_TRY(hipMemsetAsync(w.stats, 0, kStatWords * sizeof(unsigned long long), q), "memset");
  HIP_TRY(hipMemsetAsync(w.counters, 0, kNumCounters * sizeof(int32_t), q), "memset");
  HIP_TRY(hipMemcpyAsync(w.jobs_left, &left, sizeof(int32_t), hipMemcpyHostToDevice, q), "jobs_left");
  HIP_TRY(launch_init_jobs(w.jobs, jobs, njobs, k_max, w.jobs_left, q), "init");

  StepLaunch L = make_step_launch(c, w, njobs, smax, kmask, w.jobs_left, grid_mode(low_bases, tmin, smax), T);
  // the speculation bound follows this call's capacity, not a larger
  // workspace left by an earlier call
  L.ra.pair_budget = std::min<int64_t>(item_cap, w.item_cap);

  double ms_filter = 0, ms_align = 0;
  int64_t launches = 2, steps_launched = 0;
  // a job needs at most n steps; every step fixes at least one representative
  const int64_t step_limit = max_n + 2;
  // kernels per step: two filters, resolve_reps and the three resolve
  // kernels, the shape kernels of both phases, the clamp with several jobs,
  // the screen and the pigeonhole tests when on, the split kernel of each phase
  const int64_t per_step = 6 + 2 * (L.smax + 1) + (njobs > 1 ? 1 : 0) + L.screen + 2 * L.pigeon + 2 * L.split;
  if (timing) {
    // timed form: the same step launched directly with events around the
    // filter and the alignment kernels
    hipEvent_t ev[6];
    for (auto& e : ev) HIP_TRY(hipEventCreate(&e), "hipEventCreate");
    while (true) {
      HIP_TRY(launch_step(c, L, q, ev), "step launch");
      launches += per_step;
      ++steps_launched;
      HIP_TRY(hipMemcpyAsync(c->h_state, w.jobs_left, sizeof(int32_t), hipMemcpyDeviceToHost, q), "poll");
      HIP_TRY(hipStreamSynchronize(q), "step sync");
      float t01 = 0, t12 = 0, t34 = 0, t45 = 0;
      HIP_TRY(hipEventElapsedTime(&t01, ev[0], ev[1]), "event time");
      HIP_TRY(hipEventElapsedTime(&t12, ev[1], ev[2]), "event time");
      HIP_TRY(hipEventElapsedTime(&t34, ev[3], ev[4]), "event time");
      HIP_TRY(hipEventElapsedTime(&t45, ev[4], ev[5]), "event time");
      ms_filter += t01 + t34;
      ms_align += t12 + t45;
      if (*reinterpret_cast<int32_t*>(c->h_state) == 0) break;
      if (steps_launched > step_limit) return fail(GSDR_E_DEVICE, "steps did not converge");
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
  } else if (fixed) {
    // no graph (instantiating one allocates): the same steps launched
    // directly, one host synchronisation per B steps
    const int B = 16;
    int32_t* h_left = reinterpret_cast<int32_t*>(c->h_state);
    while (true) {
      for (int b = 0; b < B; ++b) HIP_TRY(launch_step(c, L, q, nullptr), "step launch");
      HIP_TRY(hipMemcpyAsync(h_left, w.jobs_left, sizeof(int32_t), hipMemcpyDeviceToHost, q), "poll");
      HIP_TRY(hipStreamSynchronize(q), "step sync");
      steps_launched += B;
      launches += (int64_t)B * per_step;
      if (*h_left == 0) break;
      if (steps_launched > step_limit + B) return fail(GSDR_E_DEVICE, "steps did not converge");
    }
  } else {
    // graph form: B steps per replay, one host synchronisation per replay
    const int B = 16;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    HIP_TRY(hipStreamBeginCapture(q, hipStreamCaptureModeThreadLocal), "begin capture");
    for (int b = 0; b < B; ++b) {
      hipError_t e = launch_step(c, L, q, nullptr);
      if (e != hipSuccess) {
        hipGraph_t g2;
        (void)hipStreamEndCapture(q, &g2);
        if (g2) (void)hipGraphDestroy(g2);
        return hip_fail(e, "capture step");
      }
    }
    HIP_TRY(hipStreamEndCapture(q, &graph), "end capture");
    HIP_TRY(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
    int32_t* h_left = reinterpret_cast<int32_t*>(c->h_state);
    int rc = GSDR_OK;
    while (true) {
      hipError_t e = hipGraphLaunch(exec, q);
      if (e == hipSuccess) e = hipMemcpyAsync(h_left, w.jobs_left, sizeof(int32_t), hipMemcpyDeviceToHost, q);
      if (e == hipSuccess) e = hipStreamSynchronize(q);
      if (e != hipSuccess) { rc = hip_fail(e, "graph replay"); break; }
      steps_launched += B;
      launches += (int64_t)B * per_step;
      if (*h_left == 0) break;
      if (steps_launched > step_limit + B) { rc = fail(GSDR_E_DEVICE, "steps did not converge"); break; }
    }
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    if (rc) return rc;
  }
  for (int j = 0; j < njobs; ++j)
    HIP_TRY(hipMemcpyAsync(out[j], jobs[j].cluster, (size_t)jobs[j].n * sizeof(int32_t), hipMemcpyDeviceToHost, q),
            "cluster copy");
  unsigned long long hall[kStatWords], hst[ST_NUM];
  HIP_TRY(hipMemcpyAsync(hall, w.stats, sizeof(hall), hipMemcpyDeviceToHost, q), "stats copy");
  JobState hs1[1];
  std::vector<JobState> hs_many((size_t)(njobs > 1 ? njobs : 0));
  JobState* hs = njobs > 1 ? hs_many.data() : hs1;
  for (int j = 0; j < njobs; ++j)
    HIP_TRY(hipMemcpyAsync(&hs[j], jobs[j].state, sizeof(JobState), hipMemcpyDeviceToHost, q), "state copy");
  HIP_TRY(hipStreamSynchronize(q), "final sync");
  stats_fold(hall, hst);
  if (hst[ST_OVERFLOW]) return fail(GSDR_E_LIMIT, "alignment list overflow");
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats_from_device(hst, stats);
    stats->n_seqs = total_n;
    stats->n_bases = total_bases;
    int64_t reps = 0, steps = 0;
    for (int j = 0; j < njobs; ++j) {
      for (int i = 0; i < jobs[j].n; ++i) reps += out[j][i] == i;
      steps = std::max<int64_t>(steps, hs[j].steps);
    }
    stats->reps = reps;
    stats->rounds = steps;
    stats->launches = launches;
    stats->ms_total = ms;
    stats->ms_filter = ms_filter;
    stats->ms_align = ms_align;
    stats->filter_launches = (2 + L.screen) * steps_launched;
    stats->align_launches = steps_launched * 2 * (L.smax + 1 + L.split);
  }
  return GSDR_OK;
}

int cluster_jobs(gsdr_ctx* c, int njobs, gsdr_store* const* stores, const gsdr_params* params,
                 int32_t* const* out, gsdr_stats* stats) {
  for (int j = 0; j < njobs; ++j)
    if (params[j].mode != GSDR_MODE_BATCHED && params[j].mode != GSDR_MODE_PLAIN)
      return fail(GSDR_E_INVALID, "unknown mode");
  if (njobs > 0 && params[0].mode == GSDR_MODE_PLAIN) {
    // plain form: the jobs one after another, one representative per round
    gsdr_stats total;
    std::memset(&total, 0, sizeof(total));
    for (int j = 0; j < njobs; ++j) {
      gsdr_stats s1;
      int st = cluster_plain(c, *stores[j], params[j], out[j], &s1);
      if (st) return st;
      total.n_seqs += s1.n_seqs; total.n_bases += s1.n_bases; total.reps += s1.reps; total.rounds += s1.rounds;
      total.launches += s1.launches; total.pairs += s1.pairs; total.rej_length += s1.rej_length;
      total.rej_comp += s1.rej_comp; total.rej_word += s1.rej_word; total.aligned += s1.aligned;
      total.rej_align += s1.rej_align; total.accepted += s1.accepted; total.cells_band += s1.cells_band;
      total.cells_exec += s1.cells_exec; total.cells_full += s1.cells_full; total.ms_total += s1.ms_total;
      total.ms_filter += s1.ms_filter; total.ms_align += s1.ms_align; total.filter_launches += s1.filter_launches;
      total.align_launches += s1.align_launches; total.filter_bytes += s1.filter_bytes;
      total.cells_decided += s1.cells_decided;
      total.rej_pigeon += s1.rej_pigeon;
      total.pigeon_checked += s1.pigeon_checked;
      total.screened += s1.screened;
    }
    if (stats) *stats = total;
    return GSDR_OK;
  }
  return cluster_batched(c, njobs, stores, params, out, stats);
}

}  // namespace gsdr
