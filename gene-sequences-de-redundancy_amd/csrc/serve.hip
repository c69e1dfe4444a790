// Serving: continuous batching of clustering jobs on one device.
//
// Jobs are admitted while others run, as long as their HBM footprint fits the
// budget and a slot is free. Every admitted job is an entry of the device job
// table, so its filter units and alignment pairs share the launches of all
// other running jobs (the kernels are the batched form's). A job's result is
// its own greedy recursion, independent of what it shares launches with.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "driver.hpp"

using namespace gsdr;

namespace {

using Clock = std::chrono::steady_clock;

struct ServeJob {
  int64_t id = 0;
  std::unique_ptr<gsdr_store> store;
  gsdr_params params;
  int status = 0;  // 0 waiting for admission to the table, 1 running, 2 done, 3 failed
  std::vector<int32_t> result;
  int64_t reserved = 0;  // footprint charged against the budget
  int32_t reps = 0, steps = 0;
  Clock::time_point t_submit, t_done;
  std::string error;
};

}  // namespace

struct gsdr_server {
  gsdr_ctx* c = nullptr;
  int64_t budget = 0;
  int32_t max_slots = 0;
  int64_t used = 0;
  int64_t seqs = 0;  // sequences of the waiting and running jobs
  int64_t next_id = 1;
  std::mutex mu;
  // held by submissions while they build a store on the device and by the
  // step while it captures its graph: loads never run inside a capture
  std::mutex dev_mu;
  std::map<int64_t, std::unique_ptr<ServeJob>> jobs;
  std::vector<ServeJob*> waiting;  // loaded, not yet in the device table
  std::vector<ServeJob*> running;  // in the device table
  int steps_per_call = 4;
  // the steps of one call, captured for the current job table and replayed
  // while the table stays the same
  hipGraphExec_t exec = nullptr;
  std::vector<ServeJob*> graph_jobs;
  int64_t graph_ws = 0;  // workspace generation the graph's pointers belong to
};

static void drop_graph(gsdr_server* srv) {
  if (srv->exec) (void)hipGraphExecDestroy(srv->exec);
  srv->exec = nullptr;
  srv->graph_jobs.clear();
}

// Captures steps_per_call steps over the current table into srv->exec.
static hipError_t capture_steps(gsdr_server* srv, const StepLaunch& L, hipStream_t q) {
  gsdr_ctx* c = srv->c;
  drop_graph(srv);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamBeginCapture(q, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  for (int s = 0; s < srv->steps_per_call && e == hipSuccess; ++s) e = launch_step(c, L, q, nullptr);
  hipError_t e2 = hipStreamEndCapture(q, &graph);
  if (e == hipSuccess) e = e2;
  if (e == hipSuccess) e = hipGraphInstantiate(&srv->exec, graph, nullptr, nullptr, 0);
  if (graph) (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    srv->exec = nullptr;
    return e;
  }
  srv->graph_jobs = srv->running;
  srv->graph_ws = c->ws.generation;
  return hipSuccess;
}

// Upper bound of a job's HBM use: packed store, k-mer profile (at most one
// entry per base), per-sequence arrays and round state, the signature screen's
// arrays for long words.
static int64_t footprint_bound(const Records& r, int k) {
  const int64_t n = r.count(), b = (int64_t)r.data.size();
  const int64_t screen = k >= kScreenMinK ? n * (4 * kSigWords + 8) + n / 8 + 1024 : 0;
  return b / 4 + 4 * b + n * 120 + screen + (1 << 16);
}

extern "C" {

int gsdr_serve_open(gsdr_ctx* c, int64_t hbm_budget, int32_t max_slots, gsdr_server** srv) {
  if (!c || !srv || hbm_budget <= 0 || max_slots < 1) return fail(GSDR_E_INVALID, "bad serving parameters");
  if (max_slots > kMaxJobsPerLaunch) return fail(GSDR_E_LIMIT, "at most 1024 slots");
  auto s = std::make_unique<gsdr_server>();
  s->c = c;
  s->budget = hbm_budget;
  s->max_slots = max_slots;
  // alignment lists shared by all running jobs (32 M items per tile shape,
  // 6 GiB); the clamp after each step keeps k_spec * nc of all jobs within
  // them
  int st = workspace_reserve(c, 32ll << 20, max_slots);
  if (st) return st;
  hipError_t e;
  if ((e = hipMemsetAsync(c->ws.counters, 0, kNumCounters * sizeof(int32_t), c->stream)) != hipSuccess ||
      (e = hipMemsetAsync(c->ws.stats, 0, kStatWords * sizeof(unsigned long long), c->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(c->stream)) != hipSuccess)
    return hip_fail(e, "serve init");
  *srv = s.release();
  return GSDR_OK;
}

int gsdr_serve_submit(gsdr_server* srv, const gsdr_records* sorted, const gsdr_params* params, int64_t* job_id) {
  if (!srv || !sorted || !params || !job_id) return fail(GSDR_E_INVALID, "null argument");
  const Records& r = sorted->r;
  const int64_t n = r.count();
  if (n <= 0) return fail(GSDR_E_INPUT, "no records");
  for (int64_t i = 0; i + 1 < n; ++i)
    if (r.len(i) < r.len(i + 1)) return fail(GSDR_E_INVALID, "records are not in length order");
  const float t = params->threshold;
  if (!(t >= 0.8f) || t >= 1.0f) return fail(GSDR_E_INVALID, "similarity out of range: 0.8<=similarity<1");
  const int k = params->word_length ? params->word_length : auto_word_length(t);
  if (k < 4 || k > 8) return fail(GSDR_E_INVALID, "word length out of range: 4<=word length<=8");
  const int64_t need = footprint_bound(r, k);
  auto job = std::make_unique<ServeJob>();
  job->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    const int64_t live = (int64_t)(srv->waiting.size() + srv->running.size());
    const int64_t cap = srv->c->ws.item_cap;  // a step's first phase queues up to nc pairs per job
    if (need > srv->budget || n > cap) return fail(GSDR_E_LIMIT, "job exceeds the whole budget");
    if (live >= srv->max_slots || srv->used + need > srv->budget || srv->seqs + n > cap)
      return fail(GSDR_E_BUSY, "no free slot or HBM budget now; retry after jobs finish");
    srv->used += need;
    srv->seqs += n;
    job->reserved = need;
    job->id = srv->next_id++;
  }
  // load outside the bookkeeping lock, on the upload stream, never during a
  // graph capture of the step
  job->store = std::make_unique<gsdr_store>();
  job->store->ctx = srv->c;
  int st;
  {
    std::lock_guard<std::mutex> dk(srv->dev_mu);
    st = store_build(srv->c, r, k, *job->store);
    int kk = 0;
    if (!st) st = check_params(*job->store, *params, &kk);
    if (!st) st = store_set_threshold(*job->store, t);
    if (st) store_release(*job->store);
  }
  if (st) {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->used -= job->reserved;
    srv->seqs -= n;
    return st;
  }
  job->params = *params;
  job->params.word_length = k;
  std::lock_guard<std::mutex> lk(srv->mu);
  // charge the real footprint
  srv->used += job->store->footprint - job->reserved;
  job->reserved = job->store->footprint;
  *job_id = job->id;
  srv->waiting.push_back(job.get());
  srv->jobs[job->id] = std::move(job);
  return GSDR_OK;
}

int gsdr_serve_step(gsdr_server* srv, int32_t* running_out) {
  if (!srv) return fail(GSDR_E_INVALID, "null server");
  gsdr_ctx* c = srv->c;
  hipStream_t q = c->stream;
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  std::vector<ServeJob*> admitted;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    admitted.swap(srv->waiting);
  }
  for (ServeJob* j : admitted) {
    j->status = 1;
    srv->running.push_back(j);
  }
  const int nj = (int)srv->running.size();
  if (nj > 0) {
    Workspace& w = c->ws;
    std::vector<JobDev> table((size_t)nj);
    int smax = 0;
    unsigned kmask = 0;
    int64_t low_bases = 0;
    float tmin = 1.f;
    for (int i = 0; i < nj; ++i) {
      table[i] = srv->running[i]->store->job(srv->running[i]->params);
      low_bases += low_t_bases(srv->running[i]->store->bases, srv->running[i]->params.threshold);
      tmin = std::min(tmin, srv->running[i]->params.threshold);
      smax = std::max(smax, max_shape_for(srv->running[i]->store->max_m));
      kmask |= 1u << srv->running[i]->store->k;
    }
    if ((e = hipMemcpyAsync(w.jobs, table.data(), sizeof(JobDev) * nj, hipMemcpyHostToDevice, q)) != hipSuccess)
      return hip_fail(e, "job table");
    // start the newly admitted jobs: their entries sit at the end of the table
    for (int i = nj - (int)admitted.size(); i < nj; ++i) {
      const gsdr_params& p = srv->running[i]->params;
      const int k_max = p.spec_reps > 0 ? std::min(p.spec_reps, kMaxSpec) : kMaxSpec;
      if ((e = launch_init_jobs(w.jobs + i, &table[i], 1, k_max, nullptr, q)) != hipSuccess)
        return hip_fail(e, "init");
    }
    if ((e = launch_clamp_spec(w.jobs, nj, w.item_cap, q)) != hipSuccess) return hip_fail(e, "clamp");
    // recapture when the table changed, or when another call on the context
    // reallocated the workspace the graph's kernels point into
    if (!srv->exec || srv->graph_jobs != srv->running || srv->graph_ws != c->ws.generation) {
      const StepLaunch L = make_step_launch(c, c->ws, nj, smax, kmask, nullptr, grid_mode(low_bases, tmin, smax));
      std::lock_guard<std::mutex> dk(srv->dev_mu);
      if ((e = capture_steps(srv, L, q)) != hipSuccess) return hip_fail(e, "serve capture");
    }
    if ((e = hipGraphLaunch(srv->exec, q)) != hipSuccess) return hip_fail(e, "serve steps");
    std::vector<JobState> hs((size_t)nj);
    for (int i = 0; i < nj; ++i)
      if ((e = hipMemcpyAsync(&hs[i], table[i].state, sizeof(JobState), hipMemcpyDeviceToHost, q)) != hipSuccess)
        return hip_fail(e, "state copy");
    unsigned long long overflow = 0;
    if ((e = hipMemcpyAsync(&overflow, w.stats + ST_OVERFLOW, sizeof(overflow), hipMemcpyDeviceToHost, q)) !=
        hipSuccess)
      return hip_fail(e, "stats copy");
    if ((e = hipStreamSynchronize(q)) != hipSuccess) return hip_fail(e, "serve sync");
    if (overflow) {
      // A step queued more pairs than a shape list holds (the clamp after
      // every step keeps k_spec * nc within the lists, so this is not
      // expected): the running jobs' round state is partly written. They
      // fail (poll reports GSDR_E_LIMIT) and release their memory, the
      // counters are cleared, and the server goes on with later jobs.
      if ((e = hipMemsetAsync(w.stats + ST_OVERFLOW, 0, sizeof(unsigned long long), q)) != hipSuccess ||
          (e = hipMemsetAsync(w.counters, 0, kNumCounters * sizeof(int32_t), q)) != hipSuccess ||
          (e = hipStreamSynchronize(q)) != hipSuccess)
        return hip_fail(e, "overflow reset");
      drop_graph(srv);
      for (ServeJob* j : srv->running) {
        j->t_done = Clock::now();
        j->error = "alignment list overflow";
        store_release(*j->store);
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->used -= j->reserved;
        srv->seqs -= j->store->n;
        j->reserved = 0;
        j->status = 3;
      }
      srv->running.clear();
      if (running_out) {
        std::lock_guard<std::mutex> lk(srv->mu);
        *running_out = (int32_t)srv->waiting.size();
      }
      return fail(GSDR_E_LIMIT, "alignment list overflow: the running jobs failed, the server goes on");
    }
    std::vector<ServeJob*> still;
    for (int i = 0; i < nj; ++i) {
      ServeJob* j = srv->running[i];
      if (!hs[i].done) {
        still.push_back(j);
        continue;
      }
      j->result.resize((size_t)j->store->n);
      if ((e = hipMemcpy(j->result.data(), j->store->cluster, (size_t)j->store->n * sizeof(int32_t),
                         hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_fail(e, "result copy");
      j->reps = hs[i].reps;
      j->steps = hs[i].steps;
      j->t_done = Clock::now();
      store_release(*j->store);
      std::lock_guard<std::mutex> lk(srv->mu);
      srv->used -= j->reserved;
      srv->seqs -= j->store->n;
      j->reserved = 0;
      j->status = 2;
    }
    srv->running.swap(still);
  }
  if (running_out) {
    std::lock_guard<std::mutex> lk(srv->mu);
    *running_out = (int32_t)(srv->running.size() + srv->waiting.size());
  }
  return GSDR_OK;
}

int gsdr_serve_poll(gsdr_server* srv, int64_t job_id, int32_t* cluster_out, gsdr_stats* stats, int32_t* done) {
  if (!srv || !done) return fail(GSDR_E_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(srv->mu);
  auto it = srv->jobs.find(job_id);
  if (it == srv->jobs.end()) return fail(GSDR_E_NOTFOUND, "unknown job id");
  ServeJob* j = it->second.get();
  if (j->status == 3) {
    // failed in a step (its memory is already released)
    *done = 1;
    const std::string why = j->error;
    srv->jobs.erase(it);
    return fail(GSDR_E_LIMIT, why.c_str());
  }
  if (j->status != 2) {
    *done = 0;
    return GSDR_OK;
  }
  *done = 1;
  if (cluster_out) std::copy(j->result.begin(), j->result.end(), cluster_out);
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->n_seqs = j->store->n;
    stats->n_bases = j->store->bases;
    stats->reps = j->reps;
    stats->rounds = j->steps;
    stats->ms_total = std::chrono::duration<double, std::milli>(j->t_done - j->t_submit).count();
  }
  srv->jobs.erase(it);
  return GSDR_OK;
}

void gsdr_serve_close(gsdr_server* srv) {
  if (!srv) return;
  drop_graph(srv);
  for (auto& kv : srv->jobs)
    if (kv.second->store) store_release(*kv.second->store);
  delete srv;
}

}  // extern "C"
