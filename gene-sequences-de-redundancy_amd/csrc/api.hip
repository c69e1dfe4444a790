// C ABI of the engine (include/gsdr.h): records, device context, HBM store,
// plain per-round clustering and the pair-level alignment entry.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gsdr.h"
#include "driver.hpp"
#include "host.hpp"
#include "kernels_impl.hpp"

using namespace gsdr;

// ------------------------------------------------------------ error state
static thread_local std::string g_last_error;

namespace gsdr {
int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}
int hip_fail(hipError_t e, const char* what) {
  g_last_error = std::string(what) + ": " + hipGetErrorString(e);
  return GSDR_E_DEVICE;
}
int host_threads() {
  unsigned hc = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(hc, 16u));
}
}  // namespace gsdr

#define HIP_TRY(expr, what)                         \
  do {                                              \
    hipError_t _e = (expr);                         \
    if (_e != hipSuccess) return hip_fail(_e, what); \
  } while (0)

extern "C" {

const char* gsdr_strerror(int32_t status) {
  switch (status) {
    case GSDR_OK: return "ok";
    case GSDR_E_INVALID: return "invalid argument";
    case GSDR_E_IO: return "i/o error";
    case GSDR_E_INPUT: return "malformed or unsupported input";
    case GSDR_E_DEVICE: return "HIP device error";
    case GSDR_E_NOMEM: return "out of memory";
    case GSDR_E_LIMIT: return "size limit exceeded";
    case GSDR_E_BUSY: return "no free slot or memory";
    case GSDR_E_NOTFOUND: return "unknown job";
    default: return "unknown status";
  }
}
const char* gsdr_last_error(void) { return g_last_error.c_str(); }

// ----------------------------------------------------------------- options
// checkOption (func.cpp:6-57) without the exit: the messages the reference
// prints before exiting with status 0 are returned as the error text.
int gsdr_parse_options(int argc, char** argv, gsdr_options* out) {
  static const char kUsage[] = "use like this: cluster i inputFile t threshold";  // func.cpp:3
  if (!out || argc < 1 || !argv) return fail(GSDR_E_INVALID, "null argument");
  gsdr_options o;
  std::memset(&o, 0, sizeof(o));
  o.input = "testData.fasta";  // func.cpp:8-11
  o.output = "result.fasta";
  o.threshold = 0.95f;
  o.word_length = 0;
  o.mode = GSDR_MODE_BATCHED;
  for (int i = 1; i < argc; i += 2) {
    if (i + 1 >= argc || !argv[i] || !argv[i + 1]) return fail(GSDR_E_INVALID, kUsage);  // func.cpp:7
    const std::string key = argv[i];
    const char* val = argv[i + 1];
    if (key.rfind("--", 0) == 0) {
      const std::string v = val;
      if (key == "--mode" && (v == "plain" || v == "batched")) o.mode = v == "plain" ? GSDR_MODE_PLAIN : GSDR_MODE_BATCHED;
      else if (key == "--spec") o.spec_reps = std::atoi(val);
      else if (key == "--membership") o.membership = val;
      else if (key == "--report") o.report = val;
      else if (key == "--device") o.device = std::atoi(val);
      else if (key == "--drop" && (v == "0" || v == "1")) o.drop_filters = v == "1";
      else return fail(GSDR_E_INVALID, kUsage);
      continue;
    }
    // the reference switches on the first letter (func.cpp:13-29); a
    // leading '-' is accepted as well
    const char* k = key.c_str();
    if (k[0] == '-' && k[1] != 0) ++k;
    try {
      switch (k[0]) {
        case 'i': o.input = val; break;
        case 'o': o.output = val; break;
        case 't': o.threshold = std::stof(val); break;
        case 'w': o.word_length = std::stoi(val); break;
        default: return fail(GSDR_E_INVALID, kUsage);
      }
    } catch (...) {
      return fail(GSDR_E_INVALID, kUsage);
    }
  }
  if (o.threshold < 0.8 || o.threshold >= 1)  // func.cpp:31-35 (float against double literals)
    return fail(GSDR_E_INVALID, std::string("similarity out of range\n0.8<=similarity<1\n") + kUsage);
  if (o.word_length == 0) {
    o.word_length = auto_word_length(o.threshold);  // func.cpp:36-45
  } else if (o.word_length < 4 || o.word_length > 8) {  // func.cpp:46-51
    return fail(GSDR_E_INVALID, std::string("word length out of range\n4<=word length<=8\n") + kUsage);
  }
  *out = o;
  return GSDR_OK;
}

// ----------------------------------------------------------------- cutoffs
int gsdr_cutoffs(float threshold, int32_t word_length, const int32_t* raw_len, int64_t n, int32_t* bc_out,
                 int32_t* wc_out) {
  if (n < 0 || (n > 0 && (!raw_len || !bc_out || !wc_out))) return fail(GSDR_E_INVALID, "null argument");
  const int k = word_length == 0 ? auto_word_length(threshold) : word_length;
  if (k < 4 || k > 8) return fail(GSDR_E_INVALID, "word length must be 4..8");
  for (int64_t i = 0; i < n; ++i) {
    if (raw_len[i] < 0) return fail(GSDR_E_INVALID, "negative length");
    bc_out[i] = base_cutoff(raw_len[i], threshold);
    wc_out[i] = word_cutoff(raw_len[i], threshold, k);
  }
  return GSDR_OK;
}

// ----------------------------------------------------------------- records
static int read_records(const char* path, bool length_order, gsdr_records** out) {
  if (!path || !out) return fail(GSDR_E_INVALID, "null argument");
  auto r = std::make_unique<gsdr_records>();
  std::string err;
  int st = read_fasta(path, r->r, err, host_threads(), length_order);
  if (st != GSDR_OK) return fail(st, err);
  *out = r.release();
  return GSDR_OK;
}

int gsdr_records_read_fasta(const char* path, gsdr_records** out) { return read_records(path, false, out); }

int gsdr_records_read_fasta_sorted(const char* path, gsdr_records** out) { return read_records(path, true, out); }

int gsdr_records_from_buffers(const char* names, const int64_t* name_off, const char* data,
                              const int64_t* data_off, int64_t count, gsdr_records** out) {
  if (!out || count < 0 || !data_off || (count > 0 && !data)) return fail(GSDR_E_INVALID, "bad buffers");
  auto r = std::make_unique<gsdr_records>();
  for (int64_t i = 0; i < count; ++i) {
    const char* nm = "";
    int64_t nl = 0;
    if (names && name_off) { nm = names + name_off[i]; nl = name_off[i + 1] - name_off[i]; }
    const int64_t dl = data_off[i + 1] - data_off[i];
    if (dl < 0 || nl < 0) return fail(GSDR_E_INVALID, "negative record length");
    r->r.push(nm, nl, data + data_off[i], dl);
  }
  *out = r.release();
  return GSDR_OK;
}

int gsdr_records_generate(int32_t law, uint64_t seed, int64_t n, double median, gsdr_records** out) {
  if (!out || (law != GSDR_LAW_FAMILIES && law != GSDR_LAW_LOGNORMAL && law != GSDR_LAW_PARETO))
    return fail(GSDR_E_INVALID, "law must be GSDR_LAW_FAMILIES, GSDR_LAW_LOGNORMAL or GSDR_LAW_PARETO");
  auto r = std::make_unique<gsdr_records>();
  GenSpec spec;
  spec.law = law;
  spec.seed = seed;
  spec.n_override = n;
  spec.median = median;
  generate(spec, r->r);
  *out = r.release();
  return GSDR_OK;
}

int gsdr_records_write_fasta(const gsdr_records* r, const char* path, int32_t line_width) {
  if (!r || !path || line_width < 1) return fail(GSDR_E_INVALID, "bad argument");
  int st = write_fasta(r->r, path, line_width);
  if (st != GSDR_OK) return fail(st, std::string("cannot write ") + path);
  return GSDR_OK;
}

int gsdr_records_sort(gsdr_records* r, int64_t* perm_out) {
  if (!r) return fail(GSDR_E_INVALID, "null records");
  Records& rec = r->r;
  if (!rec.length_ordered) {
    std::vector<int64_t> perm = length_order(rec, host_threads());
    std::vector<int64_t> prior = std::move(rec.order);
    rec = permute(rec, perm, host_threads());
    if (!prior.empty())
      for (auto& p : perm) p = prior[(size_t)p];
    rec.length_ordered = true;
    rec.order = std::move(perm);
  }
  if (perm_out) std::copy(rec.order.begin(), rec.order.end(), perm_out);
  return GSDR_OK;
}

int64_t gsdr_records_count(const gsdr_records* r) { return r ? r->r.count() : -1; }
int64_t gsdr_records_bases(const gsdr_records* r) { return r ? (int64_t)r->r.data.size() : -1; }

int gsdr_records_get(const gsdr_records* r, int64_t i, const char** name, int64_t* name_len,
                     const char** data, int64_t* data_len) {
  if (!r || i < 0 || i >= r->r.count()) return fail(GSDR_E_INVALID, "record index out of range");
  if (name) *name = r->r.names.data() + r->r.name_off[i];
  if (name_len) *name_len = r->r.name_off[i + 1] - r->r.name_off[i];
  if (data) *data = r->r.data.data() + r->r.data_off[i];
  if (data_len) *data_len = r->r.data_off[i + 1] - r->r.data_off[i];
  return GSDR_OK;
}

int gsdr_records_write_result(const gsdr_records* sorted, const int32_t* cluster, const char* path,
                              int64_t* reps_out) {
  if (!sorted || !cluster || !path) return fail(GSDR_E_INVALID, "null argument");
  int64_t reps = write_result(sorted->r, cluster, path);
  if (reps < 0) return fail(GSDR_E_IO, std::string("cannot write ") + path);
  if (reps_out) *reps_out = reps;
  return GSDR_OK;
}

int gsdr_records_write_membership(const gsdr_records* sorted, const int64_t* perm,
                                  const int32_t* cluster, const char* path) {
  if (!sorted || !cluster || !path) return fail(GSDR_E_INVALID, "null argument");
  FILE* f = fopen(path, "wb");
  if (!f) return fail(GSDR_E_IO, std::string("cannot write ") + path);
  const Records& r = sorted->r;
  for (int64_t s = 0; s < r.count(); ++s) {
    const std::string nm(r.names.data() + r.name_off[s], (size_t)(r.name_off[s + 1] - r.name_off[s]));
    fprintf(f, "%lld\t%lld\t%d\t%s\n", (long long)s, (long long)(perm ? perm[s] : s), cluster[s], nm.c_str());
  }
  if (fclose(f) != 0) return fail(GSDR_E_IO, "write failed");
  return GSDR_OK;
}

void gsdr_records_free(gsdr_records* r) { delete r; }

// ----------------------------------------------------------------- context
int gsdr_open(int32_t device, gsdr_ctx** out) {
  if (!out) return fail(GSDR_E_INVALID, "null ctx");
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) return fail(GSDR_E_DEVICE, "no HIP device available");
  if (device < 0 || device >= count) return fail(GSDR_E_INVALID, "device index out of range");
  HIP_TRY(hipSetDevice(device), "hipSetDevice");
  auto c = std::make_unique<gsdr_ctx>();
  c->device = device;
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
  c->cu_count = prop.multiProcessorCount;
  c->hbm_bytes = (int64_t)prop.totalGlobalMem;
  // some driver stacks report an empty marketing name: fall back to the
  // architecture and compute-unit count
  if (prop.name[0] != 0)
    snprintf(c->name, sizeof(c->name), "%s (%s)", prop.name, prop.gcnArchName);
  else
    snprintf(c->name, sizeof(c->name), "%s, %d CUs, %.0f GiB", prop.gcnArchName, prop.multiProcessorCount,
             (double)prop.totalGlobalMem / (double)(1ll << 30));
  if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos)
    return fail(GSDR_E_DEVICE, std::string("kernels are built for gfx950, device is ") + prop.gcnArchName);
  // The first symbol lookup loads the library's code object (tens of ms);
  // it runs on a helper thread, with the device current there, beside the
  // stream and event creation below, and is joined before the context is
  // handed out (also on every early return).
  struct Loader {
    int st = GSDR_OK;
    std::string err;
    void* ptr = nullptr;
    std::thread t;
    ~Loader() { if (t.joinable()) t.join(); }
  } ld;
  ld.t = std::thread([&ld, device] {
    if (hipSetDevice(device) != hipSuccess) {
      ld.st = GSDR_E_DEVICE;
      ld.err = "hipSetDevice (code object load)";
      return;
    }
    // this device's copy of the no-allocation scratch array
    ld.st = scratch_setup(device, &ld.ptr);
    if (ld.st != GSDR_OK) ld.err = gsdr_last_error();
  });
  HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
  HIP_TRY(hipStreamCreateWithFlags(&c->upload, hipStreamNonBlocking), "hipStreamCreate");
  HIP_TRY(hipHostMalloc((void**)&c->h_state, sizeof(JobState) * 4, hipHostMallocDefault), "hipHostMalloc");
  c->h_state_cap = 4;
  for (int i = 0; i <= kNumShapes; ++i) {
    HIP_TRY(hipStreamCreateWithFlags(&c->side[i], hipStreamNonBlocking), "hipStreamCreate");
    HIP_TRY(hipEventCreateWithFlags(&c->join[i], hipEventDisableTiming), "hipEventCreate");
  }
  HIP_TRY(hipEventCreateWithFlags(&c->fork, hipEventDisableTiming), "hipEventCreate");
  ld.t.join();
  if (ld.st != GSDR_OK) return fail(ld.st, ld.err);
  c->scratch = ld.ptr;
  *out = c.release();
  return GSDR_OK;
}

void gsdr_close(gsdr_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  workspace_release(c);
  for (int i = 0; i <= kNumShapes; ++i) {
    if (c->side[i]) (void)hipStreamDestroy(c->side[i]);
    if (c->join[i]) (void)hipEventDestroy(c->join[i]);
  }
  if (c->fork) (void)hipEventDestroy(c->fork);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  if (c->upload) (void)hipStreamDestroy(c->upload);
  if (c->h_state) (void)hipHostFree(c->h_state);
  delete c;
}

int gsdr_device_info(gsdr_ctx* c, char* name, int32_t name_len, int32_t* cu_count, int64_t* hbm_bytes) {
  if (!c) return fail(GSDR_E_INVALID, "null ctx");
  if (name && name_len > 0) snprintf(name, (size_t)name_len, "%s", c->name);
  if (cu_count) *cu_count = c->cu_count;
  if (hbm_bytes) *hbm_bytes = c->hbm_bytes;
  return GSDR_OK;
}

// ------------------------------------------------------------------- store
int gsdr_load(gsdr_ctx* c, const gsdr_records* sorted, int32_t word_length, gsdr_store** out) {
  if (!c || !sorted || !out) return fail(GSDR_E_INVALID, "null argument");
  if (word_length < 4 || word_length > 8) return fail(GSDR_E_INVALID, "word length must be 4..8");
  const int64_t n = sorted->r.count();
  if (n <= 0) return fail(GSDR_E_INPUT, "no records");
  if (n >= (int64_t)INT32_MAX / 2) return fail(GSDR_E_LIMIT, "too many records");
  for (int64_t i = 0; i + 1 < n; ++i)
    if (sorted->r.len(i) < sorted->r.len(i + 1)) return fail(GSDR_E_INVALID, "records are not in length order");
  auto s = std::make_unique<gsdr_store>();
  s->ctx = c;
  int st = store_build(c, sorted->r, word_length, *s);
  if (st != GSDR_OK) { store_release(*s); return st; }
  *out = s.release();
  return GSDR_OK;
}

void gsdr_store_free(gsdr_store* s) {
  if (!s) return;
  store_release(*s);
  delete s;
}

int64_t gsdr_store_footprint(const gsdr_store* s) { return s ? s->footprint : -1; }

// -------------------------------------------------------------- clustering
int gsdr_cluster(gsdr_ctx* c, gsdr_store* s, const gsdr_params* p, int32_t* cluster_out, gsdr_stats* stats) {
  if (!c || !s || !p || !cluster_out) return fail(GSDR_E_INVALID, "null argument");
  return cluster_store(c, *s, *p, cluster_out, stats);
}

int gsdr_cluster_noalloc(gsdr_ctx* c, gsdr_store* s, const gsdr_params* p, int32_t* cluster_out,
                         gsdr_stats* stats) {
  if (!c || !s || !p || !cluster_out) return fail(GSDR_E_INVALID, "null argument");
  if (s->ctx != c) return fail(GSDR_E_INVALID, "store belongs to another context");
  return cluster_noalloc(c, *s, *p, cluster_out, stats);
}

int gsdr_cluster_jobs(gsdr_ctx* c, int32_t njobs, gsdr_store* const* stores, const gsdr_params* params,
                      int32_t* const* cluster_out, gsdr_stats* stats) {
  if (!c || njobs < 0 || (njobs > 0 && (!stores || !params || !cluster_out)))
    return fail(GSDR_E_INVALID, "null argument");
  return cluster_jobs(c, njobs, stores, params, cluster_out, stats);
}

int gsdr_autotune(gsdr_ctx* c, gsdr_store* s, const gsdr_params* p, int32_t max_trials, gsdr_tuning* out) {
  if (!c || !s || !p) return fail(GSDR_E_INVALID, "null argument");
  if (s->ctx != c) return fail(GSDR_E_INVALID, "store belongs to another context");
  return autotune(c, *s, *p, max_trials, out);
}

int gsdr_set_tuning(gsdr_ctx* c, const gsdr_tuning* t) {
  if (!c) return fail(GSDR_E_INVALID, "null context");
  if (!t) {
    c->tuned = Tuning();
    c->has_tuning = false;
    return GSDR_OK;
  }
  if (t->swap_ratio < 0 || t->spec_max < 1 || t->spec_max > kMaxSpec || (t->filter_blocks < 1 && t->filter_blocks != -1) ||
      t->filter_blocks > 64 || t->align_blocks < -1 || t->align_blocks > 64 ||
      t->fork_shapes < -1 || t->fork_shapes > 1 || t->list_factor < 1 || t->list_factor > 256 ||
      (t->pigeon != 0 && t->pigeon != 1) || (t->screen != 0 && t->screen != 1) || t->win_period < 0 ||
      t->win_period > 16 || (t->win_period & (t->win_period - 1)) != 0 || (t->split_rows != 0 && t->split_rows != -1 && (t->split_rows < 64 || t->split_rows > 32768)))
    return fail(GSDR_E_INVALID, "tuning out of range");
  for (int b = 0; b < kLenBins; ++b)
    if (t->shape_mask[b] < 0 || (uint32_t)t->shape_mask[b] > kAllShapes || t->shape_bias[b] < 0 ||
        t->shape_bias[b] > 3)
      return fail(GSDR_E_INVALID, "tuning out of range");
  Tuning T;
  T.swap_ratio = t->swap_ratio;
  T.spec_max = t->spec_max;
  T.filter_blocks = t->filter_blocks;
  T.align_blocks = t->align_blocks;
  T.fork_shapes = t->fork_shapes;
  T.list_factor = t->list_factor;
  T.pigeon = t->pigeon;
  T.screen = t->screen;
  for (int b = 0; b < kLenBins; ++b) {
    T.shapes.mask[b] = (uint16_t)t->shape_mask[b];
    T.shapes.bias[b] = (uint8_t)t->shape_bias[b];
  }
  T.win_period = t->win_period;
  T.split_rows = t->split_rows;
  c->tuned = T;
  c->has_tuning = true;
  return GSDR_OK;
}

int gsdr_get_tuning(const gsdr_ctx* c, gsdr_tuning* out) {
  if (!c || !out) return fail(GSDR_E_INVALID, "null argument");
  const Tuning T = tuning_for(c, true);
  std::memset(out, 0, sizeof(*out));
  out->swap_ratio = T.swap_ratio;
  out->spec_max = T.spec_max;
  out->filter_blocks = T.filter_blocks;
  out->align_blocks = T.align_blocks;
  out->fork_shapes = T.fork_shapes;
  out->list_factor = T.list_factor;
  out->pigeon = T.pigeon;
  out->screen = T.screen;
  tuning_shapes_out(T, out);
  return GSDR_OK;
}

int gsdr_align_pairs(gsdr_ctx* c, const char* data, const int64_t* data_off, int32_t nseq,
                     const int32_t* pairs, const int32_t* bc, int32_t npairs, int32_t shape,
                     int32_t* accept_out, double* cells_exec_out) {
  if (!c || !data_off || nseq <= 0 || npairs < 0 || (npairs > 0 && (!pairs || !bc || !accept_out)))
    return fail(GSDR_E_INVALID, "bad argument");
  return align_pairs(c, data, data_off, nseq, pairs, bc, npairs, shape, accept_out, cells_exec_out);
}

}  // extern "C"
