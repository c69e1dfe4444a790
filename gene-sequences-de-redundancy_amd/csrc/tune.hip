// Autotuner for the launch knobs of the batched form: role-swap ratio, cap
// on speculative representatives per step, filter and alignment workgroups
// per CU, whether the tile-shape kernels run concurrently, the alignment
// list capacity (which also bounds the speculation width), and, per bin of
// text length, the tile shapes a pair may take (lanes x words per lane,
// which sets the widest window over the band) and the period of the
// window's rebase at the DP edge. None of them changes a result (every
// setting computes the same greedy clustering), so the tuner times whole
// clusterings of the caller's store and keeps the fastest setting; it also
// checks that every trial returned the same clusters. No counterpart in the
// reference, whose launch shapes are fixed (func.cpp:695-781).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <vector>

#include "driver.hpp"

namespace gsdr {

namespace {

struct Knob {
  int Tuning::*field;
  std::vector<int> values;
};

bool same(const Tuning& a, const Tuning& b) {
  bool eq = a.swap_ratio == b.swap_ratio && a.spec_max == b.spec_max && a.filter_blocks == b.filter_blocks &&
            a.align_blocks == b.align_blocks && a.fork_shapes == b.fork_shapes &&
            a.list_factor == b.list_factor && a.pigeon == b.pigeon && a.screen == b.screen &&
            a.win_period == b.win_period && a.split_rows == b.split_rows;
  for (int i = 0; i < kLenBins; ++i)
    eq = eq && a.shapes.mask[i] == b.shapes.mask[i] && a.shapes.bias[i] == b.shapes.bias[i];
  return eq;
}

// Shape choices tried per length bin: the widths between the powers of two
// left out, the one-word tiles narrower than a wave left out (such pairs
// take a 64-lane tile), and one allowed shape wider than the narrowest.
struct ShapeTry {
  uint16_t mask;
  uint8_t bias;
};
constexpr uint16_t kPow2Shapes = kAllShapes & ~((1u << 5) | (1u << 7) | (1u << 9) | (1u << 10));
constexpr uint16_t kNoNarrow = kAllShapes & ~7u;
constexpr ShapeTry kShapeTries[] = {{kPow2Shapes, 0}, {kNoNarrow, 0}, {0, 1}};
// Lower edge of each length bin (len_bin): a bin is tuned only when the
// store holds a sequence that long.
constexpr int kBinLow[kLenBins] = {0, 512, 2048, 8192};

}  // namespace

int autotune(gsdr_ctx* c, gsdr_store& s, const gsdr_params& p0, int max_trials, gsdr_tuning* out) {
  if (max_trials <= 0) max_trials = 40;
  gsdr_params p = p0;
  p.mode = GSDR_MODE_BATCHED;
  p.timing = 0;
  p.tune = 1;
  const Tuning saved = c->tuned;
  const bool had = c->has_tuning;
  const size_t n = (size_t)s.n;
  std::vector<int32_t> ref(n), got(n);
  // one clustering under knobs T; its clusters must equal the first run's
  auto run = [&](const Tuning& T, double* ms, bool first) -> int {
    c->tuned = T;
    c->has_tuning = true;
    const auto t0 = std::chrono::steady_clock::now();
    int st = cluster_store(c, s, p, first ? ref.data() : got.data(), nullptr);
    *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (st) return st;
    if (!first && std::memcmp(ref.data(), got.data(), n * sizeof(int32_t)) != 0)
      return fail(GSDR_E_DEVICE, "autotune: clusters differ between launch settings");
    return GSDR_OK;
  };
  auto finish = [&](int st) {
    if (st) {
      c->tuned = saved;
      c->has_tuning = had;
    }
    return st;
  };
  Tuning best;
  double ms = 0, ms_default = 0;
  int st = run(best, &ms, true);  // warm-up: workspace, graph, caches
  if (st) return finish(st);
  if ((st = run(best, &ms_default, false))) return finish(st);
  double ms_best = ms_default;
  int trials = 1;
  const Knob knobs[] = {
      {&Tuning::list_factor, {8, 16}},
      {&Tuning::swap_ratio, {2, 3, 6, 8, 0}},
      {&Tuning::spec_max, {16, 32}},
      {&Tuning::align_blocks, {2, 4, 8, 0}},
      {&Tuning::filter_blocks, {4, 8, 16, 32}},
      {&Tuning::fork_shapes, {0, 1}},
      {&Tuning::split_rows, {0, 128, 1024}},
  };
  for (const Knob& k : knobs) {
    for (int v : k.values) {
      if (trials >= max_trials) break;
      Tuning T = best;
      if (T.*k.field == v) continue;
      T.*k.field = v;
      if ((st = run(T, &ms, false))) return finish(st);
      ++trials;
      // keep a change only when it is clearly faster than the run-to-run noise
      if (ms < 0.98 * ms_best) {
        best = T;
        ms_best = ms;
      }
    }
  }
  // per length bin: the tile shapes a pair may take
  for (int b = 0; b < kLenBins; ++b) {
    if (s.max_m < kBinLow[b]) continue;
    for (const ShapeTry& sv : kShapeTries) {
      if (trials >= max_trials) break;
      Tuning T = best;
      if (T.shapes.mask[b] == sv.mask && T.shapes.bias[b] == sv.bias) continue;
      T.shapes.mask[b] = sv.mask;
      T.shapes.bias[b] = sv.bias;
      if ((st = run(T, &ms, false))) return finish(st);
      ++trials;
      if (ms < 0.98 * ms_best) {
        best = T;
        ms_best = ms;
      }
    }
  }
  // the period of the window's rebase at the DP edge (64-lane tiles)
  for (int v : {1, 8}) {
    if (trials >= max_trials) break;
    Tuning T = best;
    T.win_period = v;
    if ((st = run(T, &ms, false))) return finish(st);
    ++trials;
    if (ms < 0.98 * ms_best) {
      best = T;
      ms_best = ms;
    }
  }
  // confirm a changed setting once more against the defaults
  if (trials < max_trials && !same(best, Tuning())) {
    double again = 0;
    if ((st = run(best, &again, false))) return finish(st);
    ++trials;
    if (again >= ms_default) {
      best = Tuning();
      ms_best = ms_default;
    } else {
      ms_best = std::min(ms_best, again);
    }
  }
  c->tuned = best;
  c->has_tuning = true;
  if (out) {
    std::memset(out, 0, sizeof(*out));
    out->swap_ratio = best.swap_ratio;
    out->spec_max = best.spec_max;
    out->filter_blocks = best.filter_blocks;
    out->align_blocks = best.align_blocks;
    out->fork_shapes = best.fork_shapes;
    out->list_factor = best.list_factor;
    out->pigeon = best.pigeon;
    out->screen = best.screen;
    tuning_shapes_out(best, out);
    out->trials = trials;
    out->ms_default = ms_default;
    out->ms_tuned = ms_best;
  }
  return GSDR_OK;
}

}  // namespace gsdr
