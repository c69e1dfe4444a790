// gfx950 kernels of the per-round pipeline: shared-word/composition filter,
// banded bit-parallel LCS decision, and the resolve step that fixes the
// representatives of a step, assigns clusters and compacts the candidates.
//
// Reference stages replaced: kernel_makeTable/kernel_cleanTable,
// kernel_preFilter, kernel_filter, kernel_dynamic (func.cpp:499-676) and the
// host loops updateRepresentative/updateRemain/updatJobs (func.cpp:466-498).
#define GSDR_DEVICE_TU
#include "kernels.hpp"
#include "kernels_impl.hpp"

namespace gsdr {

// ----------------------------------------------------------------- resolve
// One 1024-thread workgroup per job, after the step's alignments.
//
// Speculative step with representatives r_0 < ... < r_{K-1} (the K smallest
// unclustered indices): the greedy recursion of func.cpp:466-784 makes r_i a
// representative iff no earlier true representative r_j accepts it, and puts
// every later q into the cluster of the first true representative that
// accepts it. acc[q] bit i holds P(r_i, q). With K = 1 this is exactly one
// reference round: r_0 is the representative and q joins it iff P(r_0, q).
// The unassigned candidates are then compacted in order (updateRemain,
// func.cpp:477-487), so the next step's representatives are again the
// smallest unclustered indices (updateRepresentative, func.cpp:466-476).
//
// The step runs in two phases. Phase A aligns the speculative
// representatives against each other only; resolve_reps_kernel then fixes
// which slots are true representatives (tmask). Phase B aligns the true
// representatives against the remaining candidates and resolve_kernel
// assigns and compacts them.
__global__ __launch_bounds__(64) void resolve_reps_kernel(ResolveArgs a) {
  const int tid = threadIdx.x;
  if (blockIdx.x == 0)
    for (int i = tid; i < a.ncounters; i += 64) a.counters[i] = 0;
  const JobDev& J = a.jobs[blockIdx.x];
  JobState* S = J.state;
  if (S->done) return;
  const int K = S->k_spec;
  const int32_t* cand = J.cand[S->cur];
  // lane i holds slot i's representative and accept mask (bits of earlier
  // slots only: phase A pairs each slot with the later ones)
  int r = 0;
  unsigned long long acc = 0;
  if (tid < K) {
    r = cand[tid];
    acc = J.acc[r];
  }
  // the chain over the slots, the same in every lane
  unsigned long long tm = 0;
  for (int i = 0; i < K; ++i) {
    const unsigned long long ai = __shfl(acc, i);
    if ((ai & tm) == 0) tm |= 1ull << i;
  }
  if (tid < K) {
    const unsigned long long msk = acc & tm;
    J.cluster[r] = msk == 0 ? r : cand[__builtin_ctzll(msk)];
    J.acc[r] = 0;
  }
  if (tid == 0) {
    S->tmask = tm;
    S->kt = __popcll(tm);
    S->nonrep = K - __popcll(tm);
  }
}

// Phase-B positions K..nc-1 of every running job in tiles of kResolveTile:
// the prefix of tile counts over the jobs, in LDS; returns the total.
__device__ int resolve_tiles(const ResolveArgs& a, int* s_pre, int* s_part) {
  const int tid = threadIdx.x, nj = a.njobs;
  int carry = 0;
  for (int base = 0; base < nj; base += 256) {
    const int j = base + tid;
    int u = 0;
    if (j < nj) {
      const JobState* S = a.jobs[j].state;
      const int rest = S->done ? 0 : S->nc - S->k_spec;
      u = rest > 0 ? (rest + kResolveTile - 1) / kResolveTile : 0;
    }
    int tot;
    const int ex = block_exclusive_scan<256>(u, &tot, s_part);
    if (j < nj) s_pre[j] = carry + ex;
    carry += tot;
  }
  if (tid == 0) s_pre[nj] = carry;
  __syncthreads();
  return s_pre[nj];
}

__device__ __forceinline__ int tile_job(const int* s_pre, int nj, int u) {
  int lo = 0, hi = nj - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (s_pre[mid] <= u) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Resolve, part 1 (after the step's alignments): every candidate that a
// true representative accepted joins the first such one; accept masks are
// cleared; each tile's survivors are counted.
__global__ __launch_bounds__(256) void resolve_mark_kernel(ResolveArgs a) {
  __shared__ int s_pre[kMaxJobsPerLaunch + 1];
  __shared__ int s_part[4];
  const int tid = threadIdx.x;
  if (blockIdx.x == 0)
    for (int i = tid; i < a.ncounters; i += 256) a.counters[i] = 0;
  const int total = resolve_tiles(a, s_pre, s_part);
  for (int u = blockIdx.x; u < total; u += gridDim.x) {
    const int job = tile_job(s_pre, a.njobs, u);
    const JobDev& J = a.jobs[job];
    JobState* S = J.state;
    const int K = S->k_spec, nc = S->nc;
    const int tile = u - s_pre[job];
    const int32_t* cand = J.cand[S->cur];
    const unsigned long long tm = S->tmask;
    int keep = 0, absorbed = 0;
#pragma unroll
    for (int i = 0; i < kResolveTile / 256; ++i) {
      const int p = K + tile * kResolveTile + i * 256 + tid;
      if (p < nc) {
        const int q = cand[p];
        const unsigned long long acc = J.acc[q];
        unsigned long long msk = 0;
        if (acc) {
          J.acc[q] = 0;
          msk = acc & tm;
        }
        if (msk) {
          J.cluster[q] = cand[__builtin_ctzll(msk)];
          ++absorbed;
        } else {
          ++keep;
        }
      }
    }
    keep = block_sum<256>(keep, s_part);
    absorbed = block_sum<256>(absorbed, s_part);
    if (tid == 0) {
      J.tcnt[tile] = keep;
      if (absorbed) atomicAdd(&S->absorbed, absorbed);
    }
  }
}

// Resolve, part 2: the survivors of each tile, in order, after those of the
// tiles before it (updateRemain, func.cpp:477-487); the job's last tile
// leaves the total.
__global__ __launch_bounds__(256) void resolve_scatter_kernel(ResolveArgs a) {
  __shared__ int s_pre[kMaxJobsPerLaunch + 1];
  __shared__ int s_part[4];
  const int tid = threadIdx.x;
  const int total = resolve_tiles(a, s_pre, s_part);
  constexpr int PER = kResolveTile / 256;
  for (int u = blockIdx.x; u < total; u += gridDim.x) {
    const int job = tile_job(s_pre, a.njobs, u);
    const JobDev& J = a.jobs[job];
    JobState* S = J.state;
    const int K = S->k_spec, nc = S->nc, cur = S->cur;
    const int tile = u - s_pre[job];
    const int tiles = (nc - K + kResolveTile - 1) / kResolveTile;
    int before = 0;
    for (int t = tid; t < tile; t += 256) before += J.tcnt[t];
    before = block_sum<256>(before, s_part);
    const int32_t* cand = J.cand[cur];
    int32_t* out = J.cand[cur ^ 1];
    const int p0 = K + tile * kResolveTile + tid * PER;
    int qv[PER], cnt = 0;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int p = p0 + i;
      qv[i] = -1;
      if (p < nc) {
        const int q = cand[p];
        if (J.cluster[q] < 0) {
          qv[i] = q;
          ++cnt;
        }
      }
    }
    int tot;
    int at = before + block_exclusive_scan<256>(cnt, &tot, s_part);
#pragma unroll
    for (int i = 0; i < PER; ++i)
      if (qv[i] >= 0) out[at++] = qv[i];
    if (tid == 0 && tile == tiles - 1) S->kept = before + tot;
  }
}

// Resolve, part 3: the round state of every job for the next step.
__global__ __launch_bounds__(256) void resolve_state_kernel(ResolveArgs a) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= a.njobs) return;
  const JobDev& J = a.jobs[j];
  JobState* S = J.state;
  if (S->done) return;
  const int K = S->k_spec, nc = S->nc, cur = S->cur;
  const int32_t* cand = J.cand[cur];
  const unsigned long long tm = S->tmask;
  // the true representatives are their own clusters (with one slot this is
  // the round's representative, updateRepresentative, func.cpp:466-476);
  // in the batched form resolve_reps_kernel has written them
  if (!a.reps_marked) {
    for (unsigned long long m = tm; m; m &= m - 1) {
      const int r = cand[__builtin_ctzll(m)];
      J.cluster[r] = r;
    }
  }
  const int true_reps = __popcll(tm);
  const int absorbed = S->absorbed;
  const int nc2 = nc > K ? S->kept : 0;
  S->absorbed = 0;
  S->kept = 0;
  S->reps += true_reps;
  S->steps += 1;
  S->cur = cur ^ 1;
  S->nc = nc2;
  int k2 = 1;
  if (S->k_max > 1) {
    // f = absorption per true representative and candidate. A step of K
    // slots evaluates about K*f/2 pairs per candidate whose candidate an
    // earlier slot already took (fewer where the align fetch skips them),
    // and phase A costs K^2/2 pairs: aim at K*f near 1/2.
    const double denom = (double)max(true_reps, 1) * (double)max(nc - K, 1);
    const double f = (double)absorbed / denom;
    const int up = min(2 * K, S->k_max), down = max(1, K / 2);
    if (f <= 0.0) {
      k2 = up;
    } else {
      const double want = 0.5 / f;
      k2 = want >= up ? up : (want <= down ? down : (int)want);
    }
  }
  // every survivor of a step may land in one shape list: k_spec * nc of
  // all jobs must stay within its capacity (this job alone here; the
  // clamp that follows the step bounds the sum over the jobs)
  const long long cap = a.pair_budget / (long long)max(nc2, 1);
  if ((long long)k2 > cap) k2 = (int)cap;
  if (k2 > nc2) k2 = nc2;
  if (k2 < 1) k2 = 1;
  S->k_spec = k2;
  // with one slot, phase A has nothing to decide: slot 0 is the
  // representative (updateRepresentative, func.cpp:466-476)
  S->tmask = 1ull;
  S->kt = 1;
  S->first = nc2 > 0 ? J.cand[cur ^ 1][0] : -1;
  if (nc2 == 0) {
    S->done = 1;
    if (a.jobs_left) atomicSub(a.jobs_left, 1);
  }
}

// Bounds k_spec of every running job so that k_spec * nc summed over the
// jobs fits the alignment lists: with T the total of nc over the running
// jobs, k_spec <= pair_budget / T gives sum(k_spec * nc) <= pair_budget.
// Runs after every step's resolve (whose K choice ignores the other jobs)
// and when jobs join a table. One block; any k_spec is exact.
__global__ __launch_bounds__(1024) void clamp_spec_kernel(const JobDev* __restrict__ jobs, int njobs,
                                                          long long pair_budget) {
  __shared__ int s_part[16];
  long long tot = 0;
  for (int base = 0; base < njobs; base += 1024) {
    const int j = base + threadIdx.x;
    int v = 0;
    if (j < njobs) {
      const JobState* S = jobs[j].state;
      if (!S->done) v = S->nc;
    }
    tot += block_sum<1024>(v, s_part);
  }
  const long long cap = pair_budget / (tot > 0 ? tot : 1ll);
  for (int j = threadIdx.x; j < njobs; j += 1024) {
    JobState* S = jobs[j].state;
    if (S->done) continue;
    int k = S->k_spec;
    if ((long long)k > cap) k = (int)max(cap, 1ll);
    if (k > S->nc) k = S->nc;
    if (k < 1) k = 1;
    if (k != S->k_spec) {
      S->k_spec = k;
      S->tmask = 1ull;
      S->kt = 1;
    }
  }
}

hipError_t launch_clamp_spec(const JobDev* jobs, int njobs, long long pair_budget, hipStream_t s) {
  hipLaunchKernelGGL(clamp_spec_kernel, dim3(1), dim3(1024), 0, s, jobs, njobs, pair_budget);
  return hipGetLastError();
}

// Initial state of every job: all clusters -1, candidates 0..n-1, one
// representative slot.
__global__ __launch_bounds__(256) void init_jobs_kernel(const JobDev* __restrict__ jobs, int k_max) {
  const JobDev& J = jobs[blockIdx.y];
  const int n = J.n;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    J.cluster[i] = -1;
    J.acc[i] = 0ull;
    J.cand[0][i] = i;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    JobState* S = J.state;
    S->cur = 0;
    S->nc = n;
    S->k_spec = 1;
    S->done = n > 0 ? 0 : 1;
    S->kt = 1;
    S->tmask = 1ull;
    S->nonrep = S->absorbed = S->kept = 0;
    S->pad2 = 0ull;
    S->steps = 0;
    S->reps = 0;
    S->k_max = k_max;
    S->first = n > 0 ? 0 : -1;
  }
}

hipError_t launch_pigeon(const PigeonArgs& a, int grid, hipStream_t s) {
  hipLaunchKernelGGL(pigeon_kernel, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_screen(const ScreenArgs& a, int grid, hipStream_t s) {
  hipLaunchKernelGGL(screen_kernel, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_signatures(const uint32_t* kp, const int64_t* kpoff, int n, int k, uint32_t* sig, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(signature_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, kp, kpoff, n, k, sig);
  return hipGetLastError();
}

// ---------------------------------------------------------- profile totals
// One wave per sequence: T_s = sum of the counts (modulo 2^16, as stored) of
// its distinct k-mers. The shared-word sum with any representative is at most
// what was summed so far plus the counts not yet read, which bounds the scan
// of the one-table filter.
__global__ __launch_bounds__(256) void profile_total_kernel(const uint32_t* __restrict__ kp,
                                                            const int64_t* __restrict__ kpoff, int n,
                                                            int32_t* __restrict__ ksum) {
  const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (s >= n) return;
  const int lane = threadIdx.x & 63;
  int t = 0;
  for (int64_t e = kpoff[s] + lane; e < kpoff[s + 1]; e += 64) t += (int)(kp[e] >> 16);
  t = wave_sum(t);
  if (lane == 0) ksum[s] = t;
}

hipError_t launch_profile_totals(const uint32_t* kp, const int64_t* kpoff, int n, int32_t* ksum, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(profile_total_kernel, dim3((n + 3) / 4), dim3(256), 0, s, kp, kpoff, n, ksum);
  return hipGetLastError();
}

// ----------------------------------------------------------------- cutoffs
// createCutoff (func.cpp:339-374) in float32, operation for operation like
// the reference: bc = ceilf((float)L * t) (func.cpp:368-369) and
// wc = max(L-k+1 - (int)(ceilf((float)L * (1.0f-t)) * k), 1) with the offset
// factor forced to 1 (func.cpp:353-366). Each product and difference is
// rounded on its own (no contraction), so the values equal the host's.
__global__ __launch_bounds__(256) void cutoff_kernel(const int32_t* __restrict__ raw_len, int n, float t, int k,
                                                     int32_t* __restrict__ bc, int32_t* __restrict__ wc,
                                                     SeqRec* __restrict__ rec) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int L = raw_len[i];
  const float fl = (float)L;
  const int b = (int)ceilf(__fmul_rn(fl, t));
  const float ce = ceilf(__fmul_rn(fl, __fsub_rn(1.0f, t)));
  const int cut = (int)__fmul_rn(ce, (float)k);
  const int req = L - k + 1 - cut;
  const int w = req < 1 ? 1 : req;
  bc[i] = b;
  wc[i] = w;
  if (rec) {
    rec[i].bc = b;
    rec[i].wc = w;
  }
}

hipError_t launch_cutoffs(const int32_t* raw_len, int n, float t, int k, int32_t* bc, int32_t* wc, SeqRec* rec,
                          hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(cutoff_kernel, dim3((n + 255) / 256), dim3(256), 0, s, raw_len, n, t, k, bc, wc, rec);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void pack_records_kernel(const int32_t* __restrict__ m, const uint4* __restrict__ comp,
                                                           const int64_t* __restrict__ kpoff,
                                                           const int32_t* __restrict__ ksum, int n,
                                                           SeqRec* __restrict__ rec) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint4 cp = comp[i];
  SeqRec r;
  r.m = m[i];
  r.bc = 0;  // set with the threshold (cutoff_kernel)
  r.wc = 0;
  r.ksum = ksum[i];
  r.comp_ac = cp.x | (cp.y << 16);
  r.comp_gt = cp.z | (cp.w << 16);
  r.kp0 = (uint32_t)kpoff[i];
  r.kn = (uint32_t)(kpoff[i + 1] - kpoff[i]);
  rec[i] = r;
}

hipError_t launch_pack_records(const int32_t* m, const uint4* comp, const int64_t* kpoff, const int32_t* ksum, int n,
                               SeqRec* rec, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(pack_records_kernel, dim3((n + 255) / 256), dim3(256), 0, s, m, comp, kpoff, ksum, n, rec);
  return hipGetLastError();
}

// -------------------------------------------------------------- launchers
template <int G, bool SCR>
static hipError_t launch_filter_gs(const FilterArgs& a, int max_k, int grid, hipStream_t s) {
  constexpr int ts = filter_table_stride(G);
  if (G > 1 && (1 << (2 * max_k)) > ts) return hipErrorInvalidValue;  // the tables would overlap
  const size_t lds = G > 1 ? (size_t)G * ts * sizeof(uint16_t) : (sizeof(uint16_t) << (2 * max_k));
  static bool attr_set = false;
  if (lds > 32768 && !attr_set) {
    hipError_t e = hipFuncSetAttribute((const void*)filter_kernel<G, SCR>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)(sizeof(uint16_t) << 16));
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  hipLaunchKernelGGL((filter_kernel<G, SCR>), dim3(grid), dim3(256), lds, s, a);
  return hipGetLastError();
}

// Phase B of the long-word groups has a form for the signature screen.
template <int G>
static hipError_t launch_filter_g(const FilterArgs& a, int max_k, int grid, hipStream_t s) {
  if constexpr (G <= filter_group(kScreenMinK)) {
    if (a.screen && a.phase == PHASE_MAIN) return launch_filter_gs<G, true>(a, max_k, grid, s);
  }
  return launch_filter_gs<G, false>(a, max_k, grid, s);
}

static hipError_t launch_filter_k(const FilterArgs& a, int max_k, int grid, hipStream_t s) {
  switch (filter_group(max_k)) {
    case 16: return launch_filter_g<16>(a, max_k, grid, s);
    case 8: return launch_filter_g<8>(a, max_k, grid, s);
    case 4: return launch_filter_g<4>(a, max_k, grid, s);
    case 2: return launch_filter_g<2>(a, max_k, grid, s);
    default: return launch_filter_g<1>(a, max_k, grid, s);
  }
}

// Jobs of different word lengths sharing one launch would all get the
// table group of the longest k (one table per block for k = 7); instead each
// run of word lengths with the same group size gets its own launch, whose
// tables are sized for the longest k of that run.
hipError_t launch_filter(const FilterArgs& a, unsigned kmask, int grid, hipStream_t s) {
  int k = 0;
  while (k < 32 && !((kmask >> k) & 1u)) ++k;
  while (k < 32) {
    const int g = filter_group(k);
    int hi = k;
    for (int k2 = k + 1; k2 < 32 && filter_group(k2) == g; ++k2)
      if ((kmask >> k2) & 1u) hi = k2;
    FilterArgs b = a;
    b.k_lo = k;
    b.k_hi = hi;
    hipError_t e = launch_filter_k(b, hi, grid, s);
    if (e != hipSuccess) return e;
    k = hi + 1;
    while (k < 32 && !((kmask >> k) & 1u)) ++k;
  }
  return hipSuccess;
}

template <int S, int W>
static hipError_t launch_align_sw(const AlignArgs& a, int grid, hipStream_t s) {
  hipLaunchKernelGGL((align_kernel<S, W>), dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

template <int S, int W>
static int occupancy_sw() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void*)align_kernel<S, W>, 256, 0) != hipSuccess)
    return 4;
  return nb > 0 ? nb : 1;
}

// Resident 256-thread blocks per CU of each tile-shape kernel (registers and
// LDS decide it), asked once: a grid of that many blocks per CU fills the
// chip without queueing blocks that would only find the work list empty.
int align_blocks_per_cu(int shape) {
  static int occ[kNumShapes] = {};
  static bool ready = false;
  if (!ready) {
    occ[0] = occupancy_sw<8, 1>();
    occ[1] = occupancy_sw<16, 1>();
    occ[2] = occupancy_sw<32, 1>();
    occ[3] = occupancy_sw<64, 1>();
    occ[4] = occupancy_sw<64, 2>();
    occ[5] = occupancy_sw<64, 3>();
    occ[6] = occupancy_sw<64, 4>();
    occ[7] = occupancy_sw<64, 6>();
    occ[8] = occupancy_sw<64, 8>();
    occ[9] = occupancy_sw<64, 10>();
    occ[10] = occupancy_sw<64, 12>();
    occ[11] = occupancy_sw<64, 16>();
    ready = true;
  }
  return shape >= 0 && shape < kNumShapes ? occ[shape] : 1;
}

hipError_t launch_align_split(const AlignArgs& a, int grid, hipStream_t s) {
  hipLaunchKernelGGL(align_split_kernel, dim3(grid), dim3(128), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_align(int shape, const AlignArgs& a, int grid, hipStream_t s) {
  switch (shape) {
    case 0: return launch_align_sw<8, 1>(a, grid, s);
    case 1: return launch_align_sw<16, 1>(a, grid, s);
    case 2: return launch_align_sw<32, 1>(a, grid, s);
    case 3: return launch_align_sw<64, 1>(a, grid, s);
    case 4: return launch_align_sw<64, 2>(a, grid, s);
    case 5: return launch_align_sw<64, 3>(a, grid, s);
    case 6: return launch_align_sw<64, 4>(a, grid, s);
    case 7: return launch_align_sw<64, 6>(a, grid, s);
    case 8: return launch_align_sw<64, 8>(a, grid, s);
    case 9: return launch_align_sw<64, 10>(a, grid, s);
    case 10: return launch_align_sw<64, 12>(a, grid, s);
    case 11: return launch_align_sw<64, 16>(a, grid, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_resolve_reps(const ResolveArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(resolve_reps_kernel, dim3(a.njobs), dim3(64), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_resolve(const ResolveArgs& a, hipStream_t s) {
  const int grid = a.grid > 0 ? a.grid : 1;
  hipLaunchKernelGGL(resolve_mark_kernel, dim3(grid), dim3(256), 0, s, a);
  hipLaunchKernelGGL(resolve_scatter_kernel, dim3(grid), dim3(256), 0, s, a);
  hipLaunchKernelGGL(resolve_state_kernel, dim3((a.njobs + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_init_jobs(const JobDev* jobs_dev, const JobDev* jobs_host, int njobs, int k_max,
                            int32_t* jobs_left, hipStream_t s) {
  (void)jobs_left;
  int max_n = 1;
  for (int j = 0; j < njobs; ++j) max_n = max(max_n, jobs_host[j].n);
  hipLaunchKernelGGL(init_jobs_kernel, dim3((max_n + 255) / 256, njobs), dim3(256), 0, s, jobs_dev, k_max);
  return hipGetLastError();
}

}  // namespace gsdr
