// Device code shared by the plain and batched launch forms.
#pragma once
#include "kernels.hpp"

namespace gsdr {

constexpr int kMaxJobsPerLaunch = 1024;

struct FilterArgs {
  const JobDev* jobs;
  int32_t njobs;
  int32_t min_shape;
  int32_t swap_ratio;       // role swap when n_rep >= swap_ratio * m_cand (0 = never)
  int32_t k_lo, k_hi;       // this launch filters the jobs with word length in [k_lo, k_hi]
  int32_t phase;            // FilterPhase
  AlignItem* items;         // per-shape lists: items + shape * item_cap
  int64_t item_cap;
  int32_t* bin_count;       // kNumShapes counters
  unsigned long long* stats;
  int32_t screen;           // phase B: skip what the signature screen rejected
  ShapeTable shapes = {};   // tile-shape choice per text-length bin (zero: narrowest of all)
  int32_t split_rows = 0;   // pairs with this many rows or more go to the split list (0 = none;
                            // at least 64, so that each half holds 32 rows)
  int64_t split_cap = 0;    // pairs the split list takes (it follows the kNumShapes shape lists);
                            // later ones go to their shape lists
};

struct ScreenArgs {
  const JobDev* jobs;
  int32_t njobs;
  unsigned long long* stats;
};

struct PigeonArgs {
  const JobDev* jobs;
  AlignItem* items;         // the per-shape lists: items + shape * item_cap
  int64_t item_cap;
  const int32_t* bin_count; // items per shape
  int32_t smax;             // shapes 0..smax
  unsigned long long* stats;
};

struct AlignArgs {
  const JobDev* jobs;
  const AlignItem* items;   // this shape's list
  const int32_t* count;     // number of items (device)
  int32_t* work;            // work counter (device, starts at 0)
  int32_t* result;          // non-null: result[item] = 1 accept / 0 reject (pair-level entry)
  unsigned long long* stats;
  int32_t skip_absorbed;    // skip (slot, q) when an earlier true representative already accepted q
  int32_t win_period = 0;   // 64-lane tiles: 32-row blocks between DP-edge window rebases, a power of
                            // two (0 = 2 for 6 or more words per lane, 4 for fewer)
};

struct ResolveArgs {
  const JobDev* jobs;
  int32_t njobs;
  int32_t* counters;        // bin counts and work counters, reset for the next step
  int32_t ncounters;
  int32_t* jobs_left;       // jobs not done (device)
  unsigned long long* stats;
  int64_t pair_budget;      // alignment list capacity per shape (bounds k_spec * nc over the jobs)
  int32_t reps_marked = 0;  // resolve_reps_kernel has already made the true representatives their own
                            // clusters (batched form)
  int32_t grid;             // workgroups of the resolve tile kernels
};

// ----------------------------------------------------------- block helpers
__device__ inline int wave_sum(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

template <int NT>
__device__ inline int block_sum(int v, int* part) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) part[wave] = v;
  __syncthreads();
  int s = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) s += part[w];
  return s;
}

// Exclusive scan over an NT-thread block; *total gets the block sum.
template <int NT>
__device__ inline int block_exclusive_scan(int v, int* total, int* part) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(incl, off);
    if (lane >= off) incl += y;
  }
  __syncthreads();
  if (lane == 63) part[wave] = incl;
  __syncthreads();
  int before = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) {
    const int x = part[w];
    if (w < wave) before += x;
    tot += x;
  }
  *total = tot;
  return before + incl - v;
}

// This wave's copy of the striped counters, picked by the compute unit the
// wave runs on and its place in the block (read when the counters are
// flushed, so nothing stays live through the kernel for it).
__device__ __forceinline__ unsigned long long* stat_row(unsigned long long* st) {
  return st + (((__smid() << 2) + (threadIdx.x >> 6)) % kStatStripes) * kStatStride;
}

// Kernel bodies are compiled only in the translation unit that defines
// GSDR_DEVICE_TU (kernels.hip); the other files see the launchers.
#ifdef GSDR_DEVICE_TU
// ------------------------------------------------------------------ filter
// Work unit = (job, group of up to G representative slots, chunk of
// kFilterChunk candidate positions). A 256-thread block stages, for the G
// representatives of its unit, the k-mer counts c_r(w) as u16 tables of 4^k
// entries in LDS plus their id, length and composition (rebuilt only when
// the unit's group changes). Each 16-lane group owns one candidate q at a
// position after each representative's slot and reads q's record and
// distinct k-mer profile once for the whole group:
//   length check (added, sound): c = bc_q + 1 > min(n, m_q) means LCS < c;
//   composition (func.cpp:525-543): reject sum < bc_q, plus the sound
//     tightening sum == bc_q (LCS <= sum < bc_q + 1);
//   shared words (func.cpp:544-571): sum over q's distinct k-mers of
//     min(c_r(w), c_q(w)) < wc_q rejects (one LDS lookup per representative
//     of the group that passed the checks above).
// Survivors are appended to the alignment list of their tile shape.
// Waves per SIMD the register allocation must allow: the tables take 32 KiB
// of LDS per block, so four blocks (four waves per SIMD) fit a CU; the
// widest group is held to that too.
__host__ __device__ constexpr int filter_waves_per_eu(int g) { return g >= 4 ? 4 : 1; }

// Entries per representative table in LDS: a compile-time stride, so every
// lookup addresses its table with an immediate offset (32 KiB for the group;
// 4^k must fit, which filter_group guarantees: G = 16 for k <= 5, 4 for k = 6;
// one table of 4^k entries for k >= 7).
__host__ __device__ constexpr int filter_table_stride(int g) { return g > 1 ? 16384 / g : 0; }

// SCR: the phase-B form for jobs with the signature screen (see below).
template <int G, bool SCR>
__global__ __launch_bounds__(256, filter_waves_per_eu(G)) void filter_kernel(FilterArgs a) {
  static_assert(G >= 1 && G <= 16, "one lane of a 16-lane group per representative");
  static_assert(!SCR || G <= filter_group(kScreenMinK), "the screen covers the long-word groups only");
  constexpr int kTS = filter_table_stride(G);
  HIP_DYNAMIC_SHARED(uint32_t, lds)
  uint16_t* table = reinterpret_cast<uint16_t*>(lds);
  __shared__ int s_pre[kMaxJobsPerLaunch + 1];
  __shared__ int s_part[4];
  __shared__ int s_slot[G], s_rep[G], s_n[G];
  __shared__ uint32_t s_kp0[G], s_kn[G];
  __shared__ uint4 s_comp[G];
  __shared__ int s_act[SCR ? 256 : 1];  // active units of a batch (screened phase B)
  const int tid = threadIdx.x;
  const int nj = a.njobs;
  // prefix of work units over jobs
  int total = 0;
  {
    int carry = 0;
    for (int base = 0; base < nj; base += 256) {
      const int j = base + tid;
      int u = 0;
      if (j < nj) {
        const JobDev& J = a.jobs[j];
        u = (J.k >= a.k_lo && J.k <= a.k_hi) ? phase_units(*J.state, a.phase, G) : 0;
      }
      int tot;
      const int ex = block_exclusive_scan<256>(u, &tot, s_part);
      if (j < nj) s_pre[j] = carry + ex;
      carry += tot;
    }
    if (tid == 0) s_pre[nj] = carry;
    __syncthreads();
    total = s_pre[nj];
  }
  if (total == 0) return;
  const int per = (total + gridDim.x - 1) / gridDim.x;
  const int u0 = blockIdx.x * per;
  const int u1 = min(total, u0 + per);
  const int grp = tid >> 4, gl = tid & 15;
  int ckey = -1, gcount = 0;  // job * 64 + group of the staged tables
  unsigned long long rej_len = 0, rej_comp = 0, rej_word = 0, passed = 0, bytes = 0, cband = 0, cfull = 0;
  // Phase B with the signature screen (groups of the screened word lengths):
  // the block checks 256 units at a time, one per thread, and runs only
  // those where some representative of the unit's group passed the screen
  // against some candidate of its chunk (the screen has counted the pairs of
  // the others). Without the screen every unit of the range runs.
  constexpr bool compact = SCR;
  for (int ub = u0; ub < u1; ub += compact ? 256 : u1 - u0) {
    int nact = u1 - u0;
    if (compact) {
      const int u = ub + tid;
      bool act = false;
      if (u < u1) {
        int lo = 0, hi = nj - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (s_pre[mid] <= u) lo = mid; else hi = mid - 1;
        }
        const JobDev& J = a.jobs[lo];
        const JobState* S = J.state;
        act = true;
        if (J.scr != nullptr && !(J.flags & kJobDropFilters)) {
          const int chunks = (S->nc - S->k_spec + filter_chunk(G) - 1) / filter_chunk(G);
          const int lu = u - s_pre[lo];
          const int gi = lu / chunks;
          // the group's slots: the true representatives of ranks gi*G ..
          unsigned long long tm = S->tmask;
          for (int i = 0; i < gi * G && tm; ++i) tm &= tm - 1;
          unsigned long long gslots = 0;
          for (int g = 0; g < G && tm; ++g) {
            gslots |= tm & (~tm + 1ull);
            tm &= tm - 1;
          }
          act = (J.chm[lu - gi * chunks] & gslots) != 0;
        }
      }
      int at = block_exclusive_scan<256>(act ? 1 : 0, &nact, s_part);
      if (act) s_act[at] = u;
      __syncthreads();
    }
  for (int ai = 0; ai < nact; ++ai) {
    const int u = compact ? s_act[ai] : ub + ai;
    // job of unit u: last j with s_pre[j] <= u
    int lo = 0, hi = nj - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_pre[mid] <= u) lo = mid; else hi = mid - 1;
    }
    const int job = lo;
    const JobDev& J = a.jobs[job];
    const JobState* S = J.state;
    const int nc = S->nc;
    const int K = S->k_spec;
    const int32_t* cand = J.cand[S->cur];
    const int p0 = a.phase == PHASE_REPS ? 0 : K;
    const int p1 = a.phase == PHASE_REPS ? K : nc;
    const int members = a.phase == PHASE_REPS ? K : S->kt;
    const int chunks = (p1 - p0 + filter_chunk(G) - 1) / filter_chunk(G);
    const int lu = u - s_pre[job];
    const int gi = lu / chunks;
    const int chunk = lu - gi * chunks;
    // a screened job's pairs that the screen rejected are skipped (counted there)
    const bool screened = compact && J.scr != nullptr && !(J.flags & kJobDropFilters);
    if (job * 64 + gi != ckey) {
      // stage the group: slots (phase A: ranks are slots; phase B: the
      // rank-th true representative), records and k-mer tables
      const int nw = 1 << (2 * J.k);
      gcount = min(G, members - gi * G);
      __syncthreads();
      for (int i = tid; i < (G > 1 ? G * kTS : nw) / 2; i += 256) lds[i] = 0u;
      if (tid < gcount) {
        const int rank = gi * G + tid;
        int slot = rank;
        if (a.phase == PHASE_MAIN) {
          unsigned long long tm = S->tmask;
          for (int i = 0; i < rank; ++i) tm &= tm - 1;
          slot = __builtin_ctzll(tm);
        }
        const int r = cand[slot];
        s_slot[tid] = slot;
        s_rep[tid] = r;
        const SeqRec rr = J.st.rec[r];
        s_n[tid] = rr.m;
        s_comp[tid] = make_uint4(rr.comp_ac & 0xFFFFu, rr.comp_ac >> 16, rr.comp_gt & 0xFFFFu, rr.comp_gt >> 16);
        s_kp0[tid] = rr.kp0;
        s_kn[tid] = rr.kn;
      }
      __syncthreads();
      for (int g = 0; g < gcount; ++g) {
        const int64_t k0 = s_kp0[g], k1 = k0 + s_kn[g];
        uint16_t* tg = table + g * kTS;
        for (int64_t e = k0 + tid; e < k1; e += 256) {
          const uint32_t x = J.st.kp[e];
          tg[x & 0xFFFFu] = (uint16_t)(x >> 16);
        }
      }
      __syncthreads();
      ckey = job * 64 + gi;
    }
    // every lane of the block runs the whole unit (no early exit): the
    // survivor append below is a wave-wide operation. The NS candidates of a
    // group have their index and record loads issued together.
    constexpr int NS = filter_subs(G);
    const bool drop = (J.flags & kJobDropFilters) != 0;
    // each candidate's record: (m, bc, wc, ksum) and (composition, profile)
    int qv[NS];
    uint4 rav[NS], rbv[NS];
    unsigned long long svv[NS];  // slots that passed the screen against q
#pragma unroll
    for (int sub = 0; sub < NS; ++sub) {
      const int ps = p0 + chunk * filter_chunk(G) + sub * kFilterChunk + grp;
      qv[sub] = ps < p1 ? cand[ps] : 0;
    }
#pragma unroll
    for (int sub = 0; sub < NS; ++sub) {
      const uint4* rp = reinterpret_cast<const uint4*>(J.st.rec + qv[sub]);
      rav[sub] = rp[0];
      rbv[sub] = rp[1];
      svv[sub] = screened ? J.scr[qv[sub]] : ~0ull;
    }
#pragma unroll
    for (int sub = 0; sub < NS; ++sub) {
      const int pos = p0 + chunk * filter_chunk(G) + sub * kFilterChunk + grp;
      const bool inpos = pos < p1;
      const int q = qv[sub];
      const int mq = (int)rav[sub].x;
      const int bcq = (int)rav[sub].y;
      const int c = bcq + 1;
      const uint4 cq = make_uint4(rbv[sub].x & 0xFFFFu, rbv[sub].x >> 16, rbv[sub].y & 0xFFFFu, rbv[sub].y >> 16);
      if (gl == 0 && inpos) bytes += 36;  // id, record
      // length and composition per representative of the group
      uint32_t mask = 0;
#pragma unroll
      for (int g = 0; g < G; ++g) {
        if (!inpos || g >= gcount || pos <= s_slot[g]) continue;
        if (!((svv[sub] >> s_slot[g]) & 1ull)) continue;
        if (c > min(s_n[g], mq)) {
          if (gl == 0) ++rej_len;
          continue;
        }
        const uint4 cr = s_comp[g];
        const int sum = (int)(min(cr.x, cq.x) + min(cr.y, cq.y) + min(cr.z, cq.z) + min(cr.w, cq.w));
        if (!drop && sum <= bcq) {
          if (gl == 0) ++rej_comp;
          continue;
        }
        mask |= 1u << g;
      }
      // shared words: q's profile read once, one table lookup per survivor so
      // far (none when no representative is left; drop mode reads no profile
      // and passes every pair)
      const int64_t e0 = rbv[sub].z, e1 = (drop || mask == 0) ? e0 : e0 + rbv[sub].w;
      int acc[G];
#pragma unroll
      for (int g = 0; g < G; ++g) acc[g] = 0;
      // the profile loads are latency bound: NB, then 4, in flight per lane.
      // With 16 tables every lookup runs, whether or not its representative
      // is still in play (its total is discarded below): the per-table
      // branches would hold 16 divergent exec masks in scalar registers.
      const int wcq = drop ? 0 : (int)rav[sub].z;
      int64_t e = e0 + gl;
      bool cut = false;  // the same for the 16 lanes of a group
      if constexpr (G == 1) {
        // One representative: exact early reject. With S the shared-word sum
        // over q's entries of min(c_r, c_q), T_q the sum of q's stored counts
        // and S', C' those sums over the entries read so far,
        // S <= S' + (T_q - C'): once the deficit C' - S' exceeds T_q - wc_q
        // the pair fails the test (func.cpp:567) whatever the unread entries
        // hold. Full batches of 128 entries, each followed by a 16-lane sum
        // of the deficit (the batch count is the same for the whole group).
        constexpr int NB1 = 8;
        const int slack = mask ? (int)rav[sub].w - wcq : 0;
        cut = mask != 0 && slack < 0;
        int deficit = 0;
        int64_t b = e0;
        if (!cut) {
          for (; b + NB1 * 16 <= e1; b += NB1 * 16) {
            uint32_t xs[NB1];
#pragma unroll
            for (int i = 0; i < NB1; ++i) xs[i] = J.st.kp[b + gl + 16 * i];
            int sg = 0, dg = 0;
#pragma unroll
            for (int i = 0; i < NB1; ++i) {
              const uint32_t cnt = xs[i] >> 16;
              const uint32_t m = min((uint32_t)table[xs[i] & 0xFFFFu], cnt);
              sg += (int)m;
              dg += (int)(cnt - m);
            }
            acc[0] += sg;
            deficit += dg;
            int d = deficit;
#pragma unroll
            for (int off = 8; off > 0; off >>= 1) d += __shfl_xor(d, off, 16);
            if (d > slack) {
              cut = true;
              b += NB1 * 16;  // entries read
              break;
            }
          }
        }
        e = cut ? e1 : b + gl;
        if (gl == 0 && mask != 0) bytes += 4ull * (unsigned long long)((cut ? min(b, e1) : e1) - e0);
      } else {
        // the profile loads are latency bound: NB, then 4, in flight per
        // lane. With 16 tables every lookup runs, whether or not its
        // representative is still in play (its total is discarded below):
        // the per-table branches would hold 16 divergent exec masks in
        // scalar registers.
        constexpr int NB = 8;
        for (; e + (NB - 1) * 16 < e1; e += NB * 16) {
          uint32_t xs[NB];
#pragma unroll
          for (int i = 0; i < NB; ++i) xs[i] = J.st.kp[e + 16 * i];
#pragma unroll
          for (int g = 0; g < G; ++g) {
            if (G < 16 && !((mask >> g) & 1u)) continue;
            const uint16_t* tg = table + g * kTS;
            int sg = 0;
#pragma unroll
            for (int i = 0; i < NB; ++i) sg += min((uint32_t)tg[xs[i] & 0xFFFFu], xs[i] >> 16);
            acc[g] += sg;
          }
        }
        if (gl == 0 && mask != 0) bytes += 4ull * (unsigned long long)(e1 - e0);  // the profile
      }
      for (; e + 48 < e1; e += 64) {
        const uint32_t x0 = J.st.kp[e], x1 = J.st.kp[e + 16], x2 = J.st.kp[e + 32], x3 = J.st.kp[e + 48];
#pragma unroll
        for (int g = 0; g < G; ++g) {
          if (G < 16 && !((mask >> g) & 1u)) continue;
          const uint16_t* tg = table + g * kTS;
          acc[g] += min((uint32_t)tg[x0 & 0xFFFFu], x0 >> 16) + min((uint32_t)tg[x1 & 0xFFFFu], x1 >> 16) +
                    min((uint32_t)tg[x2 & 0xFFFFu], x2 >> 16) + min((uint32_t)tg[x3 & 0xFFFFu], x3 >> 16);
        }
      }
      for (; e < e1; e += 16) {
        const uint32_t x = J.st.kp[e];
#pragma unroll
        for (int g = 0; g < G; ++g) {
          if (G < 16 && !((mask >> g) & 1u)) continue;
          acc[g] += min((uint32_t)table[g * kTS + (x & 0xFFFFu)], x >> 16);
        }
      }
      // a cut pair fails: its partial sum is dropped (wc_q >= 1)
      if (cut) acc[0] = 0;
      // per representative g of the group: the 16-lane total, kept by lane g
      int mine = 0;
#pragma unroll
      for (int g = 0; g < G; ++g) {
        if (!((mask >> g) & 1u)) continue;
        int sg = acc[g];
#pragma unroll
        for (int off = 8; off > 0; off >>= 1) sg += __shfl_xor(sg, off, 16);
        if (gl == g) mine = sg;
      }
      // lane g decides representative g; survivors of the wave are appended
      // with one list reservation per tile shape present
      const bool live = gl < G && ((mask >> gl) & 1u);
      const bool pass = live && mine >= wcq;
      if (live && !pass) ++rej_word;
      int sh = 0;
      AlignItem it{0, 0, 0, 0};
      if (pass) {
        const int r = s_rep[gl], slot = s_slot[gl], n = s_n[gl];
        ++passed;
        bytes += 16;
        cband += (unsigned long long)band_cells(n, mq, c);
        cfull += (unsigned long long)n * (unsigned long long)mq;
        const bool sw = swap_roles(n, mq, a.swap_ratio);
        sh = sw ? pick_shape(mq, n, c, a.min_shape, a.shapes) : pick_shape(n, mq, c, a.min_shape, a.shapes);
        it.text = sw ? q : r;
        it.query = sw ? r : q;
        it.job = job;
        it.slot = slot | (sw ? kSwapFlag : 0);
        // long pairs: two waves each (the split list), starting from the
        // window of their shape (at least two words per lane)
        if (a.split_rows > 0 && (sw ? n : mq) >= a.split_rows) {
          it.slot |= sh << kSplitShapeShift;
          sh = kSplitList;
        }
      }
      const int sh_own = sh == kSplitList ? (it.slot >> kSplitShapeShift) & 15 : sh;
      const int wl = tid & 63;
      bool todo = pass;
      while (true) {
        const uint64_t pend = __builtin_amdgcn_ballot_w64(todo);
        if (pend == 0) break;
        const int l = (int)__builtin_ctzll(pend);
        const int s0 = __shfl(sh, l);
        const uint64_t same = __builtin_amdgcn_ballot_w64(todo && sh == s0);
        int base = 0;
        if (wl == l) base = atomicAdd(&a.bin_count[s0], (int)__popcll(same));
        base = __shfl(base, l);
        if (todo && sh == s0) {
          const int at = base + (int)__popcll(same & ((1ull << wl) - 1ull));
          if (at < (sh == kSplitList ? a.split_cap : a.item_cap)) {
            a.items[(int64_t)sh * a.item_cap + at] = it;
            todo = false;
          } else if (sh == kSplitList) {
            // the split list holds as many pairs as the split kernel has
            // blocks: the rest go to their own shape list (next round)
            atomicSub(&a.bin_count[sh], 1);
            it.slot &= ~(15 << kSplitShapeShift);
            sh = sh_own;
          } else {
            // the list is full: count it so the host fails the call (the
            // host bounds k_spec * nc so that this does not happen)
            atomicAdd(&a.stats[ST_OVERFLOW], 1ull);
            atomicSub(&a.bin_count[sh], 1);
            todo = false;
          }
        }
      }
    }
  }
    // the next batch rewrites s_act
    if (compact) __syncthreads();
  }
  unsigned long long v[7] = {rej_len, rej_comp, rej_word, passed, bytes, cband, cfull};
  const int lane = tid & 63;
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    unsigned long long x = v[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    v[i] = x;
  }
  if (lane == 0 && (v[0] | v[1] | v[2] | v[3])) {
    atomicAdd(&stat_row(a.stats)[ST_PAIRS], v[0] + v[1] + v[2] + v[3]);
    if (v[0]) atomicAdd(&stat_row(a.stats)[ST_REJ_LEN], v[0]);
    if (v[1]) atomicAdd(&stat_row(a.stats)[ST_REJ_COMP], v[1]);
    if (v[2]) atomicAdd(&stat_row(a.stats)[ST_REJ_WORD], v[2]);
    if (v[3]) atomicAdd(&stat_row(a.stats)[ST_ALIGNED], v[3]);
    atomicAdd(&stat_row(a.stats)[ST_FILTER_BYTES], v[4]);
    if (v[5]) atomicAdd(&stat_row(a.stats)[ST_CELLS_BAND], v[5]);
    if (v[6]) atomicAdd(&stat_row(a.stats)[ST_CELLS_FULL], v[6]);
  }
}

// ------------------------------------------------------- signature screen
// Phase B, before the filter, for jobs whose store has signatures (k >= 7):
// one block per chunk of 64 candidate positions (the filter's phase-B chunk
// for these word lengths), one candidate per lane, its record and signature
// in registers of each of the block's four waves. The step's true
// representatives (records and signatures) are staged in LDS once per job
// and split over the waves. A pair is rejected here, and counted, when it
// fails the filter's length or composition test, or when the signature bound
// exceeds the candidate's slack T_q - wc_q: its shared-word sum is then below
// wc_q (func.cpp:567) whatever the profiles hold. The slots that survive go
// to scr[q], their OR over the chunk to chm[chunk]. The count over any part
// of the signature is a bound too: it runs a quarter of the words at a time
// and stops once every lane has decided (for unrelated pairs the first
// quarter already exceeds the slack).
static_assert(filter_chunk(filter_group(kScreenMinK)) == 64 && filter_chunk(filter_group(8)) == 64,
              "one block covers one phase-B filter chunk of the screened word lengths");

__global__ __launch_bounds__(256) void screen_kernel(ScreenArgs a) {
  __shared__ int s_pre[kMaxJobsPerLaunch + 1];
  __shared__ int s_part[4];
  __shared__ uint4 s_sig[kMaxSpec][kSigWords / 4];  // representatives by rank
  __shared__ int s_slot[kMaxSpec], s_m[kMaxSpec];
  __shared__ uint32_t s_cac[kMaxSpec], s_cgt[kMaxSpec];
  __shared__ unsigned long long s_sv[4][64], s_cm[4];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int nj = a.njobs;
  {
    int carry = 0;
    for (int base = 0; base < nj; base += 256) {
      const int j = base + tid;
      int u = 0;
      if (j < nj) {
        const JobDev& J = a.jobs[j];
        const JobState* S = J.state;
        if (J.scr != nullptr && !(J.flags & kJobDropFilters) && !S->done) {
          const int rest = S->nc - S->k_spec;
          u = rest > 0 ? (rest + 63) >> 6 : 0;
        }
      }
      int tot;
      const int ex = block_exclusive_scan<256>(u, &tot, s_part);
      if (j < nj) s_pre[j] = carry + ex;
      carry += tot;
    }
    if (tid == 0) s_pre[nj] = carry;
    __syncthreads();
  }
  const int total = s_pre[nj];
  unsigned long long rej_len = 0, rej_comp = 0, rej_sig = 0, cands = 0;
  int staged = -1, kt = 0;
  for (int u = blockIdx.x; u < total; u += gridDim.x) {
    int lo = 0, hi = nj - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_pre[mid] <= u) lo = mid; else hi = mid - 1;
    }
    // the same job for the whole block: scalar loads from here on
    const int job = __builtin_amdgcn_readfirstlane(lo);
    const JobDev& J = a.jobs[job];
    const JobState* S = J.state;
    const int K = S->k_spec, nc = S->nc;
    const int32_t* cand = J.cand[S->cur];
    const unsigned long long tm = S->tmask;
    if (job != staged) {
      // stage the true representatives of the step by rank: slot, length,
      // composition and signature
      __syncthreads();
      kt = __popcll(tm);
      if (tid < 64 && ((tm >> tid) & 1ull)) {
        const int rank = __popcll(tm & ((1ull << tid) - 1ull));
        const SeqRec rr = J.st.rec[cand[tid]];
        s_slot[rank] = tid;
        s_m[rank] = rr.m;
        s_cac[rank] = rr.comp_ac;
        s_cgt[rank] = rr.comp_gt;
      }
      __syncthreads();
      // four threads per representative, 64 bytes each
      const int rank = tid >> 2, part = tid & 3;
      if (rank < kt) {
        const uint4* src = reinterpret_cast<const uint4*>(J.st.sig + (int64_t)cand[s_slot[rank]] * kSigWords);
#pragma unroll
        for (int i = 0; i < 4; ++i) s_sig[rank][part * 4 + i] = src[part * 4 + i];
      }
      __syncthreads();
      staged = job;
    }
    const int chunk = u - s_pre[job];
    const int pos = K + chunk * 64 + lane;
    const bool inpos = pos < nc;
    const int q = inpos ? cand[pos] : cand[K];
    const uint4* rp = reinterpret_cast<const uint4*>(J.st.rec + q);
    const uint4 ra = rp[0], rb = rp[1];
    const int mq = (int)ra.x, bcq = (int)ra.y, c = bcq + 1;
    const int slack = (int)ra.w - (int)ra.z;  // T_q - wc_q
    const uint32_t qa = rb.x & 0xFFFFu, qc = rb.x >> 16, qg = rb.y & 0xFFFFu, qt = rb.y >> 16;
    uint32_t sq[kSigWords];
    const uint4* sp = reinterpret_cast<const uint4*>(J.st.sig + (int64_t)q * kSigWords);
#pragma unroll
    for (int w = 0; w < kSigWords / 4; ++w) {
      const uint4 v = sp[w];
      sq[4 * w] = v.x;
      sq[4 * w + 1] = v.y;
      sq[4 * w + 2] = v.z;
      sq[4 * w + 3] = v.w;
    }
    if (inpos && wid == 0) ++cands;
    unsigned long long sv = 0, cm = 0;
    for (int ri = wid; ri < kt; ri += 4) {
      const int slot = __builtin_amdgcn_readfirstlane(s_slot[ri]);
      const int rm = __builtin_amdgcn_readfirstlane(s_m[ri]);
      const uint32_t cac = __builtin_amdgcn_readfirstlane(s_cac[ri]);
      const uint32_t cgt = __builtin_amdgcn_readfirstlane(s_cgt[ri]);
      bool alive = inpos;
      if (alive && c > min(rm, mq)) {
        ++rej_len;
        alive = false;
      }
      if (alive) {
        const uint32_t sum = min(cac & 0xFFFFu, qa) + min(cac >> 16, qc) + min(cgt & 0xFFFFu, qg) + min(cgt >> 16, qt);
        if ((int)sum <= bcq) {
          ++rej_comp;
          alive = false;
        }
      }
      if (alive) {
        // buckets of q missing from r: a lower bound of the deficit
        int miss = 0;
#pragma unroll
        for (int part = 0; part < 4; ++part) {
          if (part > 0 && __builtin_amdgcn_ballot_w64(miss <= slack) == 0) break;
#pragma unroll
          for (int w = part * kSigWords / 16; w < (part + 1) * kSigWords / 16; ++w) {
            const uint4 r = s_sig[ri][w];
            miss += __builtin_popcount(sq[4 * w] & ~r.x) + __builtin_popcount(sq[4 * w + 1] & ~r.y) +
                    __builtin_popcount(sq[4 * w + 2] & ~r.z) + __builtin_popcount(sq[4 * w + 3] & ~r.w);
          }
        }
        if (miss > slack) {
          ++rej_sig;
          alive = false;
        }
      }
      if (alive) sv |= 1ull << slot;
      if (__builtin_amdgcn_ballot_w64(alive)) cm |= 1ull << slot;
    }
    // the waves' survivors of each candidate and of the chunk
    s_sv[wid][lane] = sv;
    if (lane == 0) s_cm[wid] = cm;
    __syncthreads();
    if (wid == 0) {
      if (inpos) J.scr[q] = s_sv[0][lane] | s_sv[1][lane] | s_sv[2][lane] | s_sv[3][lane];
      if (lane == 0) J.chm[chunk] = s_cm[0] | s_cm[1] | s_cm[2] | s_cm[3];
    }
    __syncthreads();
  }
  unsigned long long v[4] = {rej_len, rej_comp, rej_sig, cands};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    unsigned long long x = v[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    v[i] = x;
  }
  if (lane == 0 && (v[0] | v[1] | v[2] | v[3])) {
    atomicAdd(&stat_row(a.stats)[ST_PAIRS], v[0] + v[1] + v[2]);
    if (v[0]) atomicAdd(&stat_row(a.stats)[ST_REJ_LEN], v[0]);
    if (v[1]) atomicAdd(&stat_row(a.stats)[ST_REJ_COMP], v[1]);
    if (v[2]) {
      atomicAdd(&stat_row(a.stats)[ST_REJ_WORD], v[2]);
      atomicAdd(&stat_row(a.stats)[ST_SCREENED], v[2]);
    }
    // id, record and signature of each candidate, its survivor mask
    if (v[3]) atomicAdd(&stat_row(a.stats)[ST_FILTER_BYTES], v[3] * (4 + 32 + 4 * kSigWords + 8));
  }
}

// Signatures of every sequence: one wave per sequence, the 2048 bits built
// in LDS from the distinct k-mers with a nonzero stored count.
__global__ __launch_bounds__(256) void signature_kernel(const uint32_t* __restrict__ kp,
                                                        const int64_t* __restrict__ kpoff, int n, int k,
                                                        uint32_t* __restrict__ sig) {
  __shared__ uint32_t bits[4][kSigWords];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int s = blockIdx.x * 4 + wv;
  bits[wv][lane] = 0u;
  __syncthreads();
  if (s < n) {
    for (int64_t e = kpoff[s] + lane; e < kpoff[s + 1]; e += 64) {
      const uint32_t x = kp[e];
      if (x >> 16) {
        const uint32_t b = sig_bucket(x & 0xFFFFu, k);
        atomicOr(&bits[wv][b >> 5], 1u << (b & 31));
      }
    }
  }
  __syncthreads();
  if (s < n) sig[(int64_t)s * kSigWords + lane] = bits[wv][lane];
}

// ------------------------------------------------------- pigeonhole test
// Exact prefilter between the filter and the alignment (no counterpart in
// the reference). For a pair with gap-free lengths n, m and cutoff c = bc+1,
// a common subsequence of length >= c leaves at most E = n + m - 2c bases of
// the two sequences unmatched. Cut the shorter sequence x into E+1 disjoint
// segments: an unmatched base of x lies in one segment, and an unmatched base
// of the longer y between two matched bases of x touches only the segment
// holding both, so some segment has no unmatched base at all and occurs
// verbatim in y. Its matches lie on one diagonal d = pos_y - pos_x, and every
// match of such a path has d in [-(|x| - c), |y| - c]. So when no segment
// occurs in y at an in-band diagonal, LCS < c: the pair fails the alignment
// test (func.cpp:671) and is flagged here instead of aligned. Checked only
// where it is cheap: segments of at least 8 bases and (E+1)^2 <= 4|x|
// segment-diagonal positions.
__device__ __forceinline__ uint2 window32(const uint2* pk, int64_t off, int pos, int len) {
  // bases pos .. pos+len-1 (len <= 32, inside the sequence) of both planes
  const int w = pos >> 5, s = pos & 31;
  const uint2 a = pk[off + w];
  uint2 b = make_uint2(0u, 0u);
  if (s + len > 32) b = pk[off + w + 1];
  const uint32_t lo = s ? (a.x >> s) | (b.x << (32 - s)) : a.x;
  const uint32_t hi = s ? (a.y >> s) | (b.y << (32 - s)) : a.y;
  return make_uint2(lo, hi);
}

__global__ __launch_bounds__(256) void pigeon_kernel(PigeonArgs a) {
  const int lane = threadIdx.x & 63;
  const int gw = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int nw = gridDim.x * 4;
  unsigned long long rej = 0, checked = 0;
  for (int sh = 0; sh <= a.smax; ++sh) {
    const int cnt = a.bin_count[sh];
    AlignItem* list = a.items + (int64_t)sh * a.item_cap;
    for (int i = gw; i < cnt; i += nw) {
      const AlignItem it = list[i];
      const int text = __builtin_amdgcn_readfirstlane(it.text), query = __builtin_amdgcn_readfirstlane(it.query);
      const int slot = __builtin_amdgcn_readfirstlane(it.slot);
      const JobDev& J = a.jobs[__builtin_amdgcn_readfirstlane(it.job)];
      const StoreView& st = J.st;
      const int cand = (slot & kSwapFlag) ? text : query;
      const int n = st.m[text], mq = st.m[query], c = st.bc[cand] + 1;
      if (c > min(n, mq)) continue;  // the alignment rejects it without work
      const int x = n <= mq ? text : query, y = n <= mq ? query : text;
      const int mx = min(n, mq), ny = max(n, mq);
      const int segs = n + mq - 2 * c + 1;  // E + 1
      const int len = min(32, mx / segs);
      if (len < 8 || (long long)segs * segs > 4ll * mx) continue;
      ++checked;
      const int64_t xo = st.pkoff[x], yo = st.pkoff[y];
      const int dlo = -(mx - c);
      const uint32_t mask = len == 32 ? ~0u : (1u << len) - 1u;
      bool found = false;
      for (int j = 0; j < segs && !found; ++j) {
        const int ax = j * len;
        const uint2 xs = window32(st.pk, xo, ax, len);
        for (int d0 = 0; d0 < segs && !found; d0 += 64) {
          const int d = dlo + d0 + lane;
          const int by = ax + d;
          bool hit = false;
          if (d0 + lane < segs && by >= 0 && by + len <= ny) {
            const uint2 ys = window32(st.pk, yo, by, len);
            hit = (((xs.x ^ ys.x) | (xs.y ^ ys.y)) & mask) == 0u;
          }
          found = __builtin_amdgcn_ballot_w64(hit) != 0;
        }
      }
      if (!found) {
        if (lane == 0) list[i].slot = slot | kPigeonFlag;
        ++rej;
      }
    }
  }
  if (lane == 0 && (checked | rej)) {
    atomicAdd(&stat_row(a.stats)[ST_PIGEON_CHECKED], checked);
    if (rej) atomicAdd(&stat_row(a.stats)[ST_REJ_PIGEON], rej);
  }
}

// ------------------------------------------------------------- align body
template <int S>
__host__ __device__ constexpr uint64_t top_lanes() {
  uint64_t k = 0;
  for (int l = S - 1; l < 64; l += S) k |= 1ull << l;
  return k;
}

// Bits of text word gw that are real bases (text of n gap-free bases).
__device__ inline uint32_t valid_bits(int gw, int n) {
  const int full = n >> 5;
  if (gw < full) return ~0u;
  if (gw == full) return (1u << (n & 31)) - 1u;
  return 0u;
}

// Truth tables of v_bitop3_b32: bit i of the immediate is the result for the
// input bits (a, b, c) = (bit 2, bit 1, bit 0 of i), i.e. a = 0xF0, b = 0xCC,
// c = 0xAA.
constexpr int kBitAndXnor = 0xF0 & ~(0xCC ^ 0xAA) & 0xFF;  // a & ~(b ^ c)
constexpr int kBitOrXor = (0xF0 | (0xCC ^ 0xAA)) & 0xFF;    // a | (b ^ c)

// S = A + B on every lane, with the lanes whose add carried out returned as a
// lane mask. One v_add_co_u32, whose carry out is that mask (the compiler
// turns a ballot of the carry into two more vector operations).
__device__ __forceinline__ uint32_t add_carry_out(uint32_t a, uint32_t b, uint64_t& carries) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t s;
  asm("v_add_co_u32_e64 %0, %1, %2, %3" : "=v"(s), "=s"(carries) : "v"(a), "v"(b));
  return s;
#else
  const uint32_t s = a + b;
  carries = __builtin_amdgcn_ballot_w64(s < a);
  return s;
#endif
}

// A + 1 on the lanes set in the wave-uniform mask cin, A elsewhere. One
// v_addc_co_u32 with cin as its carry in.
__device__ __forceinline__ uint32_t add_carry_in(uint32_t a, uint64_t cin) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t s;
  uint64_t unused;
  asm("v_addc_co_u32_e64 %0, %1, %2, 0, %3" : "=v"(s), "=s"(unused) : "v"(a), "s"(cin));
  return s;
#else
  return a + (uint32_t)((cin >> (threadIdx.x & 63)) & 1u);
#endif
}

// First half of a row for a wave-uniform query base B (2-bit code): match
// mask, U = V & M and the per-lane multi-word sum S = V + U with its carry
// out. The code is fixed at compile time, so M & V is one three-input bit
// operation on (V, lo plane, hi plane).
//
// With KEEPU, U stays in u[] for the second half, which then forms
// V' = S | (V ^ U) in one three-input operation per word. Without it, V is
// replaced by V & ~M (= V ^ U) here, one more operation per word but W
// fewer registers: the widest tiles take that to keep their occupancy.
template <int B, int W, bool KEEPU>
__device__ __forceinline__ void row_sum_base(uint32_t (&v)[W], const uint32_t (&tl)[W],
                                             const uint32_t (&th)[W], uint32_t (&s)[W], uint32_t (&u)[W],
                                             uint32_t& carry) {
  carry = 0u;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const uint32_t ml = (B & 1) ? tl[w] : ~tl[w];
    const uint32_t mh = (B & 2) ? th[w] : ~th[w];
    const uint32_t uw = v[w] & ml & mh;
    uint32_t co;
    s[w] = __builtin_addc(v[w], uw, carry, &co);
    carry = co;
    if constexpr (KEEPU) u[w] = uw;
    else v[w] ^= uw;
  }
  // the carry leaves the branch as a 0/1 register value: merging a lane-mask
  // value over the four branches would cost scalar mask arithmetic per row
  asm volatile("" : "+v"(carry));
}

// Second half of a row for one 64-lane segment: carry look-ahead across the
// lanes, then V' = (S + carry in) | (V ^ U). A lane passes a carry in on only
// when all of its words of S are ones, which needs its first word to be all
// ones. When no lane has that, the carry into lane l is simply the carry out
// of lane l-1 (one lane shift of the wave, DPP wave_shr:1), and it cannot
// ripple past the first word; otherwise the full generate/propagate
// look-ahead runs, and the ripple through a lane's other words runs only
// when some lane needs it.
template <int W, bool KEEPU>
__device__ __forceinline__ void row_finish64(uint32_t (&v)[W], uint32_t (&s)[W], const uint32_t (&u)[W],
                                             uint32_t carry, int lane) {
  constexpr uint64_t KTOP = 1ull << 63;
  if (__builtin_amdgcn_ballot_w64(s[0] == ~0u) == 0) {
    // no first word is all ones: the carry in cannot ripple past it
    const uint32_t cin =
        (uint32_t)__builtin_amdgcn_update_dpp(0, (int)carry, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
    s[0] += cin;
  } else {
    const uint64_t G = __builtin_amdgcn_ballot_w64(carry != 0u) & ~KTOP;
    uint32_t pall = ~0u;
#pragma unroll
    for (int w = 0; w < W; ++w) pall &= s[w];
    const uint64_t P = __builtin_amdgcn_ballot_w64(pall == ~0u) & ~KTOP;
    const uint64_t X = G | P;
    const uint64_t C = (X + G) ^ X ^ G;
    const uint32_t cin = (uint32_t)(C >> lane) & 1u;
    uint32_t co;
    s[0] = __builtin_addc(s[0], 0u, cin, &co);
    if (W > 1 && __builtin_amdgcn_ballot_w64(co != 0u) != 0) {
#pragma unroll
      for (int w = 1; w < W; ++w) {
        uint32_t c2;
        s[w] = __builtin_addc(s[w], 0u, co, &c2);
        co = c2;
      }
    }
  }
#pragma unroll
  for (int w = 0; w < W; ++w) v[w] = KEEPU ? s[w] | (v[w] ^ u[w]) : s[w] | v[w];  // else v holds V & ~M
}

// S = V + U over a lane's W words (word 0 lowest) as one v_add_co / v_addc
// chain, returning the carry out of the top word as a lane mask. A carry
// written by one vector operation may be read as the next one's carry in only
// two wait states later, hence the s_nop between the links.
template <int W>
__device__ __forceinline__ uint64_t add_words(const uint32_t (&v)[W], const uint32_t (&u)[W], uint32_t (&s)[W]) {
  static_assert(W >= 2 && W <= 4, "chains of 2 to 4 words");
  uint64_t g;
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (W == 2) {
    asm("v_add_co_u32_e64 %0, %2, %3, %5\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e64 %1, %2, %4, %6, %2"
        : "=&v"(s[0]), "=&v"(s[1]), "=&s"(g)
        : "v"(v[0]), "v"(v[1]), "v"(u[0]), "v"(u[1]));
  } else if constexpr (W == 3) {
    asm("v_add_co_u32_e64 %0, %3, %4, %7\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e64 %1, %3, %5, %8, %3\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e64 %2, %3, %6, %9, %3"
        : "=&v"(s[0]), "=&v"(s[1]), "=&v"(s[2]), "=&s"(g)
        : "v"(v[0]), "v"(v[1]), "v"(v[2]), "v"(u[0]), "v"(u[1]), "v"(u[2]));
  } else {
    asm("v_add_co_u32_e64 %0, %4, %5, %9\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e64 %1, %4, %6, %10, %4\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e64 %2, %4, %7, %11, %4\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e64 %3, %4, %8, %12, %4"
        : "=&v"(s[0]), "=&v"(s[1]), "=&v"(s[2]), "=&v"(s[3]), "=&s"(g)
        : "v"(v[0]), "v"(v[1]), "v"(v[2]), "v"(v[3]), "v"(u[0]), "v"(u[1]), "v"(u[2]), "v"(u[3]));
  }
#else
  uint32_t c = 0u;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    uint32_t co;
    s[w] = __builtin_addc(v[w], u[w], c, &co);
    c = co;
  }
  g = __builtin_amdgcn_ballot_w64(c != 0u);
#endif
  return g;
}

// A whole row of a 64-lane tile with 2 to 4 words per lane and the query base
// as two wave-uniform masks (ql, qh all ones where the base bit is set).
// U = V & M from two v_bitop3 per word; S = V + U by add_words, whose carry
// out is already a lane mask G; then the carries between lanes on the scalar
// unit. A lane passes a carry on only when all of its words of S are ones,
// which needs its first word to be all ones: when no lane has that, the carry
// into lane l is the carry out of lane l-1 (G << 1) and stops in the first
// word; otherwise the full look-ahead ((G|P) + G) ^ (G|P) ^ G runs, with the
// ripple through the lane's words. V' = (S + carry in) | (V ^ U).
template <int W>
__device__ __forceinline__ void row_masks64(uint32_t (&v)[W], const uint32_t (&tl)[W], const uint32_t (&th)[W],
                                            uint32_t ql, uint32_t qh, int lane) {
  uint32_t u[W], s[W];
#pragma unroll
  for (int w = 0; w < W; ++w)
    u[w] = __builtin_amdgcn_bitop3_b32(__builtin_amdgcn_bitop3_b32(v[w], tl[w], ql, kBitAndXnor), th[w], qh,
                                       kBitAndXnor);
  const uint64_t g = add_words<W>(v, u, s);
  if (__builtin_amdgcn_ballot_w64(s[0] == ~0u) == 0) {
    s[0] = add_carry_in(s[0], g << 1);
  } else {
    uint32_t pall = ~0u;
#pragma unroll
    for (int w = 0; w < W; ++w) pall &= s[w];
    const uint64_t p = __builtin_amdgcn_ballot_w64(pall == ~0u);
    const uint64_t x = g | p;
    const uint64_t cm = (x + g) ^ x ^ g;
    uint32_t co = (uint32_t)(cm >> lane) & 1u;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      uint32_t c2;
      s[w] = __builtin_addc(s[w], 0u, co, &c2);
      co = c2;
    }
  }
#pragma unroll
  for (int w = 0; w < W; ++w) v[w] = __builtin_amdgcn_bitop3_b32(s[w], v[w], u[w], kBitOrXor);
}

// Banded bit-parallel LCS decision, LCS(text, query) > bc_query.
//
// Recurrence (func.cpp:596-626): V starts all ones over the text; per query
// base with match mask M, V' = (V + (V & M)) | (V & ~M); LCS = zeros of V.
// The run is exact when matches are restricted to text words of a window
// that always contains the diagonal band j-(m-c) <= i <= j+(n-c), c = bc+1,
// because every common subsequence of length >= c lies in that band: words
// that left the window keep their bits (no match, carry in 0) and words not
// yet in it stay all ones (the carry out of the window only passes through
// them), so they need no update. Each S-lane segment of a wave owns one pair
// with a window of S*W words, lane l holding words b + l*W .. b + l*W + W-1.
// The multi-word add carries across lanes by carry look-ahead on ballots
// (generate G, propagate P: carries = ((G|P) + G) ^ (G|P) ^ G) with the top
// lane of every segment cut off. Rows run in 32-row blocks aligned for all
// segments: rows past the query end use an all-mismatch mask, a no-op. At
// each block end the zero count z gives the exact early decisions z >= c
// (accept) and z + rows_left < c (reject); the window slides by one lane when
// the band's left edge has passed the lowest lane.
// True when the list is already taken: the waves still arriving leave
// without an atomic on the shared counter (checked once, before a wave's
// first reservation).
__device__ __forceinline__ bool work_taken(const int* work, int total) {
  return __hip_atomic_load(work, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= total;
}

struct RaggedRows { static constexpr bool value = true; };
struct FullRows { static constexpr bool value = false; };

template <int S, int W>
__device__ inline void align_loop(const AlignArgs& a) {
  static_assert(W == 1, "the row code below holds one text word per lane");
  // carries stop at every segment's top lane (the wave's own top lane drops
  // its carry out anyway)
  constexpr uint64_t kCut = S == 64 ? ~0ull : ~top_lanes<S>();
  const int lane = threadIdx.x & 63;
  const int sl = lane & (S - 1);
  const int seg0 = lane & ~(S - 1);

  uint32_t v[W], tl[W], th[W];
#pragma unroll
  for (int w = 0; w < W; ++w) { v[w] = ~0u; tl[w] = 0u; th[w] = 0u; }

  int item = -1;
  bool exhausted = false;
  // tx: bit-vector sequence, qx: row sequence, q: the candidate (owner of
  // the cutoff and of the accept mark)
  int tx = 0, qx = 0, q = 0, job = 0, slot = 0, n = 0, mq = 0, c = 0, b = 0, jb = 0, frozen = 0, ntw = 0;
  const uint2* pk = nullptr;
  int64_t toff = 0, qoff = 0;
  unsigned long long cells = 0, nacc = 0, nrej = 0, cdec = 0, nskip = 0;
  const int total = *a.count;
  if (total == 0 || work_taken(a.work, total)) return;

  while (true) {
    // ---- fetch a pair for every segment without one; pairs decided without
    // alignment (LCS <= min(n, m) < c, or q already taken by an earlier true
    // representative of this step) are settled here and the fetch repeats
    while (true) {
      const bool need = (item < 0) && !exhausted;
      // one work-counter reservation for every segment of the wave that
      // needs a pair; segment i of them takes the i-th index
      const uint64_t needm = __builtin_amdgcn_ballot_w64(need && sl == 0);
      if (needm == 0) break;
      int base = 0;
      if (lane == 0) base = atomicAdd(a.work, (int)__popcll(needm));
      base = __shfl(base, 0);
      const int got = base + (int)__popcll(needm & ((1ull << seg0) - 1ull));
      if (!need) continue;
      if (got >= total) {
        exhausted = true;
        continue;
      }
      const AlignItem it = a.items[got];
      if (it.slot & kPigeonFlag) continue;  // rejected (and counted) by the pigeonhole test
      const JobDev& J = a.jobs[it.job];
      const StoreView& st = J.st;
      const int islot = it.slot & (kSwapFlag - 1);
      const int cand = (it.slot & kSwapFlag) ? it.text : it.query;  // the candidate owns the cutoff
      const int n2 = st.m[it.text], m2 = st.m[it.query], c2 = st.bc[cand] + 1;
      if (a.skip_absorbed && (J.acc[cand] & J.state->tmask & ((1ull << islot) - 1ull))) {
        if (sl == 0) ++nskip;
        continue;
      }
      if (c2 > min(n2, m2)) {
        if (sl == 0) {
          if (a.result) a.result[got] = 0;
          ++nrej;
        }
        continue;
      }
      item = got;
      tx = it.text;
      q = cand;
      qx = it.query;
      job = it.job;
      slot = islot;
      pk = st.pk;
      n = n2;
      mq = m2;
      c = c2;
      toff = st.pkoff[tx];
      qoff = st.pkoff[qx];
      ntw = (n >> 5) + 1;
      b = 0;
      jb = 0;
      frozen = 0;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const int gw = sl * W + w;
        uint2 t = make_uint2(0u, 0u);
        if (gw < ntw) t = pk[toff + gw];
        tl[w] = t.x;
        th[w] = t.y;
        v[w] = ~0u;
      }
    }
    if (__builtin_amdgcn_ballot_w64(!exhausted) == 0) break;
    const bool active = item >= 0;
    const bool hopeless = false;

    // ---- slide the window by one lane when the band has left the lowest lane
    const int lo0 = jb * 32 - (mq - c);
    const bool slide = active && !hopeless && lo0 >= 0 && (lo0 >> 5) >= b + W && b + S * W < ntw;
    if (__builtin_amdgcn_ballot_w64(slide) != 0) {
      int dz = 0;
#pragma unroll
      for (int w = 0; w < W; ++w) dz += __popc(~v[w] & valid_bits(b + sl * W + w, n));
      dz = __shfl(dz, seg0);
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const uint32_t nv = __shfl(v[w], lane + 1);
        const uint32_t nl = __shfl(tl[w], lane + 1);
        const uint32_t nh = __shfl(th[w], lane + 1);
        if (slide) { v[w] = nv; tl[w] = nl; th[w] = nh; }
      }
      if (slide) {
        frozen += dz;
        if (sl == S - 1) {
#pragma unroll
          for (int w = 0; w < W; ++w) {
            const int gw = b + S * W + w;
            uint2 t = make_uint2(0u, 0u);
            if (gw < ntw) t = pk[toff + gw];
            tl[w] = t.x;
            th[w] = t.y;
            v[w] = ~0u;
          }
        }
        b += W;
      }
    }

    // ---- query word for this 32-row block
    uint32_t qlo = 0u, qhi = 0u, qnv = ~0u;
    if (active && !hopeless) {
      const uint2 qw = pk[qoff + jb];
      qlo = qw.x;
      qhi = qw.y;
      const int rows = mq - jb * 32;
      qnv = rows >= 32 ? 0u : ~((1u << rows) - 1u);
    }
    if constexpr (S == 64) {
      // one pair per wave: the query bits are wave-uniform, so the per-row
      // masks are made on the scalar unit
      qlo = __builtin_amdgcn_readfirstlane(qlo);
      qhi = __builtin_amdgcn_readfirstlane(qhi);
      qnv = __builtin_amdgcn_readfirstlane(qnv);
    }

    // ---- 32 rows, unrolled in chunks of UR rows with constant bit offsets
    constexpr int UR = 8;
    // Whether any segment ends its query inside this block: only such a
    // block needs the query-end mask on its rows.
    const bool ragged = __builtin_amdgcn_ballot_w64(active && qnv != 0u) != 0;
    auto run_rows = [&](auto tag) {
      constexpr bool kRagged = decltype(tag)::value;
      for (int r0 = 0; r0 < 32; r0 += UR) {
        const uint32_t qlc = qlo >> r0, qhc = qhi >> r0, qnc = qnv >> r0;
#pragma unroll
        for (int r = 0; r < UR; ++r) {
          // bit r as an all-zeros / all-ones mask (one signed bit-field extract)
          const uint32_t ql = (uint32_t)((int32_t)(qlc << (31 - r)) >> 31);
          const uint32_t qh = (uint32_t)((int32_t)(qhc << (31 - r)) >> 31);
          // U = V & M, M = [text base == query base] = ~(tl ^ ql) & ~(th ^ qh):
          // two three-input bit operations; rows past a query end match nothing
          uint32_t u = __builtin_amdgcn_bitop3_b32(
              __builtin_amdgcn_bitop3_b32(v[0], tl[0], ql, kBitAndXnor), th[0], qh, kBitAndXnor);
          if constexpr (kRagged) u &= ~(uint32_t)((int32_t)(qnc << (31 - r)) >> 31);
          // S = V + U per lane (carry-out lanes G), then the carries between
          // lanes by look-ahead on the scalar unit: a lane passes a carry on
          // only where S is all ones (P), which never holds where it
          // generates one (V + U < 2^33 - 1). With X = G | P, bit l of X + G
          // is then P_l ^ (carry into lane l), so the carries into the lanes
          // are (X + G) ^ P, with every segment's top lane cut off in G and P
          // (four scalar operations; the 64-bit AND of the cut is free for a
          // 64-lane segment). Rows past a query end and segments without a
          // pair have U = 0 and generate nothing.
          uint64_t g;
          const uint32_t s = add_carry_out(v[0], u, g);
          const uint64_t p = __builtin_amdgcn_ballot_w64(s == ~0u);
          const uint64_t gm = g & kCut, pm = p & kCut;
          const uint64_t cin = ((gm | pm) + gm) ^ pm;
          // V' = (S + carry in) | (V ^ U)
          v[0] = __builtin_amdgcn_bitop3_b32(add_carry_in(s, cin), v[0], u, kBitOrXor);
        }
      }
    };
    if (ragged) {
      run_rows(RaggedRows{});
    } else {
      run_rows(FullRows{});
    }

    // ---- block end: exact early decisions on the zero counts of all
    // columns (z) and of the columns < n-R (R rows left): accept z >= c,
    // reject D[j][n-R] + R < c (columns left of the window all counted, an
    // upper bound), never weaker than z + R < c
    const int jl = min(jb * 32 + 31, mq - 1);
    const int R = mq - 1 - jl;
    const int istar = max(0, n - R);
    uint32_t zx = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const int g = b + sl * W + w;
      zx += (uint32_t)__popc(~v[w] & valid_bits(g, n)) | ((uint32_t)__popc(~v[w] & valid_bits(g, istar)) << 16);
    }
#pragma unroll
    for (int off = S / 2; off > 0; off >>= 1) zx += __shfl_xor(zx, off);
    if (active) {
      const int z = frozen + (int)(zx & 0xffffu);
      const int dstar = frozen + (int)(zx >> 16);
      bool accept = false, reject = false;
      if (hopeless) {
        reject = true;
      } else if (z >= c) {
        accept = true;
      } else if (dstar + R < c) {
        reject = true;
      }
      if (sl == 0) cells += (unsigned long long)S * W * 32ull * 32ull;
      if (accept || reject) {
        if (sl == 0) {
          if (!hopeless) cdec += (unsigned long long)band_cells_rows(n, mq, c, jb * 32 + 32);
          if (a.result) {
            a.result[item] = accept ? 1 : 0;
          } else if (accept) {
            atomicOr(&a.jobs[job].acc[q], 1ull << slot);
          }
          if (accept) ++nacc; else ++nrej;
        }
        item = -1;
      } else {
        ++jb;
      }
    }
  }
  // wave totals
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    cells += __shfl_xor(cells, off);
    nacc += __shfl_xor(nacc, off);
    nrej += __shfl_xor(nrej, off);
    cdec += __shfl_xor(cdec, off);
    nskip += __shfl_xor(nskip, off);
  }
  if (lane == 0 && a.stats) {
    unsigned long long* sr = stat_row(a.stats);
    if (cells) atomicAdd(&sr[ST_CELLS_EXEC], cells);
    if (cdec) atomicAdd(&sr[ST_CELLS_DECIDED], cdec);
    if (nskip) atomicAdd(&sr[ST_SPEC_WASTE], nskip);
    if (nacc) atomicAdd(&sr[ST_ACCEPT], nacc);
    if (nrej) atomicAdd(&sr[ST_REJ_ALIGN], nrej);
  }
}

// ------------------------------------------------ 64-lane tiles, W >= 2
//
// One pair per wave, so every row is wave-uniform. On top of the banded
// recurrence above, the window follows the band's left edge as the DP values
// sharpen it, and its width follows the band:
//
// * Left edge from the DP values. After row j (R rows left) a path that
//   leaves row j at text column i ends with at most D[j][i] + min(R, n-i)
//   matches, D[j][i] being the zero count of columns < i. D is
//   non-decreasing in i, so the columns whose D + R < c form a prefix. A path
//   of at least c matches keeps its prefix inside the window (induction over
//   the blocks), so its value at its row-j column is at least its prefix
//   length and that column is not in the prefix: the prefix columns can be
//   frozen like the columns left of the diagonal band. The same bound gives
//   the reject test D[j][n-R] + R < c (the maximum over i sits at i = n-R),
//   never weaker than z + R < c.
// * Width. At each 32-row block end the window is rebased at the first live
//   word and given the narrowest width of the list (2, 3, 4, 6, 8, 10, 12, 16
//   words per lane, at most the tile's) that reaches the band's right edge at
//   the last row of the next block; the content moves through LDS. The tile's
//   own width always suffices (it covers the whole diagonal band). Words
//   entering on the right start all ones; words dropped on the right lie
//   right of the band's right edge, so on no band path yet, and re-enter all
//   ones, a state no larger than the true one: every value stays a lower
//   bound of the LCS while every band path stays inside the window, so the
//   decisions stay exact.

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ int64_t uniform64(int64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)x >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

struct Pair64 {
  const uint2* pk;
  int64_t toff, qoff;  // first packed word of the bit-vector (text) and row (query) sequences
  int n, mq, c, ntw;   // text length, row count, cutoff + 1, text words
  int every;           // blocks between DP-edge window rebases, minus one (a power of two; -1 = by width)
  // A half of a split pair (align_split_kernel) runs only its first `rows`
  // rows (the bounds still count all mq rows), on the reversed pair when
  // `rev` is set, and stops early when the other half has decided (*stop).
  int rows;
  int rev;
  volatile int* stop;
};

// Text word g of a sequence read backwards (bases len-1, len-2, ... from
// bit 0 of word 0), both planes: a 32-base window of the forward packing,
// bit-reversed; the reversed sequence's padding bits are zero like the
// forward one's.
__device__ __forceinline__ uint2 rev_word(const uint2* pk, int64_t off, int len, int g) {
  const int p = len - 32 * (g + 1);
  if (p <= -32) return make_uint2(0u, 0u);
  const uint2 w = window32(pk, off, p > 0 ? p : 0, p > 0 ? 32 : 32 + p);
  const int sh = p > 0 ? 0 : -p;
  return make_uint2(__builtin_bitreverse32(w.x) >> sh, __builtin_bitreverse32(w.y) >> sh);
}

__device__ __forceinline__ uint2 text_word64(const Pair64& P, int g) {
  return P.rev ? rev_word(P.pk, P.toff, P.n, g) : P.pk[P.toff + g];
}
__device__ __forceinline__ uint2 query_word64(const Pair64& P, int jb) {
  return P.rev ? rev_word(P.pk, P.qoff, P.mq, jb) : P.pk[P.qoff + jb];
}

// Per-wave LDS staging used to re-lay a window.
struct Stage64 {
  uint32_t* v;
  uint32_t* tl;
  uint32_t* th;
};

// Narrowest width of the list (at most wmax) whose window based at word
// `base` reaches word `last`.
__device__ inline int width_to(int base, int last, int wmax) {
  const int ws[8] = {2, 3, 4, 6, 8, 10, 12, 16};
#pragma unroll
  for (int i = 0; i < 8; ++i)
    if (ws[i] >= wmax || base + 64 * ws[i] > last) return min(ws[i], wmax);
  return wmax;
}

// Loads a window of W words per lane based at word b. Words that were in
// the staged window (based at ob, ow words per lane) come from LDS, the rest
// start all ones with their text from the store.
template <int W, bool SPLIT>
__device__ __forceinline__ void load_window64(const Pair64& P, const Stage64& X, int lane, int b, int ob, int ow,
                                              uint32_t (&v)[W], uint32_t (&tl)[W], uint32_t (&th)[W]) {
  // selects rather than branches: writing the word arrays in divergent
  // branches would move them to scratch memory. Two passes (staged words
  // from LDS, then new words from the store), so that each word's loads go
  // straight to its registers: one pass held every word's LDS and store
  // values at once, the register peak of the wide tiles.
  const int base = b + lane * W - ob;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const int rel = base + w;
    const bool staged = rel < 64 * ow;
    const int ri = staged ? rel : 0;
    const uint32_t sv = X.v[ri], sl = X.tl[ri], sh = X.th[ri];
    v[w] = staged ? sv : ~0u;
    tl[w] = staged ? sl : 0u;
    th[w] = staged ? sh : 0u;
  }
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const int g = b + lane * W + w;
    if (g - ob >= 64 * ow && g < P.ntw) {
      const uint2 t = SPLIT ? text_word64(P, g) : P.pk[P.toff + g];
      tl[w] = t.x;
      th[w] = t.y;
    }
  }
  __builtin_amdgcn_wave_barrier();
}

template <int W>
__device__ __forceinline__ void stage_window64(const Stage64& X, int lane, const uint32_t (&v)[W],
                                               const uint32_t (&tl)[W], const uint32_t (&th)[W]) {
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int w = 0; w < W; ++w) {
    X.v[lane * W + w] = v[w];
    X.tl[lane * W + w] = tl[w];
    X.th[lane * W + w] = th[w];
  }
  __builtin_amdgcn_wave_barrier();
}

// Window state of a pair between width changes.
struct Win64 {
  int b, jb, frozen;  // window base word, next 32-row block, zeros left of the window
  int ob, ow;         // staged window (base word, words per lane; ow = 0: none)
  int w;              // width to run next
  int w0;             // widest width (the tile's)
};

// Runs 32-row blocks of the pair with W words per lane, from the state in
// S (window contents staged in X). Returns 1 (accept) or 0 (reject), or -1
// after staging the window when the next block needs another width (S.w).
template <int W, bool SPLIT, bool KU>
__device__ __forceinline__ int run_blocks64(const Pair64& P, const Stage64& X, int lane, Win64& S,
                                            unsigned long long& cells, unsigned long long& cdec) {
  const int n = P.n, mq = P.mq, c = P.c;
  int b = S.b, jb = S.jb, frozen = S.frozen;
  uint32_t v[W], tl[W], th[W];
  load_window64<W, SPLIT>(P, X, lane, b, S.ob, S.ow, v, tl, th);
  while (true) {
    // ---- the 32 rows of block jb (fewer in the last block)
    const uint2 qw = SPLIT ? query_word64(P, jb) : P.pk[P.qoff + jb];
    const uint32_t qls = __builtin_amdgcn_readfirstlane(qw.x);
    const uint32_t qhs = __builtin_amdgcn_readfirstlane(qw.y);
    const int nrows = min(32, (SPLIT ? P.rows : mq) - jb * 32);
    if constexpr (W >= 6) {
      // With 6 or more words per lane each row branches (scalar) to the
      // code specialised for its base; the empty asm keeps the four match
      // masks from being hoisted out of the loop (4*W more registers).
      auto row = [&](uint32_t base) {
        uint32_t s[W], u[W], carry;
#pragma unroll
        for (int w = 0; w < W; ++w) asm volatile("" : "+v"(tl[w]), "+v"(th[w]));
        if (base == 0u) row_sum_base<0, W, KU>(v, tl, th, s, u, carry);
        else if (base == 1u) row_sum_base<1, W, KU>(v, tl, th, s, u, carry);
        else if (base == 2u) row_sum_base<2, W, KU>(v, tl, th, s, u, carry);
        else row_sum_base<3, W, KU>(v, tl, th, s, u, carry);
        row_finish64<W, KU>(v, s, u, carry, lane);
      };
      if (W <= 8 && nrows == 32) {
        // a full block in chunks of 8 rows with constant bit offsets (the
        // wider tiles would spill more registers so, and keep the row loop)
#pragma unroll 1
        for (int r0 = 0; r0 < 32; r0 += 8) {
          const uint32_t qlc = qls >> r0, qhc = qhs >> r0;
#pragma unroll
          for (int r = 0; r < 8; ++r) row(((qlc >> r) & 1u) | (((qhc >> r) & 1u) << 1));
        }
      } else {
#pragma unroll 1
        for (int r = 0; r < nrows; ++r) row(((qls >> r) & 1u) | (((qhs >> r) & 1u) << 1));
      }
    } else if (nrows == 32) {
      // With fewer words the base enters as uniform masks: measured faster
      // there than the branch, whose scalar issue outweighs the two bit
      // operations per word it saves. A full block runs in chunks of 8 rows
      // with constant bit offsets: each mask is one scalar bit-field
      // extract and the loop counts chunks, not rows (the scalar unit also
      // issues the carries' look-ahead, at the vector unit's rate).
#pragma unroll 1
      for (int r0 = 0; r0 < 32; r0 += 8) {
        const uint32_t qlc = qls >> r0, qhc = qhs >> r0;
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          const uint32_t ql = (uint32_t)((int32_t)(qlc << (31 - r)) >> 31);
          const uint32_t qh = (uint32_t)((int32_t)(qhc << (31 - r)) >> 31);
          row_masks64<W>(v, tl, th, ql, qh, lane);
        }
      }
    } else {
#pragma unroll 1
      for (int r = 0; r < nrows; ++r) {
        const uint32_t ql = (uint32_t)((int32_t)(qls << (31 - r)) >> 31);
        const uint32_t qh = (uint32_t)((int32_t)(qhs << (31 - r)) >> 31);
        row_masks64<W>(v, tl, th, ql, qh, lane);
      }
    }
    cells += 64ull * W * 32ull * 32ull;  // the same in every lane: a scalar register

    // ---- block end: zero counts per lane (all columns, and columns < n-R)
    const int jl = jb * 32 + nrows - 1;  // last row done
    const int R = mq - 1 - jl;           // rows left
    const int istar = max(0, n - R);
    uint32_t x = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const int g = b + lane * W + w;
      x += (uint32_t)__popc(~v[w] & valid_bits(g, n)) | ((uint32_t)__popc(~v[w] & valid_bits(g, istar)) << 16);
    }
    uint32_t tot = x;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
    tot = __builtin_amdgcn_readfirstlane(tot);
    const int z = frozen + (int)(tot & 0xffffu);
    // columns < n-R left of the window are all counted: an upper bound
    const int dstar = frozen + (int)(tot >> 16);
    if (z >= c || dstar + R < c || R == 0) {
      cdec += (unsigned long long)uniform64((int64_t)band_cells_rows(n, mq, c, jl + 1));
      return z >= c ? 1 : 0;
    }
    if constexpr (SPLIT) {
      if (jl + 1 >= P.rows) {
        // a half of a split pair at its last row, undecided: its window goes
        // to LDS for the join at the middle row
        stage_window64<W>(X, lane, v, tl, th);
        S.b = b; S.jb = jb + 1; S.frozen = frozen; S.ob = b; S.ow = W; S.w = W;
        return -2;
      }
      if (*P.stop) return -3;  // the other half has decided
    }
    ++jb;

    // ---- window for the next block: from the first live word, reaching
    // the band's right edge at the block's last row. The DP edge (a prefix
    // scan of the lane counts) is taken every few blocks, or when the
    // window must move anyway; otherwise the diagonal edge alone is used
    // (dropping fewer columns is always exact).
    const int jn = jb * 32;
    const int hi_col = min(n - 1, jn + 31 + n - c);  // right edge at the last row of the next block
    const int last = hi_col >> 5;
    int lo_col = max(0, jn - (mq - c));
    const bool covered = b + 64 * W > last;
    // periods are powers of two: a mask (P.every = period - 1, or -1 for
    // the default of this width)
    constexpr int kEvery = W >= 6 ? 2 : 4;
    const bool rebase = (jb & (P.every >= 0 ? P.every : kEvery - 1)) == 0;
    if (!covered || (W > 2 && rebase)) {
      // inclusive prefix over the lanes: D at the end of each lane's words;
      // lanes whose last column has D + R < c are dead (a prefix of lanes)
      uint32_t pre = x & 0xffffu;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(pre, d);
        if (lane >= d) pre += y;
      }
      const bool dead = frozen + (int)pre + R < c;
      const int ndead = __popcll(__builtin_amdgcn_ballot_w64(dead));
      const int dead_col = (b + ndead * W) * 32;  // columns before it are dead after row jl
      if (dead_col > min(n - 1, jl + n - c)) {
        // every column of row jl's band is dead: no path of c matches exists
        cdec += (unsigned long long)uniform64((int64_t)band_cells_rows(n, mq, c, jl + 1));
        return 0;
      }
      lo_col = max(lo_col, dead_col);
      const int nb = max(b, lo_col >> 5);
      const int w2 = width_to(nb, last, S.w0);
      if (w2 < W || !covered) {
        // zeros of the words dropped on the left
        int dz = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) {
          const int g = b + lane * W + w;
          if (g < nb) dz += __popc(~v[w] & valid_bits(g, n));
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) dz += __shfl_xor(dz, off);
        stage_window64<W>(X, lane, v, tl, th);
        frozen += uniform(dz);
        // the window moves (and may change width): the caller re-enters at
        // the new base, whose first act reloads the window from LDS (one
        // place where the word arrays are defined keeps them in the same
        // registers; a reload here made the compiler copy them at the join)
        S.b = nb; S.jb = jb; S.frozen = frozen; S.ob = b; S.ow = W; S.w = w2;
        return -1;
      }
    }
  }
}

// Decides LCS >= c for a pair on a 64-lane tile of at most W words per lane,
// from the window state S (width S.w, at most S.w0 <= W). Returns 1 / 0, or
// for a half of a split pair -2 (rows done, window staged in X) or -3
// (stopped: the other half decided).
template <int W, bool SPLIT>
__device__ __forceinline__ int decide64(const Pair64& P, const Stage64& X, int lane, Win64& S,
                                        unsigned long long& cells, unsigned long long& cdec) {
  // Tiles of up to 8 words a lane keep U through the row (one operation per
  // word fewer); the wider tiles give those registers to occupancy
  constexpr bool KU = W <= 8;
  while (true) {
    int r = -1;
    switch (S.w) {
      case 16: if constexpr (W >= 16) r = run_blocks64<16, SPLIT, KU>(P, X, lane, S, cells, cdec); break;
      case 12: if constexpr (W >= 12) r = run_blocks64<12, SPLIT, KU>(P, X, lane, S, cells, cdec); break;
      case 10: if constexpr (W >= 10) r = run_blocks64<10, SPLIT, KU>(P, X, lane, S, cells, cdec); break;
      case 8: if constexpr (W >= 8) r = run_blocks64<8, SPLIT, KU>(P, X, lane, S, cells, cdec); break;
      case 6: if constexpr (W >= 6) r = run_blocks64<6, SPLIT, KU>(P, X, lane, S, cells, cdec); break;
      case 4: if constexpr (W >= 4) r = run_blocks64<4, SPLIT, KU>(P, X, lane, S, cells, cdec); break;
      case 3: if constexpr (W >= 3) r = run_blocks64<3, SPLIT, KU>(P, X, lane, S, cells, cdec); break;
      default: r = run_blocks64<2, SPLIT, KU>(P, X, lane, S, cells, cdec); break;
    }
    if (r != -1) return r;
  }
}

template <int W>
__device__ inline void align_loop64(const AlignArgs& a) {
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  __shared__ uint32_t stage[4][3][64 * W];
  const Stage64 X{stage[wid][0], stage[wid][1], stage[wid][2]};
  unsigned long long cells = 0, cdec = 0;
  uint32_t nacc = 0, nrej = 0, nskip = 0;  // pairs per wave (32-bit: the u64 forms spilled in 64x8)
  const int total = *a.count;
  if (total == 0 || work_taken(a.work, total)) return;
  while (true) {
    int got = 0;
    if (lane == 0) got = atomicAdd(a.work, 1);
    got = __builtin_amdgcn_readfirstlane(__shfl(got, 0));
    if (got >= total) break;
    // the pair's scalars are wave-uniform: reading them through the first
    // lane keeps them (and the control flow that depends on them) scalar
    const AlignItem it = a.items[got];
    const int text = uniform(it.text), query = uniform(it.query), jobi = uniform(it.job), rslot = uniform(it.slot);
    if (rslot & kPigeonFlag) continue;  // rejected (and counted) by the pigeonhole test
    const JobDev& J = a.jobs[jobi];
    const StoreView& st = J.st;
    const int islot = rslot & (kSwapFlag - 1);
    const int cand = (rslot & kSwapFlag) ? text : query;  // the candidate owns the cutoff
    const int n = uniform(st.m[text]), mq = uniform(st.m[query]), c = uniform(st.bc[cand]) + 1;
    if (a.skip_absorbed && uniform64(J.acc[cand] & J.state->tmask & ((1ull << islot) - 1ull))) {
      ++nskip;  // per-wave counters are uniform: scalar registers
      continue;
    }
    int res = 0;
    if (c <= min(n, mq)) {
      const Pair64 P{st.pk, uniform64(st.pkoff[text]), uniform64(st.pkoff[query]), n, mq, c, (n >> 5) + 1,
                     a.win_period - 1, mq, 0, nullptr};
      // a step ends when its longest pair does: waves on long pairs get the
      // SIMD's issue slots first, the short pairs fill the gaps
      if (mq >= 16384) __builtin_amdgcn_s_setprio(3);
      else if (mq >= 4096) __builtin_amdgcn_s_setprio(2);
      else if (mq >= 1024) __builtin_amdgcn_s_setprio(1);
      Win64 S{0, 0, 0, 0, 0, W, W};
      res = decide64<W, false>(P, X, lane, S, cells, cdec);
      __builtin_amdgcn_s_setprio(0);
    }
    if (lane == 0) {
      if (a.result) {
        a.result[got] = res;
      } else if (res) {
        atomicOr(&J.acc[cand], 1ull << islot);
      }
    }
    if (res) ++nacc; else ++nrej;
  }
  if (lane == 0 && a.stats) {
    unsigned long long* sr = stat_row(a.stats);
    if (cells) atomicAdd(&sr[ST_CELLS_EXEC], cells);
    if (cdec) atomicAdd(&sr[ST_CELLS_DECIDED], cdec);
    if (nskip) atomicAdd(&sr[ST_SPEC_WASTE], (unsigned long long)nskip);
    if (nacc) atomicAdd(&sr[ST_ACCEPT], (unsigned long long)nacc);
    if (nrej) atomicAdd(&sr[ST_REJ_ALIGN], (unsigned long long)nrej);
  }
}

// ----------------------------------------------------------- split pairs
//
// A long pair of the step's speculative representatives against each other
// ends its phase when it runs alone: one wave's chain of dependent rows. Its
// rows split at h (a multiple of 32 near m/2): wave 0 of a 128-thread block
// runs rows 0..h-1 of (text, query), wave 1 the other m-h rows on the
// reversed pair (reversing both sequences keeps every common subsequence),
// each with the banded window, exact early decisions and rebases of the
// 64-lane tile code, the bounds counting the rows of both halves (R = rows
// left of all m). A half that decides alone decides the pair: its accept is
// a common subsequence of c bases, its reject bounds every path through all
// m rows. Otherwise the halves join at the middle row (Hirschberg's split):
// with F(i) the forward value over text columns < i and B(i) the backward
// value over columns >= i, LCS >= F(i) + B(i) for every i. For a path of at
// least c matches, let A be one past its last forward match and Bc its
// first backward match column: no column in [A, Bc] is ever frozen by
// either half (each such column keeps the path's count plus what it can
// still gain at c or more), the window values there are at least the
// path's forward and backward match counts (monotone in the column), so
// some i in [A, Bc] inside both windows has F(i) + B(i) >= c. Hence the pair
// is accepted iff the maximum over the columns both windows hold reaches c.

// Zero flags (columns with a match recorded) of a staged window, as text
// word g of the staged sequence's own orientation; 0 outside the window and
// in padding columns.
__device__ __forceinline__ uint32_t staged_zeros(const uint32_t* v, int base, int width, int n, int g) {
  const int k = g - base;
  if (k < 0 || k >= 64 * width) return 0u;
  return ~v[k] & valid_bits(g, n);
}

// The join at the middle row (wave 0): forward window (base bf words,
// width wf, ff zeros frozen left of it), backward window in reversed
// coordinates (bb, wb, fb). Columns i range over [32 bf, n - 32 bb]; F(i) =
// ff + forward zeros in [32 bf, i), B(i) = fb + backward zeros in
// [i, n - 32 bb). Returns 1 when max_i F(i) + B(i) >= c.
__device__ int join_halves(const uint32_t* vf, int bf, int wf, int ff, const uint32_t* vb, int bb, int wb, int fb,
                           int n, int c, int lane) {
  const int L = 32 * bf, U = n - 32 * bb;
  if (L > U) return 0;  // no column is live in both halves: no path of c matches
  const int g0 = bf, g1 = (U + 31) >> 5;  // text words holding columns [L, U)
  const int per = (g1 - g0 + 63) >> 6;
  const int ga = g0 + lane * per, gz = min(g1, ga + per);
  int run = 0, best = 0, btot = 0;
  for (int g = ga; g < gz; ++g) {
    uint32_t za = staged_zeros(vf, bf, wf, n, g);
    // backward zeros of columns 32g .. 32g+31: reversed columns n-1-32g
    // down to n-32-32g, the 32 reversed bits from p = n - 32(g+1), bit-reversed
    const int p = n - 32 * (g + 1);
    const int rw = p >> 5, sft = p & 31;  // floor division for negative p
    const uint32_t lo = staged_zeros(vb, bb, wb, n, rw), hi = staged_zeros(vb, bb, wb, n, rw + 1);
    uint32_t zb = __builtin_bitreverse32(sft ? (lo >> sft) | (hi << (32 - sft)) : lo);
    if (g == (U >> 5)) {  // the word holding U: columns < U only
      const uint32_t keep = (1u << (U & 31)) - 1u;
      za &= keep;
      zb &= keep;
    }
    btot += __popc(zb);
    for (int t = 0; t < 32; ++t) {
      run += (int)((za >> t) & 1u) - (int)((zb >> t) & 1u);
      best = max(best, run);
    }
  }
  // value(i) = ff + fb + (backward zeros of [L, U)) + sum over [L, i) of
  // (forward - backward zeros): the best prefix over all lanes
  int pre = run;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(pre, d);
    if (lane >= d) pre += y;
  }
  int m = pre - run + best;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    m = max(m, __shfl_xor(m, off));
    btot += __shfl_xor(btot, off);
  }
  return ff + fb + btot + m >= c ? 1 : 0;
}

// One split pair per 128-thread block at a time, from the split list (phase
// A: no skipping of absorbed candidates). With more pairs than blocks the
// halves would only queue behind each other, so each wave then aligns whole
// pairs.
__global__ __launch_bounds__(128) void align_split_kernel(AlignArgs a) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __shared__ uint32_t stage[2][3][64 * 16];
  __shared__ int s_item, s_stop;
  __shared__ int s_res[2], s_b[2], s_w[2], s_frozen[2];
  const Stage64 X{stage[wid][0], stage[wid][1], stage[wid][2]};
  unsigned long long cells = 0, nacc = 0, nrej = 0, cdec = 0;
  const int total = *a.count;
  if (total == 0) return;
  const bool whole = total > (int)gridDim.x;  // the same in every block
  while (whole) {
    int got = 0;
    if (lane == 0) got = atomicAdd(a.work, 1);
    got = __builtin_amdgcn_readfirstlane(__shfl(got, 0));
    if (got >= total) break;
    const AlignItem it = a.items[got];
    const int text = uniform(it.text), query = uniform(it.query), jobi = uniform(it.job), rslot = uniform(it.slot);
    const JobDev& J = a.jobs[jobi];
    const StoreView& st = J.st;
    const int islot = rslot & (kSwapFlag - 1);
    const int cand = (rslot & kSwapFlag) ? text : query;
    const int n = uniform(st.m[text]), mq = uniform(st.m[query]), c = uniform(st.bc[cand]) + 1;
    int res = 0;
    if (!(rslot & kPigeonFlag) && c <= min(n, mq)) {
      const int w0 = max(2, kShapes[(rslot >> kSplitShapeShift) & 15].w);
      const Pair64 P{st.pk, uniform64(st.pkoff[text]), uniform64(st.pkoff[query]), n, mq, c, (n >> 5) + 1,
                     a.win_period - 1, mq, 0, nullptr};
      Win64 S{0, 0, 0, 0, 0, w0, w0};
      res = decide64<16, false>(P, X, lane, S, cells, cdec);
    }
    if (lane == 0) {
      if (a.result) {
        a.result[got] = res;
      } else if (res) {
        atomicOr(&J.acc[cand], 1ull << islot);
      }
      if (res) ++nacc; else ++nrej;
    }
  }
  while (!whole) {
    __syncthreads();  // the previous pair's staged windows are no longer read
    if (threadIdx.x == 0) {
      s_item = atomicAdd(a.work, 1);
      s_stop = 0;
    }
    __syncthreads();
    const int got = s_item;
    if (got >= total) break;  // the same for both waves
    const AlignItem it = a.items[got];
    const int text = uniform(it.text), query = uniform(it.query), jobi = uniform(it.job), rslot = uniform(it.slot);
    const JobDev& J = a.jobs[jobi];
    const StoreView& st = J.st;
    const int islot = rslot & (kSwapFlag - 1);
    const int cand = (rslot & kSwapFlag) ? text : query;
    const int n = uniform(st.m[text]), mq = uniform(st.m[query]), c = uniform(st.bc[cand]) + 1;
    // both halves hold at least 32 rows (the split list takes m >= 64)
    const bool work = !(rslot & kPigeonFlag) && c <= min(n, mq) && mq >= 64;
    if (work) {
      const int h = ((mq >> 1) >> 5) << 5;
      const int w0 = max(2, kShapes[(rslot >> kSplitShapeShift) & 15].w);
      const Pair64 P{st.pk, uniform64(st.pkoff[text]), uniform64(st.pkoff[query]), n, mq, c, (n >> 5) + 1,
                     a.win_period - 1, wid == 0 ? h : mq - h, wid, &s_stop};
      Win64 S{0, 0, 0, 0, 0, w0, w0};
      const int r = decide64<16, true>(P, X, lane, S, cells, cdec);
      if (lane == 0) {
        if (r >= 0) s_stop = 1;  // the other half may stop
        s_res[wid] = r;
        s_b[wid] = S.b;
        s_w[wid] = S.ow;
        s_frozen[wid] = S.frozen;
      }
    }
    __syncthreads();
    if (wid == 0) {
      int res = 0;
      if (work) {
        const int r0 = s_res[0], r1 = s_res[1];
        if (r0 == 1 || r1 == 1) {
          res = 1;
        } else if (r0 == 0 || r1 == 0) {
          res = 0;
        } else {
          res = join_halves(stage[0][0], s_b[0], s_w[0], s_frozen[0], stage[1][0], s_b[1], s_w[1], s_frozen[1], n,
                            c, lane);
        }
      }
      if (lane == 0) {
        if (a.result) {
          a.result[got] = res;
        } else if (res) {
          atomicOr(&J.acc[cand], 1ull << islot);
        }
        if (res) ++nacc; else ++nrej;
      }
    }
  }
  if (lane == 0 && a.stats) {
    unsigned long long* sr = stat_row(a.stats);
    if (cells) atomicAdd(&sr[ST_CELLS_EXEC], cells);
    if (cdec) atomicAdd(&sr[ST_CELLS_DECIDED], cdec);
    if (nacc) atomicAdd(&sr[ST_ACCEPT], nacc);
    if (nrej) atomicAdd(&sr[ST_REJ_ALIGN], nrej);
  }
}

// Waves per SIMD the register allocation must allow (HIP's second launch
// bound): 8 for the narrow one-word tiles (62 registers, no spills; lets a
// grid of 8 blocks per CU be resident), 4 up to 10 words per lane (128
// registers: the per-wave counters in scalar registers, the window reloaded
// in one place), 3 for 12 and 16 words; none spills.
__host__ __device__ constexpr int align_waves_per_eu(int s, int w) {
  return s < 64 ? 8 : w <= 10 ? 4 : 3;
}

template <int S, int W>
__global__ __launch_bounds__(256, align_waves_per_eu(S, W)) void align_kernel(AlignArgs a) {
  if constexpr (S == 64 && W >= 2) {
    align_loop64<W>(a);
  } else {
    align_loop<S, W>(a);
  }
}

#endif  // GSDR_DEVICE_TU

// host launchers (kernels.hip)
// kmask: bit k set for every word length k among the jobs; one launch per
// class of word lengths that share a table group size.
hipError_t launch_filter(const FilterArgs& a, unsigned kmask, int grid, hipStream_t s);
hipError_t launch_align(int shape, const AlignArgs& a, int grid, hipStream_t s);
// The split list (kSplitList): one pair per 128-thread block at a time.
hipError_t launch_align_split(const AlignArgs& a, int grid, hipStream_t s);
int align_blocks_per_cu(int shape);
hipError_t launch_resolve(const ResolveArgs& a, hipStream_t s);
hipError_t launch_resolve_reps(const ResolveArgs& a, hipStream_t s);
hipError_t launch_clamp_spec(const JobDev* jobs, int njobs, long long pair_budget, hipStream_t s);
// Per sequence, the sum of the counts stored in its k-mer profile.
hipError_t launch_profile_totals(const uint32_t* kp, const int64_t* kpoff, int n, int32_t* ksum, hipStream_t s);
hipError_t launch_pigeon(const PigeonArgs& a, int grid, hipStream_t s);
// Signature screen of phase B (jobs with signatures), before the filter.
hipError_t launch_screen(const ScreenArgs& a, int grid, hipStream_t s);
// Word-presence signatures of n sequences (kSigWords words each).
hipError_t launch_signatures(const uint32_t* kp, const int64_t* kpoff, int n, int k, uint32_t* sig, hipStream_t s);
// Cutoffs of n sequences for threshold t and word length k (func.cpp:339-374).
// (also into the records' bc and wc when rec is not null)
hipError_t launch_cutoffs(const int32_t* raw_len, int n, float t, int k, int32_t* bc, int32_t* wc, SeqRec* rec,
                          hipStream_t s);
// Fills the records' length, profile total, composition and profile place.
hipError_t launch_pack_records(const int32_t* m, const uint4* comp, const int64_t* kpoff, const int32_t* ksum, int n,
                               SeqRec* rec, hipStream_t s);
hipError_t launch_init_jobs(const JobDev* jobs_dev, const JobDev* jobs_host, int njobs, int k_max,
                            int32_t* jobs_left, hipStream_t s);

}  // namespace gsdr
