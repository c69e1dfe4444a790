// Device code shared by the plain and batched launch forms.
#pragma once
#include "kernels.hpp"

namespace gsdr {

struct FilterArgs {
  StoreView st;
  const int32_t* cand;      // candidate sorted indices (ascending)
  const int32_t* ncp;       // number of candidates (device)
  const int32_t* repp;      // representative (device)
  int2* items;              // per-shape pair lists: items + shape * item_cap
  int32_t item_cap;
  int32_t min_shape;
  int32_t* bin_count;       // kNumShapes counters
  unsigned long long* stats;
};

struct AlignArgs {
  StoreView st;
  const int2* items;        // (text, query) sorted indices
  const int32_t* count;     // number of items (device)
  int32_t* work;            // work counter (device, starts at 0)
  int32_t* cluster;         // non-null: cluster[query] = text on accept
  int32_t* result;          // non-null: result[item] = 1 accept / 0 reject
  unsigned long long* stats;
};

// ----------------------------------------------------------- block helpers
__device__ inline int wave_sum(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

__device__ inline int block_sum_256(int v) {
  __shared__ int part[4];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) part[wave] = v;
  __syncthreads();
  return part[0] + part[1] + part[2] + part[3];
}

// Exclusive scan over a 256-thread block; *total gets the block sum.
__device__ inline int block_exclusive_scan_256(int v, int* total) {
  __shared__ int part[4];
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
  int before = 0;
  for (int w = 0; w < wave; ++w) before += part[w];
  *total = part[0] + part[1] + part[2] + part[3];
  return before + incl - v;
}

// ------------------------------------------------------------ filter body
// Wave loop of the filter kernel; `table` holds c_r(w) for representative r.
__device__ inline void filter_wave_loop(const FilterArgs& a, const uint16_t* table, int r, int nc) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int waves_per_block = blockDim.x >> 6;
  const int n = a.st.m[r];
  const uint4 cr = a.st.comp[r];
  unsigned long long rej_len = 0, rej_comp = 0, rej_word = 0, passed = 0;
  unsigned long long bytes = 0, cband = 0, cfull = 0;
  for (int base = (blockIdx.x * waves_per_block + wave) * 64; base < nc;
       base += gridDim.x * waves_per_block * 64) {
    const int idx = base + lane;
    const bool in = idx < nc;
    int q = 0, mq = 0, bcq = 0;
    bool ok = false;
    if (in) {
      q = a.cand[idx];
      mq = a.st.m[q];
      bcq = a.st.bc[q];
      bytes += 28;  // id, m, bc, comp
      const int c = bcq + 1;
      if (c > min(n, mq)) {
        ++rej_len;  // LCS <= min(n, m) < c
      } else {
        const uint4 cq = a.st.comp[q];
        const int sum = (int)(min(cr.x, cq.x) + min(cr.y, cq.y) + min(cr.z, cq.z) + min(cr.w, cq.w));
        // reference F1 rejects sum < bc; LCS <= sum also rejects sum == bc
        if (sum <= bcq) ++rej_comp;
        else ok = true;
      }
    }
    uint64_t surv = __builtin_amdgcn_ballot_w64(ok);
    while (surv) {
      const int bit = __builtin_ctzll(surv);
      surv &= surv - 1;
      const int q2 = __shfl(q, bit);
      const int m2 = __shfl(mq, bit);
      const int b2 = __shfl(bcq, bit);
      const int64_t e0 = a.st.kpoff[q2], e1 = a.st.kpoff[q2 + 1];
      int acc = 0;
      for (int64_t e = e0 + lane; e < e1; e += 64) {
        const uint32_t x = a.st.kp[e];
        acc += min((uint32_t)table[x & 0xFFFFu], x >> 16);
      }
      acc = wave_sum(acc);
      if (lane == 0) {
        bytes += 20 + 4ull * (unsigned long long)(e1 - e0);  // kpoff pair, wc, profile
        if (acc < a.st.wc[q2]) {
          ++rej_word;
        } else {
          ++passed;
          bytes += 8;
          cband += (unsigned long long)band_cells(n, m2, b2 + 1);
          cfull += (unsigned long long)n * (unsigned long long)m2;
          const int sh = pick_shape(n, m2, b2 + 1, a.min_shape);
          const int pos = atomicAdd(&a.bin_count[sh], 1);
          a.items[(int64_t)sh * a.item_cap + pos] = make_int2(r, q2);
        }
      }
    }
  }
  // per-wave counters
  unsigned long long v[7] = {rej_len, rej_comp, rej_word, passed, bytes, cband, cfull};
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    unsigned long long x = v[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    v[i] = x;
  }
  if (lane == 0) {
    const unsigned long long pairs = v[0] + v[1] + v[2] + v[3];
    if (pairs) {
      atomicAdd(&a.stats[ST_PAIRS], pairs);
      if (v[0]) atomicAdd(&a.stats[ST_REJ_LEN], v[0]);
      if (v[1]) atomicAdd(&a.stats[ST_REJ_COMP], v[1]);
      if (v[2]) atomicAdd(&a.stats[ST_REJ_WORD], v[2]);
      if (v[3]) atomicAdd(&a.stats[ST_ALIGNED], v[3]);
      atomicAdd(&a.stats[ST_FILTER_BYTES], v[4]);
      if (v[5]) atomicAdd(&a.stats[ST_CELLS_BAND], v[5]);
      if (v[6]) atomicAdd(&a.stats[ST_CELLS_FULL], v[6]);
    }
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
template <int S, int W>
__device__ inline void align_loop(const AlignArgs& a) {
  constexpr uint64_t KTOP = top_lanes<S>();
  const int lane = threadIdx.x & 63;
  const int sl = lane & (S - 1);
  const int seg0 = lane & ~(S - 1);

  uint32_t v[W], tl[W], th[W];
#pragma unroll
  for (int w = 0; w < W; ++w) { v[w] = ~0u; tl[w] = 0u; th[w] = 0u; }

  int item = -1;
  bool exhausted = false;
  int tx = 0, q = 0, n = 0, mq = 0, c = 0, b = 0, jb = 0, frozen = 0, ntw = 0;
  int64_t toff = 0, qoff = 0;
  unsigned long long cells = 0, nacc = 0, nrej = 0, cdec = 0;
  const int total = *a.count;

  while (true) {
    // ---- fetch a pair for every segment without one
    const bool need = (item < 0) && !exhausted;
    int got = 0;
    if (need && sl == 0) got = atomicAdd(a.work, 1);
    got = __shfl(got, seg0);
    if (need) {
      if (got < total) {
        item = got;
        const int2 it = a.items[got];
        tx = it.x;
        q = it.y;
        n = a.st.m[tx];
        mq = a.st.m[q];
        c = a.st.bc[q] + 1;
        toff = a.st.pkoff[tx];
        qoff = a.st.pkoff[q];
        ntw = (n >> 5) + 1;
        b = 0;
        jb = 0;
        frozen = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) {
          const int gw = sl * W + w;
          uint2 t = make_uint2(0u, 0u);
          if (gw < ntw) t = a.st.pk[toff + gw];
          tl[w] = t.x;
          th[w] = t.y;
          v[w] = ~0u;
        }
      } else {
        exhausted = true;
      }
    }
    if (__builtin_amdgcn_ballot_w64(!exhausted) == 0) break;
    const bool active = item >= 0;
    // trivially impossible pairs end at the first block end (LCS <= min(n, m))
    const bool hopeless = active && (c > min(n, mq));

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
            if (gw < ntw) t = a.st.pk[toff + gw];
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
      const uint2 qw = a.st.pk[qoff + jb];
      qlo = qw.x;
      qhi = qw.y;
      const int rows = mq - jb * 32;
      qnv = rows >= 32 ? 0u : ~((1u << rows) - 1u);
    }

    // ---- 32 rows, unrolled in chunks of UR rows with constant bit offsets
    constexpr int UR = W <= 4 ? 32 : 8;
    for (int r0 = 0; r0 < 32; r0 += UR) {
     const uint32_t qlc = qlo >> r0, qhc = qhi >> r0, qnc = qnv >> r0;
#pragma unroll
     for (int r = 0; r < UR; ++r) {
      const uint32_t ql = 0u - ((qlc >> r) & 1u);
      const uint32_t qh = 0u - ((qhc >> r) & 1u);
      const uint32_t qn = 0u - ((qnc >> r) & 1u);
      uint32_t x[W], s[W];
      uint32_t carry = 0u, pall = ~0u;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        x[w] = (tl[w] ^ ql) | (th[w] ^ qh) | qn;
        const uint32_t u = v[w] & ~x[w];
        uint32_t co;
        s[w] = __builtin_addc(v[w], u, carry, &co);
        carry = co;
        pall &= s[w];
      }
      const uint64_t G = __builtin_amdgcn_ballot_w64(carry != 0u) & ~KTOP;
      const uint64_t P = __builtin_amdgcn_ballot_w64(pall == ~0u) & ~KTOP;
      const uint64_t X = G | P;
      const uint64_t C = (X + G) ^ X ^ G;
      uint32_t cin = (uint32_t)(C >> lane) & 1u;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        uint32_t co;
        const uint32_t t = __builtin_addc(s[w], 0u, cin, &co);
        cin = co;
        v[w] = t | (v[w] & x[w]);
      }
     }
    }

    // ---- block end: exact early decisions on the zero count
    int zl = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) zl += __popc(~v[w] & valid_bits(b + sl * W + w, n));
#pragma unroll
    for (int off = S / 2; off > 0; off >>= 1) zl += __shfl_xor(zl, off);
    if (active) {
      const int z = frozen + zl;
      const int jl = min(jb * 32 + 31, mq - 1);
      bool accept = false, reject = false;
      if (hopeless) {
        reject = true;
      } else if (z >= c) {
        accept = true;
      } else if (z + (mq - 1 - jl) < c) {
        reject = true;
      }
      if (sl == 0) cells += (unsigned long long)S * W * 32ull * 32ull;
      if (accept || reject) {
        if (sl == 0) {
          if (!hopeless) cdec += (unsigned long long)band_cells_rows(n, mq, c, jb * 32 + 32);
          if (accept && a.cluster) a.cluster[q] = tx;
          if (a.result) a.result[item] = accept ? 1 : 0;
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
  }
  if (lane == 0 && a.stats) {
    if (cells) atomicAdd(&a.stats[ST_CELLS_EXEC], cells);
    if (cdec) atomicAdd(&a.stats[ST_CELLS_DECIDED], cdec);
    if (nacc) atomicAdd(&a.stats[ST_ACCEPT], nacc);
    if (nrej) atomicAdd(&a.stats[ST_REJ_ALIGN], nrej);
  }
}

template <int S, int W>
__global__ __launch_bounds__(256) void align_kernel(AlignArgs a) {
  align_loop<S, W>(a);
}

// host launchers (kernels.hip)
hipError_t launch_filter(int k, const FilterArgs& a, int grid, hipStream_t s);
hipError_t launch_align(int shape, const AlignArgs& a, int grid, hipStream_t s);
hipError_t launch_compact(const int32_t* cand, RoundState* st, int32_t* cluster, int32_t* bcount,
                          int32_t* out, int nc, int32_t* reset, int nreset, hipStream_t s);
hipError_t launch_init(int32_t* cluster, int32_t* cand, int n, RoundState* st, hipStream_t s);

}  // namespace gsdr
