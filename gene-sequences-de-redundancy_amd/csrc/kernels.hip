// gfx950 kernels of the per-round pipeline: shared-word/composition filter,
// banded bit-parallel LCS decision, and order-preserving candidate
// compaction with next-representative selection.
//
// Reference stages replaced: kernel_makeTable/kernel_cleanTable,
// kernel_preFilter, kernel_filter, kernel_dynamic (func.cpp:499-676) and the
// host loops updateRepresentative/updateRemain/updatJobs (func.cpp:466-498).
#include "kernels.hpp"
#include "kernels_impl.hpp"

namespace gsdr {

// ------------------------------------------------------------------ filter
// One representative per launch. Each workgroup stages the representative's
// word counts c_r(w) in LDS (a u16 table of 4^k entries, the reference's
// global table at func.cpp:449), then each wave takes 64 candidates:
//   lane-per-candidate:  length check (added, sound), composition filter
//                        (func.cpp:525-543) plus its sound tightening;
//   wave-per-survivor:   sum over the candidate's distinct words of
//                        min(c_r(w), c_q(w)) against wc_q (func.cpp:544-571).
// Survivors are appended to the alignment list of their tile shape.
template <int K>
__global__ __launch_bounds__(256) void filter_kernel(FilterArgs a) {
  extern __shared__ uint32_t lds[];
  uint16_t* table = reinterpret_cast<uint16_t*>(lds);
  constexpr int NW = 1 << (2 * K);
  const int nc = *a.ncp;
  const int r = *a.repp;
  if (r < 0 || nc <= 0) return;
  const int tid = threadIdx.x;
  for (int i = tid; i < NW / 2; i += blockDim.x) lds[i] = 0u;
  __syncthreads();
  const int64_t k0 = a.st.kpoff[r], k1 = a.st.kpoff[r + 1];
  for (int64_t e = k0 + tid; e < k1; e += blockDim.x) {
    const uint32_t x = a.st.kp[e];
    table[x & 0xFFFFu] = (uint16_t)(x >> 16);
  }
  __syncthreads();
  filter_wave_loop(a, table, r, nc);
}

// --------------------------------------------------------------- compaction
// Order-preserving removal of absorbed candidates. The first survivor is the
// next representative (updateRepresentative, func.cpp:466-476); the rest form
// the next candidate list (updateRemain, func.cpp:477-487).
__global__ __launch_bounds__(256) void compact_count(const int32_t* __restrict__ cand, int nc,
                                                     const int32_t* __restrict__ cluster,
                                                     int32_t* __restrict__ bcount) {
  const int base = blockIdx.x * 1024 + threadIdx.x * 4;
  int cnt = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int i = base + k;
    if (i < nc && cluster[cand[i]] == -1) ++cnt;
  }
  cnt = block_sum_256(cnt);
  if (threadIdx.x == 0) bcount[blockIdx.x] = cnt;
}

__global__ __launch_bounds__(256) void compact_scatter(const int32_t* __restrict__ cand, int nc,
                                                       RoundState* __restrict__ st,
                                                       int32_t* __restrict__ cluster,
                                                       const int32_t* __restrict__ bcount,
                                                       int32_t* __restrict__ out, int nblocks,
                                                       int32_t* __restrict__ reset, int nreset) {
  // prefix over preceding blocks
  int pre = 0;
  for (int b = threadIdx.x; b < (int)blockIdx.x; b += 256) pre += bcount[b];
  pre = block_sum_256(pre);
  const int base = blockIdx.x * 1024 + threadIdx.x * 4;
  int32_t q[4];
  int f[4];
  int cnt = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int i = base + k;
    q[k] = i < nc ? cand[i] : -1;
    f[k] = (q[k] >= 0 && cluster[q[k]] == -1) ? 1 : 0;
    cnt += f[k];
  }
  int total_block = 0;
  const int excl = block_exclusive_scan_256(cnt, &total_block);
  int pos = pre + excl;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (f[k]) {
      if (pos == 0) {
        st->rep = q[k];
        cluster[q[k]] = q[k];
      } else {
        out[pos - 1] = q[k];
      }
      ++pos;
    }
  }
  if (blockIdx.x == (unsigned)(nblocks - 1) && threadIdx.x == 0) {
    const int total = pre + total_block;
    st->pad[0] = total;
    if (total == 0) st->rep = -1;
    st->nc = total > 0 ? total - 1 : 0;
  }
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < nreset; i += 256) reset[i] = 0;
}

// Initial state: representative 0, candidates 1..n-1, every cluster -1.
__global__ __launch_bounds__(256) void init_round(int32_t* __restrict__ cluster,
                                                  int32_t* __restrict__ cand, int n,
                                                  RoundState* __restrict__ st) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    cluster[i] = (i == 0) ? 0 : -1;
    if (i > 0) cand[i - 1] = i;
  }
  if (i == 0) {
    st->rep = n > 0 ? 0 : -1;
    st->nc = n > 0 ? n - 1 : 0;
    st->pad[0] = n;
    st->pad[1] = 0;
  }
}

// -------------------------------------------------------------- launchers
template <int K>
static hipError_t launch_filter_k(const FilterArgs& a, int grid, hipStream_t s) {
  const size_t lds = sizeof(uint16_t) << (2 * K);
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute((const void*)filter_kernel<K>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(filter_kernel<K>, dim3(grid), dim3(256), lds, s, a);
  return hipGetLastError();
}

hipError_t launch_filter(int k, const FilterArgs& a, int grid, hipStream_t s) {
  switch (k) {
    case 4: return launch_filter_k<4>(a, grid, s);
    case 5: return launch_filter_k<5>(a, grid, s);
    case 6: return launch_filter_k<6>(a, grid, s);
    case 7: return launch_filter_k<7>(a, grid, s);
    case 8: return launch_filter_k<8>(a, grid, s);
    default: return hipErrorInvalidValue;
  }
}

template <int S, int W>
static hipError_t launch_align_sw(const AlignArgs& a, int grid, hipStream_t s) {
  hipLaunchKernelGGL((align_kernel<S, W>), dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_align(int shape, const AlignArgs& a, int grid, hipStream_t s) {
  switch (shape) {
    case 0: return launch_align_sw<8, 1>(a, grid, s);
    case 1: return launch_align_sw<16, 1>(a, grid, s);
    case 2: return launch_align_sw<32, 1>(a, grid, s);
    case 3: return launch_align_sw<64, 1>(a, grid, s);
    case 4: return launch_align_sw<64, 2>(a, grid, s);
    case 5: return launch_align_sw<64, 4>(a, grid, s);
    case 6: return launch_align_sw<64, 8>(a, grid, s);
    case 7: return launch_align_sw<64, 16>(a, grid, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_compact(const int32_t* cand, RoundState* st, int32_t* cluster, int32_t* bcount,
                          int32_t* out, int nc, int32_t* reset, int nreset, hipStream_t s) {
  const int nblocks = nc > 0 ? (nc + 1023) / 1024 : 1;
  hipLaunchKernelGGL(compact_count, dim3(nblocks), dim3(256), 0, s, cand, nc, cluster, bcount);
  hipLaunchKernelGGL(compact_scatter, dim3(nblocks), dim3(256), 0, s, cand, nc, st, cluster, bcount,
                     out, nblocks, reset, nreset);
  return hipGetLastError();
}

hipError_t launch_init(int32_t* cluster, int32_t* cand, int n, RoundState* st, hipStream_t s) {
  const int grid = n > 0 ? (n + 255) / 256 : 1;
  hipLaunchKernelGGL(init_round, dim3(grid), dim3(256), 0, s, cluster, cand, n, st);
  return hipGetLastError();
}

}  // namespace gsdr
