// Device-side construction of the HBM store from the raw, length-sorted
// sequence bytes: base codes, composition, gap-free length, the gap-free
// bases as two bit planes and the distinct k-mer profile of every sequence.
//
// Reference stages replaced: baseToNumber, createPrefix, createWords,
// sortWords, mergeWords, deleteGap, compressData (func.cpp:146-446). The
// results equal the host restatement in host.cpp (preprocess): codes A/a 0,
// C/c 1, G/g 2, T/t/U/u 3, else 4 (func.cpp:155-188); composition of codes
// 0..3 (func.cpp:207-232); k-mers over raw windows without a code-4 base
// (func.cpp:252-338) counted modulo 2^16; bit j of plane word i is gap-free
// base 32i+j (func.cpp:404-446). Profile entries are stored in word order
// (the filter only sums over them).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "driver.hpp"
#include "kernels_impl.hpp"

namespace gsdr {

namespace {

__constant__ uint8_t c_code[256];

struct CodeTable {
  uint8_t v[256];
  CodeTable() {
    for (int i = 0; i < 256; ++i) v[i] = 4;
    v['A'] = v['a'] = 0;
    v['C'] = v['c'] = 1;
    v['G'] = v['g'] = 2;
    v['T'] = v['t'] = v['U'] = v['u'] = 3;
  }
};

constexpr int kPrepThreads = 256;

// Per-thread contiguous slice of a sequence of length L.
__device__ inline void slice(int64_t L, int64_t& i0, int64_t& i1) {
  const int64_t seg = (L + kPrepThreads - 1) / kPrepThreads;
  i0 = min(L, (int64_t)threadIdx.x * seg);
  i1 = min(L, i0 + seg);
}

// Counts the k-mers of the thread's slice into the LDS table (u16 counts,
// two per 32-bit word; a sequence has at most 32767 windows, so a half never
// carries into its neighbour). Returns the words this thread saw first.
__device__ inline int count_kmers(const uint8_t* d, int64_t i0, int64_t i1, int k, uint32_t* table) {
  const uint32_t mask = (1u << (2 * k)) - 1u;
  uint32_t word = 0;
  int valid = 0, first = 0;
  for (int64_t i = max((int64_t)0, i0 - k + 1); i < i1; ++i) {
    const uint32_t c = c_code[d[i]];
    if (c == 4) {
      valid = 0;
      word = 0;
      continue;
    }
    word = ((word << 2) | c) & mask;
    if (++valid >= k && i >= i0) {
      const uint32_t sh = (word & 1u) * 16u;
      const uint32_t old = atomicAdd(&table[word >> 1], 1u << sh);
      if (((old >> sh) & 0xFFFFu) == 0u) ++first;
    }
  }
  return first;
}

__device__ inline void clear(uint32_t* p, int words) {
  for (int i = threadIdx.x; i < words; i += kPrepThreads) p[i] = 0u;
}

// Pass 1: gap-free length, composition and number of distinct k-mers.
__global__ __launch_bounds__(kPrepThreads) void prep_count_kernel(const uint8_t* __restrict__ data,
                                                                 const int64_t* __restrict__ off, int k,
                                                                 int32_t* __restrict__ m_out,
                                                                 uint4* __restrict__ comp_out,
                                                                 int32_t* __restrict__ distinct_out) {
  HIP_DYNAMIC_SHARED(uint32_t, table)
  __shared__ int s_part[kPrepThreads / 64];
  const int64_t s = blockIdx.x;
  const int64_t a = off[s];
  const int64_t L = off[s + 1] - a;
  const uint8_t* d = data + a;
  clear(table, 1 << (2 * k - 1));
  __syncthreads();
  int64_t i0, i1;
  slice(L, i0, i1);
  int c4[4] = {0, 0, 0, 0};
  for (int64_t i = i0; i < i1; ++i) {
    const uint32_t c = c_code[d[i]];
    if (c < 4) ++c4[c];
  }
  const int first = count_kmers(d, i0, i1, k, table);
  const int ca = block_sum<kPrepThreads>(c4[0], s_part);
  const int cc = block_sum<kPrepThreads>(c4[1], s_part);
  const int cg = block_sum<kPrepThreads>(c4[2], s_part);
  const int ct = block_sum<kPrepThreads>(c4[3], s_part);
  const int dist = block_sum<kPrepThreads>(first, s_part);
  if (threadIdx.x == 0) {
    m_out[s] = ca + cc + cg + ct;
    comp_out[s] = make_uint4((uint32_t)ca, (uint32_t)cc, (uint32_t)cg, (uint32_t)ct);
    distinct_out[s] = dist;
  }
}

// Pass 2: the two bit planes and the (word | count << 16) profile entries.
// LDS: the k-mer table (4^k u16) followed by the planes (2 * (m/32 + 1)
// words, m <= 32767).
__global__ __launch_bounds__(kPrepThreads) void prep_write_kernel(const uint8_t* __restrict__ data,
                                                                 const int64_t* __restrict__ off, int k,
                                                                 const int64_t* __restrict__ pkoff,
                                                                 const int64_t* __restrict__ kpoff,
                                                                 uint2* __restrict__ pk, uint32_t* __restrict__ kp) {
  HIP_DYNAMIC_SHARED(uint32_t, lds)
  __shared__ int s_part[kPrepThreads / 64];
  const int tableWords = 1 << (2 * k - 1);
  uint32_t* table = lds;
  uint32_t* planes = lds + tableWords;
  const int64_t s = blockIdx.x;
  const int64_t a = off[s];
  const int64_t L = off[s + 1] - a;
  const uint8_t* d = data + a;
  const int pw = (int)(pkoff[s + 1] - pkoff[s]);  // m/32 + 1 words per plane
  clear(table, tableWords);
  clear(planes, 2 * pw);
  __syncthreads();
  int64_t i0, i1;
  slice(L, i0, i1);
  // gap-free index of the slice's first base
  int ng = 0;
  for (int64_t i = i0; i < i1; ++i) ng += c_code[d[i]] < 4;
  int total;
  int j = block_exclusive_scan<kPrepThreads>(ng, &total, s_part);
  for (int64_t i = i0; i < i1; ++i) {
    const uint32_t c = c_code[d[i]];
    if (c == 4) continue;
    const uint32_t bit = 1u << (j & 31);
    if (c & 1u) atomicOr(&planes[2 * (j >> 5)], bit);
    if (c & 2u) atomicOr(&planes[2 * (j >> 5) + 1], bit);
    ++j;
  }
  count_kmers(d, i0, i1, k, table);
  __syncthreads();
  uint2* out = pk + pkoff[s];
  for (int w = threadIdx.x; w < pw; w += kPrepThreads) out[w] = make_uint2(planes[2 * w], planes[2 * w + 1]);
  // profile entries in word order: each wave takes a quarter of the words,
  // counts its entries, and writes them after the waves before it
  const int nw = 1 << (2 * k);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int per = nw / (kPrepThreads / 64);
  const int w0 = wave * per;
  int mine = 0;
  for (int b = w0; b < w0 + per; b += 64) {
    const int wd = b + lane;
    const uint32_t cnt = (table[wd >> 1] >> ((wd & 1) * 16)) & 0xFFFFu;
    mine += __popcll(__builtin_amdgcn_ballot_w64(cnt != 0u));
  }
  if (lane == 0) s_part[wave] = mine;
  __syncthreads();
  int at = 0;
  for (int w = 0; w < wave; ++w) at += s_part[w];
  uint32_t* kout = kp + kpoff[s];
  for (int b = w0; b < w0 + per; b += 64) {
    const int wd = b + lane;
    const uint32_t cnt = (table[wd >> 1] >> ((wd & 1) * 16)) & 0xFFFFu;
    const uint64_t bal = __builtin_amdgcn_ballot_w64(cnt != 0u);
    if (cnt != 0u) kout[at + __popcll(bal & ((1ull << lane) - 1ull))] = (uint32_t)wd | (cnt << 16);
    at += __popcll(bal);
  }
}

}  // namespace

// Builds the store's sequence arrays on the device from the sorted records:
// uploads the raw bytes, counts (pass 1), sizes the packed and profile arrays
// from the counts, writes them (pass 2). Fills s.m (host), s.raw_len, s.max_m
// and the device arrays pk, pkoff, m_d, comp, kp, kpoff.
int store_build_device(gsdr_ctx* c, const Records& sorted, int k, gsdr_store& s,
                       int (*alloc)(gsdr_store&, void**, size_t)) {
  static const CodeTable kCodes;
  // the code table and the LDS limits are set once per device
  static bool ready[64] = {};
  if (c->device < 0 || c->device >= 64) return fail(GSDR_E_INVALID, "device index above 63");
  hipStream_t q = c->upload;
  const int64_t n = sorted.count();
  auto fail_hip = [](hipError_t e, const char* what) { return hip_fail(e, what); };
  hipError_t e;
  const size_t table_bytes = sizeof(uint32_t) << (2 * k - 1);
  const size_t plane_bytes = 2 * sizeof(uint32_t) * (GSDR_MAX_GAPFREE / 32 + 1);
  if (!ready[c->device]) {
    // k = 8 needs a 128 KiB table
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_code), kCodes.v, sizeof(kCodes.v))) != hipSuccess ||
        (e = hipFuncSetAttribute((const void*)prep_count_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)(sizeof(uint32_t) << 15))) != hipSuccess ||
        (e = hipFuncSetAttribute((const void*)prep_write_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)((sizeof(uint32_t) << 15) + plane_bytes))) != hipSuccess)
      return fail_hip(e, "prep setup");
    ready[c->device] = true;
  }
  // raw bytes and offsets, freed at the end
  uint8_t* d_data = nullptr;
  int64_t* d_off = nullptr;
  int32_t* d_dist = nullptr;
  const size_t nbytes = std::max<size_t>(sorted.data.size(), 1);
  if ((e = hipMalloc((void**)&d_data, nbytes)) != hipSuccess) return fail(GSDR_E_NOMEM, "hipMalloc raw bytes");
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(q);
    (void)hipFree(d_data);
    if (d_off) (void)hipFree(d_off);
    if (d_dist) (void)hipFree(d_dist);
  };
  int st = GSDR_OK;
  do {
    if ((e = hipMalloc((void**)&d_off, sizeof(int64_t) * (size_t)(n + 1))) != hipSuccess ||
        (e = hipMalloc((void**)&d_dist, sizeof(int32_t) * (size_t)std::max<int64_t>(n, 1))) != hipSuccess) {
      st = fail(GSDR_E_NOMEM, "hipMalloc prep buffers");
      break;
    }
    if ((st = alloc(s, (void**)&s.m_d, sizeof(int32_t) * (size_t)n)) ||
        (st = alloc(s, (void**)&s.comp, sizeof(uint4) * (size_t)n)) ||
        (st = alloc(s, (void**)&s.pkoff, sizeof(int64_t) * (size_t)(n + 1))) ||
        (st = alloc(s, (void**)&s.kpoff, sizeof(int64_t) * (size_t)(n + 1))))
      break;
    if ((e = hipMemcpyAsync(d_data, sorted.data.data(), sorted.data.size(), hipMemcpyHostToDevice, q)) != hipSuccess ||
        (e = hipMemcpyAsync(d_off, sorted.data_off.data(), sizeof(int64_t) * (size_t)(n + 1), hipMemcpyHostToDevice,
                            q)) != hipSuccess) {
      st = fail_hip(e, "upload raw bytes");
      break;
    }
    hipLaunchKernelGGL(prep_count_kernel, dim3((unsigned)n), dim3(kPrepThreads), table_bytes, q, d_data, d_off, k,
                       s.m_d, s.comp, d_dist);
    if ((e = hipGetLastError()) != hipSuccess) { st = fail_hip(e, "prep count launch"); break; }
    std::vector<int32_t> dist((size_t)n);
    s.m.assign((size_t)n, 0);
    if ((e = hipMemcpyAsync(s.m.data(), s.m_d, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, q)) != hipSuccess ||
        (e = hipMemcpyAsync(dist.data(), d_dist, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, q)) !=
            hipSuccess ||
        (e = hipStreamSynchronize(q)) != hipSuccess) {
      st = fail_hip(e, "prep counts");
      break;
    }
    std::vector<int64_t> pkoff((size_t)n + 1, 0), kpoff((size_t)n + 1, 0);
    s.raw_len.assign((size_t)n, 0);
    s.max_m = 0;
    for (int64_t i = 0; i < n; ++i) {
      pkoff[(size_t)i + 1] = pkoff[(size_t)i] + (s.m[(size_t)i] / 32 + 1);
      kpoff[(size_t)i + 1] = kpoff[(size_t)i] + dist[(size_t)i];
      s.raw_len[(size_t)i] = (int32_t)sorted.len(i);
      s.max_m = std::max(s.max_m, s.m[(size_t)i]);
    }
    if (s.max_m > GSDR_MAX_GAPFREE) {
      st = fail(GSDR_E_LIMIT, "sequence longer than 32767 gap-free bases");
      break;
    }
    s.kp_total = kpoff[(size_t)n];
    if ((st = alloc(s, (void**)&s.pk, sizeof(uint2) * (size_t)pkoff[(size_t)n])) ||
        (st = alloc(s, (void**)&s.kp, sizeof(uint32_t) * (size_t)std::max<int64_t>(kpoff[(size_t)n], 1))))
      break;
    if ((e = hipMemcpyAsync(s.pkoff, pkoff.data(), sizeof(int64_t) * (size_t)(n + 1), hipMemcpyHostToDevice, q)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(s.kpoff, kpoff.data(), sizeof(int64_t) * (size_t)(n + 1), hipMemcpyHostToDevice, q)) !=
            hipSuccess) {
      st = fail_hip(e, "upload offsets");
      break;
    }
    hipLaunchKernelGGL(prep_write_kernel, dim3((unsigned)n), dim3(kPrepThreads), table_bytes + plane_bytes, q,
                       d_data, d_off, k, s.pkoff, s.kpoff, s.pk, s.kp);
    if ((e = hipGetLastError()) != hipSuccess) { st = fail_hip(e, "prep write launch"); break; }
    // the host vectors above must outlive the copies
    if ((e = hipStreamSynchronize(q)) != hipSuccess) { st = fail_hip(e, "prep sync"); break; }
  } while (false);
  cleanup();
  return st;
}

}  // namespace gsdr
