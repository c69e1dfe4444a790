// Device-side data structures shared by the kernels and the host driver.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace gsdr {

// Alignment tile shapes: S lanes of one wave per pair, W 32-bit text words
// per lane. A pair fits a shape when the S*W-word window covers the whole
// text, or when the sliding window ((S-1)*W words plus one lane of lag)
// covers the diagonal band of the pair.
constexpr int kNumShapes = 8;
struct Shape { int s, w; };
constexpr Shape kShapes[kNumShapes] = {{8, 1}, {16, 1}, {32, 1}, {64, 1},
                                       {64, 2}, {64, 4}, {64, 8}, {64, 16}};

// Smallest shape that holds a pair with text length n, query length m and
// minimum accepted LCS c (both gap-free lengths). E = n + m - 2c bounds the
// band width; the window slides at 32-row block boundaries, hence +31.
__host__ __device__ inline int pick_shape(int n, int m, int c, int min_shape) {
  const int ntw = (n >> 5) + 1;
  const int e = n + m - 2 * c;
  const int need = ((e + 31) >> 5) + 1;
  for (int i = min_shape; i < kNumShapes; ++i) {
    const int s = kShapes[i].s, w = kShapes[i].w;
    if (s * w >= ntw || (s - 1) * w >= need) return i;
  }
  return kNumShapes - 1;
}

// Cells of the exact decision band: rows j = 1..m, text columns
// max(1, j-(m-c)) .. min(n, j+n-c); equals c(n-c) + n(m-c) + c for
// 1 <= c <= min(n, m).
__host__ __device__ inline double band_cells(int n, int m, int c) {
  if (c < 1 || c > n || c > m) return 0.0;
  return (double)c * (double)(n - c) + (double)n * (double)(m - c) + (double)c;
}

// Band cells of rows 1..R only (R <= m): the band work an exact decision
// taken after R query rows needs.
__host__ __device__ inline double band_cells_rows(int n, int m, int c, int R) {
  if (c < 1 || c > n || c > m || R <= 0) return 0.0;
  if (R > m) R = m;
  const double r = R;
  // sum of hi(j) = min(n, j + n - c)
  const double rc = R < c ? R : c;
  double hi = rc * (rc + 1) / 2.0 + rc * (double)(n - c) + (R > c ? (double)(R - c) * n : 0.0);
  // sum of lo(j) = max(1, j - (m - c))
  const int d = m - c;  // rows with lo = 1: j <= d + 1
  double lo;
  if (R <= d + 1) {
    lo = r;
  } else {
    const double a = d + 1;                   // rows with lo == 1
    const double u0 = 2.0, u1 = (double)(R - d);  // lo values of the other rows
    lo = a + (u0 + u1) * (u1 - u0 + 1) / 2.0;
  }
  return hi - lo + r;
}

// Per-round state kept on the device (plain form).
struct RoundState {
  int32_t rep;   // current representative, -1 when done
  int32_t nc;    // candidates after rep
  int32_t pad[2];
};

// Counters accumulated on the device.
enum StatSlot {
  ST_PAIRS = 0, ST_REJ_LEN, ST_REJ_COMP, ST_REJ_WORD, ST_ALIGNED, ST_ACCEPT, ST_REJ_ALIGN,
  ST_CELLS_EXEC, ST_CELLS_BAND, ST_CELLS_FULL, ST_FILTER_BYTES, ST_SPEC_WASTE, ST_CELLS_DECIDED,
  ST_NUM = 16
};

// Store arrays resident in HBM (indexed by sorted sequence index).
struct StoreView {
  const uint2* pk;         // (lo plane, hi plane) per 32 gap-free bases
  const int64_t* pkoff;    // first packed word of each sequence
  const int32_t* m;        // gap-free length
  const uint4* comp;       // counts of A, C, G, T
  const uint32_t* kp;      // distinct k-mers: word | count << 16
  const int64_t* kpoff;    // n+1 offsets into kp
  const int32_t* bc;       // ceil(L*t) (func.cpp:368)
  const int32_t* wc;       // shared-word cutoff (func.cpp:353-366)
};

}  // namespace gsdr
