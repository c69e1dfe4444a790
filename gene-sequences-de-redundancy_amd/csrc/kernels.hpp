// Device-side data structures shared by the kernels and the host driver.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace gsdr {

// Alignment tile shapes: S lanes of one wave per pair, W 32-bit text words
// per lane. A pair fits a shape when the S*W-word window covers the whole
// text, or when the sliding window ((S-1)*W words plus one lane of lag)
// covers the diagonal band of the pair. Widths between the powers of two
// keep the executed window close to the band: a row costs about W word
// updates per lane whatever part of the window the band uses.
constexpr int kNumShapes = 12;
struct Shape { int s, w; };
constexpr Shape kShapes[kNumShapes] = {{8, 1},  {16, 1}, {32, 1}, {64, 1},  {64, 2},  {64, 3},
                                       {64, 4}, {64, 6}, {64, 8}, {64, 10}, {64, 12}, {64, 16}};

// Tile-shape choice per bin of text length (gap-free text length < 512,
// < 2,048, < 8,192, longer): the shapes a pair held by its band may take
// (bit i = shape i; 0 = all) and how many allowed shapes past the narrowest
// one to go. A shape that holds the whole text is always allowed and always
// ends the search, so no choice goes past the text's whole-text shape (the
// bound the launches use). Every choice is exact: a wider shape also holds
// the band, and any window that holds the band gives the same decision.
constexpr int kLenBins = 4;
constexpr uint32_t kAllShapes = (1u << kNumShapes) - 1u;
struct ShapeTable {
  uint16_t mask[kLenBins];
  uint8_t bias[kLenBins];
};
__host__ __device__ inline int len_bin(int n) { return n < 512 ? 0 : n < 2048 ? 1 : n < 8192 ? 2 : 3; }

// Shape for a pair with text length n, query length m and minimum accepted
// LCS c (both gap-free lengths): the narrowest allowed shape that holds it,
// then `bias` allowed shapes further. E = n + m - 2c bounds the band width;
// the window slides at 32-row block boundaries, hence +31. A shape holds
// the pair when its S*W words cover the whole text, or when its sliding
// window ((S-1)*W words plus one lane of lag) covers the band; both sizes
// grow along kShapes, so every later shape holds it too.
__host__ __device__ inline int pick_shape(int n, int m, int c, int min_shape, uint32_t mask = kAllShapes,
                                          int bias = 0) {
  const int ntw = (n >> 5) + 1;
  const int e = n + m - 2 * c;
  const int need = ((e + 31) >> 5) + 1;
  if (mask == 0u) mask = kAllShapes;
  int left = -1;  // allowed shapes still to pass once the band is held
  for (int i = min_shape; i < kNumShapes; ++i) {
    const int s = kShapes[i].s, w = kShapes[i].w;
    if (s * w >= ntw) return i;
    if (!((mask >> i) & 1u)) continue;
    if (left < 0 && (s - 1) * w >= need) left = bias;
    if (left == 0) return i;
    if (left > 0) --left;
  }
  return kNumShapes - 1;
}

__host__ __device__ inline int pick_shape(int n, int m, int c, int min_shape, const ShapeTable& t) {
  const int b = len_bin(n);
  return pick_shape(n, m, c, min_shape, t.mask[b], t.bias[b]);
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

// Counters accumulated on the device.
enum StatSlot {
  ST_PAIRS = 0, ST_REJ_LEN, ST_REJ_COMP, ST_REJ_WORD, ST_ALIGNED, ST_ACCEPT, ST_REJ_ALIGN,
  ST_CELLS_EXEC, ST_CELLS_BAND, ST_CELLS_FULL, ST_FILTER_BYTES, ST_SPEC_WASTE, ST_CELLS_DECIDED,
  ST_OVERFLOW, ST_REJ_PIGEON, ST_PIGEON_CHECKED, ST_SCREENED, ST_NUM = 17
};
// The counters are striped: kStatStripes copies kStatStride words apart,
// each wave adding to the copy its block and wave pick, and the host sums
// the copies (thousands of waves adding to one word serialise on it at the
// end of every launch). Overflow is only ever added to copy 0.
constexpr int kStatStripes = 64, kStatStride = 32, kStatWords = kStatStripes * kStatStride;
static_assert(ST_NUM <= kStatStride, "one stripe holds every counter");

// The filter's view of one sequence, one 32-byte record (a candidate costs
// one cache line instead of one per array): gap-free length, cutoffs, the
// sum of its stored k-mer counts, composition and its profile's place.
struct __align__(16) SeqRec {
  int32_t m, bc, wc, ksum;
  uint32_t comp_ac, comp_gt;  // counts of A | C << 16, G | T << 16 (each <= 32,767)
  uint32_t kp0, kn;           // first profile entry and entries
};
static_assert(sizeof(SeqRec) == 32, "one 32-byte record per sequence");

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
  const int32_t* ksum;     // sum of the stored k-mer counts (the profile's total)
  const SeqRec* rec;       // the filter's per-sequence records
  const uint32_t* sig;     // word-presence signatures (kSigWords per sequence), or null
};

// Word-presence signature: bit sig_bucket(w) of kSigWords * 32 for every
// distinct k-mer w of a sequence with a nonzero stored count. Short words map
// one to one; longer ones through a multiplicative hash (Knuth's 2^32/phi).
// Each distinct word of q absent from r adds at least 1 to the shared-word
// deficit of the pair, and a bucket set in q's signature but not in r's holds
// at least one such word, so popcount(sig_q & ~sig_r) bounds the deficit from
// below. The bound only fires for long words, where the cutoff leaves little
// slack (t >= 0.94): shorter words fill the signatures of random sequences.
constexpr int kSigWords = 64;
constexpr int kScreenMinK = 7;
__host__ __device__ inline uint32_t sig_bucket(uint32_t w, int k) {
  return k <= 5 ? w : (w * 0x9E3779B1u) >> 21;
}

// Speculative representatives per step are bounded by the accept-mask width.
constexpr int kMaxSpec = 64;

// Split list: phase-A pairs (speculative representatives against each
// other) with many rows, each aligned by two waves of one block, the first
// half of the rows forward and the second half backward on the reversed
// pair, joined at the middle row (align_split_kernel). It sits after the
// kNumShapes shape lists; a job holds at most kMaxSpec (kMaxSpec - 1) / 2
// phase-A pairs, so it never overflows. An item's slot field carries the
// tile shape its band picked (the widest window the halves start from).
constexpr int kSplitList = kNumShapes;
constexpr int kSplitPerJob = kMaxSpec * (kMaxSpec - 1) / 2;
constexpr int kSplitShapeShift = 12;
// Workspace counters: the item count of every list (shape lists, then the
// split list), then the work counter of every list.
constexpr int kWorkCounter0 = kNumShapes + 1;
constexpr int kNumCounters = 2 * (kNumShapes + 1);
// Candidate positions per resolve tile (256 threads, 8 each).
constexpr int kResolveTile = 2048;
// Candidates per filter work unit (one per 16-lane group of a 256-thread block).
constexpr int kFilterChunk = 16;
// With one table per block (long words) a unit takes four candidates per
// 16-lane group, whose record loads are issued together.
__host__ __device__ constexpr int filter_subs(int group) { return group <= 4 ? 4 : 1; }
__host__ __device__ constexpr int filter_chunk(int group) { return kFilterChunk * filter_subs(group); }

// Per-job round state on the device. The candidate list cand[cur][0..nc)
// holds the unclustered sorted indices in ascending order; its first k_spec
// entries are this step's (speculative) representatives.
struct JobState {
  int32_t cur;      // current candidate buffer
  int32_t nc;       // entries in the current list
  int32_t k_spec;   // speculative representatives this step
  int32_t done;     // 1 when every sequence has a cluster
  int32_t steps;    // steps taken
  int32_t reps;     // representatives fixed so far
  int32_t k_max;    // upper bound for k_spec (1 = plain form)
  int32_t first;    // cand[cur][0], the next representative (-1 when done)
  int32_t kt;       // true representatives among the k_spec slots (after phase A)
  int32_t nonrep;   // speculative slots that an earlier slot absorbed (phase A)
  int32_t absorbed; // candidates absorbed in phase B (summed over the resolve tiles)
  int32_t kept;     // candidates left after the step (from the last resolve tile)
  unsigned long long tmask;  // slots that are true representatives (after phase A)
  unsigned long long pad2;
};

// Filter phases of a step: A = speculative representatives against each
// other (slots 0..k_spec-1), B = true representatives against the rest.
enum FilterPhase { PHASE_REPS = 0, PHASE_MAIN = 1 };

// Work units of a job in a phase: (group of up to `group` representative
// slots, filter_chunk(group) candidate positions). Phase A: slots 0..K-1 against
// positions 0..K-1; phase B: the kt true representatives against positions
// K..nc-1.
__host__ __device__ inline int phase_units(const JobState& s, int phase, int group) {
  if (s.done) return 0;
  const int K = s.k_spec;
  if (phase == PHASE_REPS)
    return K > 1 ? ((K + group - 1) / group) * ((K + filter_chunk(group) - 1) / filter_chunk(group)) : 0;
  const int rest = s.nc - K;
  return rest > 0 ? ((s.kt + group - 1) / group) * ((rest + filter_chunk(group) - 1) / filter_chunk(group)) : 0;
}

// Representatives whose shared-word tables a filter block holds at once:
// as many as fit 32 KiB of LDS (u16 counts over 4^k words), at most 16.
__host__ __device__ constexpr int filter_group(int k) {
  const int bytes = 2 << (2 * k);
  int g = 32768 / bytes;
  if (g >= 16) return 16;
  if (g >= 8) return 8;
  if (g >= 4) return 4;
  if (g >= 2) return 2;
  return 1;
}

// One clustering job as the kernels see it.
struct JobDev {
  StoreView st;
  int32_t* cluster;          // n entries, -1 until assigned
  int32_t* cand[2];          // candidate lists (ping-pong)
  unsigned long long* acc;   // n accept masks: bit i = P(rep slot i, q)
  JobState* state;
  // signature screen (jobs whose store has signatures): per sequence, the
  // step's slots whose representative passed the screen against it; per
  // phase-B chunk of filter_chunk positions, the OR of those masks
  unsigned long long* scr;
  unsigned long long* chm;
  int32_t* tcnt;             // resolve: survivors per tile of kResolveTile positions
  int32_t n;
  int32_t k;                 // word length of the stored profiles
  int32_t flags;             // kJobDropFilters
};

// Job flag: the cuda-version's drop mode (cuda-version/func.cu:1068-1097):
// no composition or shared-word filter, every pair past the length test is
// aligned. Opt-in; clusters then differ from the root reference's.
constexpr int kJobDropFilters = 1;

// Alignment work item: text and query sorted indices, the job and the
// representative's slot in this step. Normally text = representative and
// query = candidate. LCS is symmetric, so when the representative is at
// least kSwapRatio times longer the roles swap (flag kSwapFlag in slot): the
// bit-vector spans the short candidate and rows run over the long
// representative, where the exact early accept fires after about 4m rows
// instead of costing n/32 words per row. The cutoff always belongs to the
// candidate.
struct AlignItem {
  int32_t text, query, job, slot;
};
constexpr int kSwapFlag = 256;
// Set in AlignItem.slot by the pigeonhole prefilter: the pair cannot reach
// the cutoff and the alignment kernels skip it.
constexpr int kPigeonFlag = 512;
constexpr int kSwapRatio = 4;  // default; a tuned context may choose another (0 = never)

__host__ __device__ inline bool swap_roles(int n_rep, int m_cand, int ratio = kSwapRatio) {
  return ratio > 0 && (long long)n_rep >= (long long)ratio * m_cand;
}

}  // namespace gsdr
