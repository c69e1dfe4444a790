// CPU oracle for the exact greedy de-redundancy path.
//
// TEST INFRASTRUCTURE ONLY. This file restates, on the CPU and independently
// of the product code, the canonical behaviour of the reference program
// (JuZhenCS/gene-sequences-de-redundancy, root variant: main.cpp, func.h,
// func.cpp). Only tests, the smoke check and the CPU-baseline leg of the
// benchmark may load it; the product never links or calls it.
//
// Pinning (what checks each function, see tests/test_oracle.py):
//   orc_kmap, orc_cutoffs   exhaustive tables vs exact rational arithmetic and
//                           the float32 quirk counts; k boundaries 0.88/0.94
//   orc_summin              brute-force multiset intersection (Python Counter)
//   orc_shared_words_pass   early stop == full scan at wc = sum-1, sum, sum+1
//   orc_lcs                 textbook O(nm) LCS DP; random/adversarial lengths
//   orc_decide_fast         == (orc_lcs > bc) on random + near-threshold pairs
//   orc_sort_order          std::sort over records holding std::string data
//   read_fasta              hand-built edge-case files (CRLF, blank lines,
//                           no trailing newline, header-only records, IUPAC)
//   greedy driver           brute-force recursion (Python) on small inputs and
//                           the idempotence known answer on the reference's
//                           shipped result.fasta (subset fixture); its drop
//                           mode against the same recursion without F1/F2
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>
#include <omp.h>

namespace {

struct Read {
  std::string name;
  std::string data;
};

// func.cpp:155-188.
inline uint8_t code(unsigned char c) {
  switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': case 'U': case 'u': return 3;
    default: return 4;
  }
}

// func.cpp:70-104 (getline loop; on a well-formed file with a trailing
// newline this is the reference's record split; where the reference never
// terminates this one stops at end of file).
bool read_fasta(const char* path, std::vector<Read>& reads) {
  std::ifstream file(path, std::ios::binary);
  if (!file) return false;
  std::string line;
  if (!std::getline(file, line)) return true;  // empty file: no records
  Read r;
  r.name = line;
  while (std::getline(file, line)) {
    if (!line.empty() && line[0] == '>') {
      reads.push_back(r);
      r.name = line;
      r.data.clear();
    } else {
      r.data += line;
    }
  }
  reads.push_back(r);
  return true;
}

// func.cpp:36-45.
int kmap(float t) {
  if (t < 0.88) return 4;
  else if (t < 0.94) return 5;
  else if (t < 0.97) return 6;
  return 7;
}

// func.cpp:353-369, in float32.
void cutoffs(int length, float t, int k, int* bc, int* wc) {
  int required = length - k + 1;
  float om = 1.0f - t;
  float prod = (float)length * om;
  float ce = std::ceil(prod);
  float sc = ce * (float)k;
  int cutoff = (int)sc;
  required -= cutoff;
  required = std::max(required, 1);
  *wc = required;
  float p2 = (float)length * t;
  *bc = (int)std::ceil(p2);
}

// Per-sequence data after preprocessing (func.cpp:197-446).
struct Seq {
  int L = 0;                 // raw length
  int m = 0;                 // gap-free length
  int comp[4] = {0, 0, 0, 0};
  std::vector<uint64_t> lo, hi;          // gap-free bit planes, 64 bases/word
  std::vector<std::pair<uint16_t, uint16_t>> words;  // (word, count) distinct
  int total = 0;             // sum of the stored counts
  int bc = 0, wc = 0;
};

void prepare(const std::string& data, int k, float t, Seq& s) {
  const int L = (int)data.size();
  s.L = L;
  std::vector<uint8_t> c(L);
  for (int i = 0; i < L; ++i) c[i] = code((unsigned char)data[i]);
  // createPrefix func.cpp:207-232
  for (int i = 0; i < L; ++i) if (c[i] < 4) s.comp[c[i]]++;
  // createWords func.cpp:252-270: raw windows ending at i, no code 4 inside.
  std::vector<uint16_t> w;
  for (int i = k - 1; i < L; ++i) {
    unsigned word = 0;
    bool flag = false;
    for (int j = 0; j < k; ++j) {
      unsigned b = c[i - j];
      word += b << (2 * j);
      if (b == 4) flag = true;
    }
    if (!flag) w.push_back((uint16_t)word);
  }
  // sortWords / mergeWords func.cpp:276-338: sorted runs, count per run
  // (unsigned short, so counts wrap modulo 2^16).
  std::sort(w.begin(), w.end());
  for (size_t i = 0; i < w.size();) {
    size_t j = i;
    while (j < w.size() && w[j] == w[i]) ++j;
    s.words.emplace_back(w[i], (uint16_t)(j - i));
    s.total += (uint16_t)(j - i);
    i = j;
  }
  // deleteGap + compressData func.cpp:375-446 (64 bases per word here).
  std::vector<uint8_t> y;
  y.reserve(L);
  for (int i = 0; i < L; ++i) if (c[i] < 4) y.push_back(c[i]);
  s.m = (int)y.size();
  int nw = s.m / 64 + 1;
  s.lo.assign(nw, 0);
  s.hi.assign(nw, 0);
  for (int i = 0; i < s.m; ++i) {
    s.lo[i >> 6] |= (uint64_t)(y[i] & 1) << (i & 63);
    s.hi[i >> 6] |= (uint64_t)(y[i] >> 1) << (i & 63);
  }
  cutoffs(L, t, k, &s.bc, &s.wc);
}

inline uint64_t match_mask(uint64_t tl, uint64_t th, int qc) {
  uint64_t ql = (qc & 1) ? ~0ull : 0ull;
  uint64_t qh = (qc & 2) ? ~0ull : 0ull;
  return ~(tl ^ ql) & ~(th ^ qh);
}
inline int query_base(const Seq& q, int j) {
  return (int)(((q.lo[j >> 6] >> (j & 63)) & 1) | (((q.hi[j >> 6] >> (j & 63)) & 1) << 1));
}

// kernel_dynamic func.cpp:572-669: bit-vector LCS of text (representative)
// and query, V starting all ones, per query base U = V & M,
// V' = (V + U) | (V & ~M) with the carry rippling from low to high text
// words; LCS = zeros among the first n bits. The reference blocks 32 query
// bases per text word with a carry per query base; this blocks 64.
int lcs(const Seq& text, const Seq& query) {
  const int n = text.m, m = query.m;
  const int tw = n / 64 + 1;
  std::vector<uint64_t> line(tw, ~0ull);
  for (int i0 = 0; i0 < m; i0 += 64) {
    const int kb = std::min(64, m - i0);
    uint64_t column[64];
    int qb[64];
    for (int k = 0; k < kb; ++k) { column[k] = 0; qb[k] = query_base(query, i0 + k); }
    for (int j = 0; j < tw; ++j) {
      const uint64_t tl = text.lo[j], th = text.hi[j];
      uint64_t row = line[j];
      for (int k = 0; k < kb; ++k) {
        uint64_t match = match_mask(tl, th, qb[k]);
        uint64_t u = row & match;
        uint64_t keep = row & ~match;
        uint64_t s = row + column[k];
        uint64_t c1 = s < row;
        s += u;
        uint64_t c2 = s < u;
        column[k] = c1 | c2;
        row = s | keep;
      }
      line[j] = row;
    }
  }
  int sum = 0;
  for (int i = 0; i < n; ++i) sum += !((line[i >> 6] >> (i & 63)) & 1);
  return sum;
}

// Exact decision LCS > bc computed on a diagonal band with early exit.
// With c = bc + 1, every common subsequence of length >= c uses only matches
// (i, j) with j - (m - c) <= i <= j + (n - c) (i text, j query). Running the
// same bit-vector recurrence with matches restricted to 64-bit text words
// that overlap that band gives a value v with LCS_band <= v <= LCS, hence
// [v >= c] == [LCS >= c]. Words left of the band keep their value (carry in
// 0, no match), words right of it stay all ones, so only band words are
// updated. Since v grows by at most one per query base, the run stops as
// soon as v >= c or v + (remaining bases) < c.
bool decide_fast(const Seq& text, const Seq& query, int bc, int64_t* cells) {
  const int n = text.m, m = query.m;
  const int c = bc + 1;
  if (c > std::min(n, m)) return false;
  if (c <= 0) return true;
  const int tw = n / 64 + 1;
  std::vector<uint64_t> line(tw, ~0ull);
  const uint64_t lastmask = (n % 64) ? ((1ull << (n % 64)) - 1) : 0ull;
  auto zeros = [&]() {
    int z = 0;
    for (int w = 0; w < n / 64; ++w) z += __builtin_popcountll(~line[w]);
    if (n % 64) z += __builtin_popcountll(~line[n / 64] & lastmask);
    return z;
  };
  for (int j = 0; j < m; ++j) {
    int lo = std::max(0, j - (m - c));
    int hi = std::min(n - 1, j + (n - c));
    int wlo = lo >> 6, whi = hi >> 6;
    const int qc = query_base(query, j);
    uint64_t carry = 0;
    for (int w = wlo; w <= whi; ++w) {
      uint64_t row = line[w];
      uint64_t match = match_mask(text.lo[w], text.hi[w], qc);
      uint64_t u = row & match;
      uint64_t s = row + carry;
      uint64_t c1 = s < row;
      s += u;
      uint64_t c2 = s < u;
      carry = c1 | c2;
      line[w] = s | (row & ~match);
    }
    if (cells) *cells += (int64_t)(whi - wlo + 1) * 64;
    if ((j & 63) == 63 || j == m - 1) {
      int z = zeros();
      if (z >= c) return true;
      if (z + (m - 1 - j) < c) return false;
    }
  }
  return zeros() >= c;
}

// kernel_filter func.cpp:552-567: sum over q's distinct words of
// min(table[w], count) >= wc_q. With `early`, the scan stops as soon as the
// deficit (sum of count - min) exceeds total - wc: the sum over the words
// still unread can then not lift the shared count to wc (exact; the decision
// equals the full scan's, tests/test_oracle.py compares the two).
bool shared_words_pass(const std::vector<uint16_t>& table, const Seq& s, bool early) {
  if (!early) {
    int shared = 0;
    for (auto& e : s.words) shared += std::min(table[e.first], e.second);
    return shared >= s.wc;
  }
  const int slack = s.total - s.wc;
  if (slack < 0) return false;
  int deficit = 0;
  for (auto& e : s.words) {
    deficit += e.second - std::min(table[e.first], e.second);
    if (deficit > slack) return false;
  }
  return true;
}

struct Stats {
  int64_t reps = 0, pairs = 0, rej_comp = 0, rej_word = 0, aligned = 0, accepted = 0;
  double seconds = 0;
};

// clustering func.cpp:677-784 with the per-round kernels func.cpp:499-676.
// drop: the cuda-version's drop mode (cuda-version/func.cu:1068-1097), where
// each round aligns every remaining sequence without the composition and
// shared-word filters; the rest of the semantics stays root's.
void cluster_seqs(const std::vector<Seq>& seqs, int k, int threads, bool fast, bool drop,
                  std::vector<int32_t>& cluster, Stats& st, int64_t max_rounds) {
  const int n = (int)seqs.size();
  cluster.assign(n, -1);
  std::vector<int32_t> remain(n);
  for (int i = 0; i < n; ++i) remain[i] = i;
  std::vector<uint16_t> table((size_t)1 << (2 * k), 0);
  int rep = -1;
  if (threads > 0) omp_set_num_threads(threads);
  int64_t rounds = 0;
  auto t0 = std::chrono::steady_clock::now();
  while (true) {
    // updateRepresentative func.cpp:466-476
    rep += 1;
    while (rep < n && cluster[rep] != -1) rep += 1;
    if (rep >= n) break;
    if (max_rounds >= 0 && rounds >= max_rounds) break;
    ++rounds;
    cluster[rep] = rep;
    st.reps++;
    // updateRemain func.cpp:477-487
    size_t cnt = 0;
    for (size_t i = 0; i < remain.size(); ++i)
      if (cluster[remain[i]] == -1) remain[cnt++] = remain[i];
    remain.resize(cnt);
    const Seq& r = seqs[rep];
    for (auto& e : r.words) table[e.first] = e.second;  // kernel_makeTable
    int64_t rc = 0, rw = 0, al = 0, ac = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : rc, rw, al, ac)
    for (long long idx = 0; idx < (long long)remain.size(); ++idx) {
      const int q = remain[idx];
      const Seq& s = seqs[q];
      if (!drop) {
        // kernel_preFilter func.cpp:525-543
        int sum = 0;
        for (int b = 0; b < 4; ++b) sum += std::min(r.comp[b], s.comp[b]);
        if (sum < s.bc) { rc++; continue; }
        // kernel_filter func.cpp:544-571
        if (!shared_words_pass(table, s, fast)) { rw++; continue; }
      }
      // kernel_dynamic func.cpp:658-675
      al++;
      bool acc = fast ? decide_fast(r, s, s.bc, nullptr) : (lcs(r, s) > s.bc);
      if (acc) { cluster[q] = rep; ac++; }
    }
    for (auto& e : r.words) table[e.first] = 0;  // kernel_cleanTable
    st.pairs += (int64_t)remain.size();
    st.rej_comp += rc; st.rej_word += rw; st.aligned += al; st.accepted += ac;
  }
  st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

Seq seq_from_codes(const uint8_t* a, int n) {
  // codes 0..4 mapped back to letters so prepare() runs the same path
  static const char letters[5] = {'A', 'C', 'G', 'T', 'N'};
  std::string d(n, 'A');
  for (int i = 0; i < n; ++i) d[i] = letters[a[i] > 4 ? 4 : a[i]];
  Seq s;
  prepare(d, 4, 0.95f, s);
  return s;
}

}  // namespace

extern "C" {

int orc_kmap(float t) { return kmap(t); }

void orc_cutoffs(int length, float t, int k, int* bc, int* wc) { cutoffs(length, t, k, bc, wc); }

// LCS of two code strings (codes 0..3 kept, 4 removed as gaps).
int orc_lcs(const uint8_t* a, int n, const uint8_t* b, int m) {
  Seq x = seq_from_codes(a, n), y = seq_from_codes(b, m);
  return lcs(x, y);
}

int orc_decide_fast(const uint8_t* a, int n, const uint8_t* b, int m, int bc, int64_t* cells) {
  Seq x = seq_from_codes(a, n), y = seq_from_codes(b, m);
  return decide_fast(x, y, bc, cells) ? 1 : 0;
}

// Sum over k-mers w of min(c_a(w), c_b(w)) for raw code strings (windows
// containing a code 4 are skipped).
long long orc_summin(const uint8_t* a, int n, const uint8_t* b, int m, int k) {
  static const char letters[5] = {'A', 'C', 'G', 'T', 'N'};
  std::string da(n, 'A'), db(m, 'A');
  for (int i = 0; i < n; ++i) da[i] = letters[a[i] > 4 ? 4 : a[i]];
  for (int i = 0; i < m; ++i) db[i] = letters[b[i] > 4 ? 4 : b[i]];
  Seq x, y;
  prepare(da, k, 0.95f, x);
  prepare(db, k, 0.95f, y);
  std::vector<uint16_t> table((size_t)1 << (2 * k), 0);
  for (auto& e : x.words) table[e.first] = e.second;
  long long s = 0;
  for (auto& e : y.words) s += std::min(table[e.first], e.second);
  return s;
}

// Shared-word test of one pair (raw code strings; q's cutoff wc given):
// early = 0 scans every word, 1 stops at the exact deficit bound.
int orc_shared_words_pass(const uint8_t* a, int n, const uint8_t* b, int m, int k, int wc, int early) {
  static const char letters[5] = {'A', 'C', 'G', 'T', 'N'};
  std::string da(n, 'A'), db(m, 'A');
  for (int i = 0; i < n; ++i) da[i] = letters[a[i] > 4 ? 4 : a[i]];
  for (int i = 0; i < m; ++i) db[i] = letters[b[i] > 4 ? 4 : b[i]];
  Seq x, y;
  prepare(da, k, 0.95f, x);
  prepare(db, k, 0.95f, y);
  y.wc = wc;
  std::vector<uint16_t> table((size_t)1 << (2 * k), 0);
  for (auto& e : x.words) table[e.first] = e.second;
  return shared_words_pass(table, y, early != 0) ? 1 : 0;
}

// Length order: perm[s] = original index of the s-th sorted record, by
// std::sort over records that hold their data as std::string
// (func.cpp:105-107).
void orc_sort_order(const int64_t* lens, int64_t n, int64_t* perm) {
  struct R { std::string data; int64_t idx; };
  std::vector<R> v(n);
  for (int64_t i = 0; i < n; ++i) { v[i].data.assign((size_t)lens[i], 'A'); v[i].idx = i; }
  std::sort(v.begin(), v.end(), [](R& a, R& b) { return a.data.size() > b.data.size(); });
  for (int64_t i = 0; i < n; ++i) perm[i] = v[i].idx;
}

// Whole pipeline on a FASTA file: read, sort, preprocess, cluster, write.
// cluster_out (may be null) receives cluster[] over sorted indices (n_cap
// entries at most); perm_out (may be null) the original index per sorted
// index. max_rounds < 0 runs to completion. Returns the record count or a
// negative value on error.
long long orc_cluster_file(const char* in_path, const char* out_path, float t, int k, int threads,
                           int fast, int32_t* cluster_out, int64_t* perm_out, long long n_cap,
                           long long max_rounds, double* stats_out, int drop) {
  std::vector<Read> reads;
  if (!read_fasta(in_path, reads)) return -1;
  if (reads.empty()) return -2;
  if (k == 0) k = kmap(t);
  const int64_t n = (int64_t)reads.size();
  std::vector<int64_t> lens(n), perm(n);
  for (int64_t i = 0; i < n; ++i) lens[i] = (int64_t)reads[i].data.size();
  orc_sort_order(lens.data(), n, perm.data());
  std::vector<Read> sorted(n);
  for (int64_t i = 0; i < n; ++i) sorted[i] = std::move(reads[perm[i]]);
  std::vector<Seq> seqs(n);
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 64)
  for (long long i = 0; i < (long long)n; ++i) prepare(sorted[i].data, k, t, seqs[i]);
  std::vector<int32_t> cluster;
  Stats st;
  cluster_seqs(seqs, k, threads, fast != 0, drop != 0, cluster, st, max_rounds);
  if (out_path && out_path[0]) {
    std::ofstream f(out_path, std::ios::binary);
    for (int64_t s = 0; s < n; ++s)
      if (cluster[s] == s) f << sorted[s].name << '\n' << sorted[s].data << '\n';
  }
  if (cluster_out) for (int64_t s = 0; s < std::min<int64_t>(n, n_cap); ++s) cluster_out[s] = cluster[s];
  if (perm_out) for (int64_t s = 0; s < std::min<int64_t>(n, n_cap); ++s) perm_out[s] = perm[s];
  if (stats_out) {
    stats_out[0] = (double)st.reps; stats_out[1] = (double)st.pairs; stats_out[2] = (double)st.rej_comp;
    stats_out[3] = (double)st.rej_word; stats_out[4] = (double)st.aligned; stats_out[5] = (double)st.accepted;
    stats_out[6] = st.seconds;
  }
  return (long long)n;
}

// Reads a FASTA file and reports the record count and total data bytes;
// names/data are concatenated into caller buffers when given (with their
// offsets), so tests can compare record splits.
long long orc_read_fasta(const char* path, char* names, int64_t* name_off, char* data, int64_t* data_off,
                         long long cap_records, long long cap_bytes, long long* total_bytes) {
  std::vector<Read> reads;
  if (!read_fasta(path, reads)) return -1;
  long long nb = 0, db = 0;
  for (auto& r : reads) { nb += (long long)r.name.size(); db += (long long)r.data.size(); }
  if (total_bytes) *total_bytes = db;
  if (names && data && (long long)reads.size() <= cap_records && nb <= cap_bytes && db <= cap_bytes) {
    long long a = 0, b = 0;
    name_off[0] = 0; data_off[0] = 0;
    for (size_t i = 0; i < reads.size(); ++i) {
      memcpy(names + a, reads[i].name.data(), reads[i].name.size()); a += (long long)reads[i].name.size();
      memcpy(data + b, reads[i].data.data(), reads[i].data.size()); b += (long long)reads[i].data.size();
      name_off[i + 1] = a; data_off[i + 1] = b;
    }
  }
  return (long long)reads.size();
}

}  // extern "C"

#ifdef ORC_MAIN
// orc_cluster <in.fasta> <out.fasta> <t> [k] [threads] [fast] [cluster.bin]
int main(int argc, char** argv) {
  if (argc < 4) { fprintf(stderr, "usage: orc_cluster in out t [k] [threads] [fast] [cluster.bin]\n"); return 2; }
  float t = std::stof(argv[3]);
  int k = argc > 4 ? atoi(argv[4]) : 0;
  int threads = argc > 5 ? atoi(argv[5]) : 0;
  int fast = argc > 6 ? atoi(argv[6]) : 1;
  std::vector<int32_t> cl(1 << 26);
  double st[8] = {0};
  long long n = orc_cluster_file(argv[1], argv[2], t, k, threads, fast, cl.data(), nullptr, (long long)cl.size(), -1, st, 0);
  if (n < 0) { fprintf(stderr, "error %lld\n", n); return 1; }
  if (argc > 7) { FILE* f = fopen(argv[7], "wb"); fwrite(cl.data(), 4, (size_t)n, f); fclose(f); }
  printf("{\"n\": %lld, \"reps\": %.0f, \"pairs\": %.0f, \"rej_comp\": %.0f, \"rej_word\": %.0f, \"aligned\": %.0f, \"accepted\": %.0f, \"seconds\": %.3f}\n",
         n, st[0], st[1], st[2], st[3], st[4], st[5], st[6]);
  return 0;
}
#endif
