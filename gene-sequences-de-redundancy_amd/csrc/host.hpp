// Host-side data model of the de-redundancy path: FASTA records, the
// reference's length order, per-sequence preprocessing (2-bit packing,
// composition, k-mer profiles, cutoffs) and the result writer.
//
// Everything here is plain C++ (no device code) and is compiled into the
// same shared library as the HIP kernels.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace gsdr {

// One FASTA record as the reference stores it (func.h:17-20): the header line
// verbatim (including '>') and the raw concatenation of the following lines.
// Large host buffers: 2 MiB aligned and marked for huge pages, so that
// filling them costs few page faults.
void* big_alloc(size_t bytes);
void big_free(void* p, size_t bytes) noexcept;

// Byte buffer whose resize leaves new bytes uninitialised, so that large
// buffers are first touched by the threads that fill them.
template <class T>
struct NoInitAlloc {
  using value_type = T;
  NoInitAlloc() = default;
  template <class U> NoInitAlloc(const NoInitAlloc<U>&) noexcept {}
  T* allocate(size_t n) { return (T*)big_alloc(n * sizeof(T)); }
  void deallocate(T* p, size_t n) noexcept { big_free(p, n * sizeof(T)); }
  template <class U> void construct(U* p) noexcept { ::new ((void*)p) U; }
  template <class U, class... A> void construct(U* p, A&&... a) { ::new ((void*)p) U(std::forward<A>(a)...); }
  template <class U> bool operator==(const NoInitAlloc<U>&) const noexcept { return true; }
  template <class U> bool operator!=(const NoInitAlloc<U>&) const noexcept { return false; }
};
using Bytes = std::vector<char, NoInitAlloc<char>>;
// Per-record offsets: sized without zeroing, first touched by the threads
// that fill them.
using Index = std::vector<int64_t, NoInitAlloc<int64_t>>;

struct Records {
  Bytes names;                      // concatenated header bytes
  Index name_off;                   // size n+1
  Bytes data;                       // concatenated raw sequence bytes
  Index data_off;                   // size n+1
  // Set once the records are in the reference's length order; `order` then
  // holds the input position of each record.
  bool length_ordered = false;
  std::vector<int64_t> order;
  int64_t count() const { return (int64_t)data_off.size() - 1; }
  int64_t len(int64_t i) const { return data_off[i + 1] - data_off[i]; }
  void clear() {
    names.clear(); name_off.assign(1, 0); data.clear(); data_off.assign(1, 0);
    length_ordered = false; order.clear();
  }
  void push(const char* nm, int64_t nl, const char* d, int64_t dl) {
    names.insert(names.end(), nm, nm + nl); name_off.push_back((int64_t)names.size());
    data.insert(data.end(), d, d + dl); data_off.push_back((int64_t)data.size());
  }
  Records() { clear(); }
};

// Reads a FASTA file with the reference's record semantics (func.cpp:70-104)
// on `threads` threads. With `length_order` the records come out in the
// reference's length order (func.cpp:105-107) with `order` set, in the same
// pass that copies them. Returns 0 on success, a gsdr status code otherwise;
// `err` gets a message.
int read_fasta(const char* path, Records& out, std::string& err, int threads = 1, bool length_order = false);

// Stable-free length order of the reference (func.cpp:105-107): libstdc++
// std::sort with "longer first", on `threads` threads with the same result.
// Returns perm with perm[sorted] = original.
std::vector<int64_t> length_order(const Records& r, int threads = 1);
Records permute(const Records& r, const std::vector<int64_t>& perm, int threads = 1);

// Writes representatives (cluster[s] == s) in sorted order, unwrapped
// (func.cpp:785-796). Returns the number written, or -1 on I/O error.
int64_t write_result(const Records& sorted, const int32_t* cluster, const char* path);

// FASTA writer with fixed line width (generator output).
int write_fasta(const Records& r, const char* path, int width);

// Word length chosen from the threshold (func.cpp:36-45); float t is
// compared against double literals exactly as the reference does.
int auto_word_length(float t);
// Cutoffs (func.cpp:353-369), evaluated in float32 like the reference.
int base_cutoff(int32_t raw_len, float t);
int word_cutoff(int32_t raw_len, float t, int k);

// Per-sequence preprocessing for one word length k. Arrays are indexed by
// the sorted index s.
struct Prep {
  int32_t n = 0;
  int k = 0;
  std::vector<int32_t> raw_len;     // L_s
  std::vector<int32_t> m;           // gap-free length
  std::vector<uint32_t> comp;       // 4 per sequence: counts of codes 0..3
  std::vector<int64_t> pk_off;      // offset in 64-bit packed words (uint2)
  std::vector<uint32_t> pk;         // pairs (lo plane, hi plane) per 32 bases
  std::vector<int64_t> kp_off;      // size n+1, offset into kp
  std::vector<uint32_t> kp;         // (word | count << 16) per distinct k-mer
  int32_t max_m = 0;
};
// Builds Prep from sorted records using `threads` host threads.
void preprocess(const Records& sorted, int k, int threads, Prep& p);

// Synthetic generator (seeded, deterministic). law: 1 = 500 parents with 19
// mutated children each, 2 = log-normal-length parents with 10 children each,
// 4 = Pareto(1.5) lengths with singletons and families (values match
// GSDR_LAW_* in gsdr.h); n_override > 0 rescales the record count.
struct GenSpec {
  int law = 1;
  uint64_t seed = 1001;
  int64_t n_override = 0;
  double median = 0;     // override log-normal median (law 2 and per-job workloads)
};
void generate(const GenSpec& spec, Records& out);
// One job of law 2: n sequences, log-normal lengths with the given
// median (sigma 0.5) clipped to [150, 6000], 10 children per parent.
void generate_family_job(uint64_t seed, int64_t n, double median, Records& out);

uint64_t splitmix64(uint64_t& x);
uint64_t job_seed(uint64_t base, uint64_t j);

}  // namespace gsdr
