// Host-side loader, length order, preprocessing, generator and writer.
// Reference semantics cited per function (files of JuZhenCS/gene-sequences-de-redundancy).
#include "host.hpp"

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <thread>
#include <unistd.h>

#include "status.hpp"

namespace gsdr {

namespace {
// memcpy that accepts empty ranges of empty containers (whose data() may be
// null, which memcpy itself does not allow even for zero bytes)
inline void copy_bytes(void* dst, const void* src, size_t n) {
  if (n) std::memcpy(dst, src, n);
}
}  // namespace


// ------------------------------------------------------------------ memory
static constexpr size_t kHuge = size_t(2) << 20;

void* big_alloc(size_t bytes) {
  if (bytes < 16 * kHuge) {
    void* p = std::malloc(bytes ? bytes : 1);
    if (!p) throw std::bad_alloc();
    return p;
  }
  const size_t len = (bytes + kHuge - 1) & ~(kHuge - 1);
  void* p = std::aligned_alloc(kHuge, len);
  if (!p) throw std::bad_alloc();
  madvise(p, len, MADV_HUGEPAGE);
  return p;
}

void big_free(void* p, size_t) noexcept { std::free(p); }

// ---------------------------------------------------------------- encoding
static uint8_t code_of(unsigned char c) {
  switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': case 'U': case 'u': return 3;
    default: return 4;
  }
}
struct CodeTable {
  uint8_t v[256];
  CodeTable() { for (int i = 0; i < 256; ++i) v[i] = code_of((unsigned char)i); }
};
static const CodeTable kTable;
static inline uint8_t base_code(unsigned char c) { return kTable.v[c]; }

namespace {

// Runs body(lo, hi) over [0, n) in chunks handed out dynamically.
template <class F>
void parallel_chunks(int64_t n, int64_t chunk, int threads, F body) {
  if (threads <= 1 || n <= chunk) {
    if (n > 0) body((int64_t)0, n);
    return;
  }
  std::atomic<int64_t> next{0};
  auto run = [&]() {
    for (int64_t lo; (lo = next.fetch_add(chunk)) < n;) body(lo, std::min(n, lo + chunk));
  };
  std::vector<std::thread> th;
  for (int t = 1; t < threads; ++t) th.emplace_back(run);
  run();
  for (auto& x : th) x.join();
}

// ------------------------------------------------------------------- order
// func.cpp:105-107: std::sort(reads, [](a, b) { return a.data.size() >
// b.data.size(); }). libstdc++'s introsort makes the same moves for any
// element type given the same comparison results, so sorting (length, index)
// keys with the same comparator yields the reference's permutation,
// including its (unstable) order of equal lengths.
//
// std::sort is __introsort_loop (median-of-three partitions until a range
// holds 16 elements or the depth limit 2*lg(n) is spent, then heapsort)
// followed by one insertion pass over the whole array. This runs the same
// steps on threads: the partitions of large ranges (each range's partition
// touches only that range, so the order in which disjoint ranges are
// partitioned changes no move), then __introsort_loop on each resulting
// range with its remaining depth, then the insertion pass per range. The
// insertion pass never moves an element across a partition point (every
// element left of one is ordered before or equal to every element right of
// it), and insertion sort is stable, so per range it gives the arrangement
// of the whole pass.
template <class It, class Cmp>
void sort_like_std(It first, It last, Cmp comp, int threads) {
  const ptrdiff_t n = last - first;
  if (n < 2) return;
  const auto c = __gnu_cxx::__ops::__iter_comp_iter(comp);
  struct Range {
    It f, l;
    ptrdiff_t depth;
  };
  const ptrdiff_t grain = std::max<ptrdiff_t>(ptrdiff_t(1) << 15, n / (8 * std::max(threads, 1)));
  if (threads <= 1 || n <= grain) {
    std::sort(first, last, comp);
    return;
  }
  // A shared stack of ranges: a thread takes one, partitions it and pushes
  // both parts while it is large and has depth left, else finishes it.
  // Ranges never wait for one another, so lopsided partitions of heavily
  // tied lengths do not hold up the rest.
  std::vector<Range> stack{Range{first, last, (ptrdiff_t)std::__lg(n) * 2}};
  std::mutex mu;
  std::condition_variable cv;
  int busy = 0;
  auto work = [&]() {
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
      cv.wait(lk, [&] { return !stack.empty() || busy == 0; });
      if (stack.empty()) break;  // and no thread can push more
      const Range r = stack.back();
      stack.pop_back();
      ++busy;
      lk.unlock();
      if (r.l - r.f > grain && r.depth > 0) {
        const It cut = std::__unguarded_partition_pivot(r.f, r.l, c);
        lk.lock();
        stack.push_back(Range{r.f, cut, r.depth - 1});
        stack.push_back(Range{cut, r.l, r.depth - 1});
        --busy;
        cv.notify_all();
      } else {
        std::__introsort_loop(r.f, r.l, r.depth, c);
        std::__insertion_sort(r.f, r.l, c);
        lk.lock();
        --busy;
        if (stack.empty() && busy == 0) cv.notify_all();
      }
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < threads; ++t) th.emplace_back(work);
  work();
  for (auto& x : th) x.join();
}

template <class I>
std::vector<int64_t> sort_keys(const int64_t* lens, int64_t n, int threads) {
  struct Key { I len; I idx; };
  std::vector<Key, NoInitAlloc<Key>> keys((size_t)n);
  parallel_chunks(n, 1 << 16, threads, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i) keys[(size_t)i] = Key{(I)lens[i], (I)i};
  });
  sort_like_std(keys.begin(), keys.end(), [](Key& a, Key& b) { return a.len > b.len; }, threads);
  std::vector<int64_t> perm((size_t)n);
  parallel_chunks(n, 1 << 16, threads, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i) perm[(size_t)i] = (int64_t)keys[(size_t)i].idx;
  });
  return perm;
}

// The moves depend only on the comparison results, so 32-bit keys (when
// lengths and indices fit) give the same permutation with half the traffic.
std::vector<int64_t> order_by_length(const int64_t* lens, int64_t n, int threads) {
  bool narrow = n <= INT32_MAX;
  for (int64_t i = 0; narrow && i < n; ++i) narrow = lens[i] <= INT32_MAX;
  return narrow ? sort_keys<int32_t>(lens, n, threads) : sort_keys<int64_t>(lens, n, threads);
}

// Copies the bytes of [p, e) other than '\n' to dst; returns the bytes written.
int64_t copy_lines(const char* p, const char* e, char* dst) {
  char* d = dst;
  while (p < e) {
    const char* nl = (const char*)memchr(p, '\n', (size_t)(e - p));
    const char* q = nl ? nl : e;
    copy_bytes(d, p, (size_t)(q - p));
    d += q - p;
    p = q + 1;
  }
  return d - dst;
}

}  // namespace

std::vector<int64_t> length_order(const Records& r, int threads) {
  const int64_t n = r.count();
  std::vector<int64_t> lens((size_t)n);
  for (int64_t i = 0; i < n; ++i) lens[(size_t)i] = r.len(i);
  return order_by_length(lens.data(), n, threads);
}

// ------------------------------------------------------------------ reader
// func.cpp:70-104. The first line of the file is the first record's name
// (verbatim). Each later line that starts with '>' starts a new record;
// every other line (blank lines included) is appended raw to the current
// record's data, so a '\r' of CRLF files stays in the data. Where the
// reference loops forever (no trailing newline, a header as the last line,
// an empty file) this reader ends: the last partial line counts as a line
// and a trailing header becomes a record with empty data.
//
// A record therefore starts at offset 0 and at every '>' that follows a
// '\n'; its name runs to the first '\n' after the start and its data is
// every byte up to the next start except '\n'. The reader finds the starts
// in parallel slices of the file, measures each record, and copies the
// records (in file or length order) in parallel.
int read_fasta(const char* path, Records& out, std::string& err, int threads, bool length_order) {
  out.clear();
  int fd = ::open(path, O_RDONLY);
  if (fd < 0) { err = std::string("cannot open input file: ") + path; return GSDR_E_IO; }
  struct stat st;
  if (fstat(fd, &st) != 0) { ::close(fd); err = "cannot stat input file"; return GSDR_E_IO; }
  const int64_t size = (int64_t)st.st_size;
  if (size == 0) { ::close(fd); err = "input file is empty"; return GSDR_E_INPUT; }
  const char* buf = (const char*)mmap(nullptr, (size_t)size, PROT_READ, MAP_PRIVATE, fd, 0);
  ::close(fd);
  if (buf == MAP_FAILED) { err = "cannot map input file"; return GSDR_E_IO; }
  if (threads < 1) threads = 1;

  // One pass over slices of the file finds the record starts and counts
  // the '\n' bytes of each record. Slice sl scans the '\n' bytes at
  // x in [a - 1, b - 1) (the last slice up to the end of the file): each
  // belongs to the record holding x, and one followed by '>' starts a record
  // at x + 1. `head` counts the ones that belong to a record of an earlier
  // slice.
  struct Slice {
    std::vector<int64_t> start, hdr, nl;
    int64_t head = 0;
  };
  const int64_t nslice = std::max<int64_t>(1, std::min<int64_t>(threads * 4, size >> 20));
  std::vector<Slice> part((size_t)nslice);
  parallel_chunks(nslice, 1, threads, [&](int64_t lo, int64_t hi) {
    for (int64_t sl = lo; sl < hi; ++sl) {
      const int64_t a = size * sl / nslice, b = size * (sl + 1) / nslice;
      Slice& v = part[(size_t)sl];
      int64_t* cnt = &v.head;
      bool want_hdr = false;
      if (a == 0) {
        v.start.push_back(0); v.hdr.push_back(-1); v.nl.push_back(0);
        cnt = &v.nl.back();
        want_hdr = true;
      }
      const char* p = buf + std::max<int64_t>(a - 1, 0);
      const char* e = buf + (sl + 1 == nslice ? size : b - 1);
      while (p < e) {
        const char* nl = (const char*)memchr(p, '\n', (size_t)(e - p));
        if (!nl) break;
        ++*cnt;
        if (want_hdr) { v.hdr.back() = nl - buf; want_hdr = false; }
        if (nl + 1 < buf + size && nl[1] == '>') {
          v.start.push_back(nl + 1 - buf); v.hdr.push_back(-1); v.nl.push_back(0);
          cnt = &v.nl.back();
          want_hdr = true;
        }
        p = nl + 1;
      }
    }
  });
  std::vector<int64_t> first((size_t)nslice + 1, 0);
  for (int64_t sl = 0; sl < nslice; ++sl) first[(size_t)sl + 1] = first[(size_t)sl] + (int64_t)part[(size_t)sl].start.size();
  const int64_t n = first[(size_t)nslice];
  Index start((size_t)n + 1), hdr((size_t)n), dlen((size_t)n);
  parallel_chunks(nslice, 1, threads, [&](int64_t lo, int64_t hi) {
    for (int64_t sl = lo; sl < hi; ++sl) {
      Slice& v = part[(size_t)sl];
      const size_t r0 = (size_t)first[(size_t)sl];
      std::copy(v.start.begin(), v.start.end(), start.begin() + r0);
      std::copy(v.hdr.begin(), v.hdr.end(), hdr.begin() + r0);
      std::copy(v.nl.begin(), v.nl.end(), dlen.begin() + r0);
    }
  });
  // '\n' bytes at the head of a slice belong to the last record before it
  for (int64_t sl = 1; sl < nslice; ++sl)
    if (first[(size_t)sl] > 0) dlen[(size_t)first[(size_t)sl] - 1] += part[(size_t)sl].head;
  part.clear();
  start[(size_t)n] = size;
  // bytes of [s, e) = name + every '\n' + data; a name without '\n' ends
  // at e (the last record); a name cut by a slice edge is found here
  parallel_chunks(n, 4096, threads, [&](int64_t lo, int64_t hi) {
    for (int64_t r = lo; r < hi; ++r) {
      const int64_t s = start[(size_t)r], e = start[(size_t)r + 1];
      int64_t h = hdr[(size_t)r];
      if (h < 0) {
        const char* nl = (const char*)memchr(buf + s, '\n', (size_t)(e - s));
        h = nl ? nl - buf : e;
        hdr[(size_t)r] = h;
      }
      dlen[(size_t)r] = (e - s) - (h - s) - dlen[(size_t)r];
    }
  });

  std::vector<int64_t> perm;
  if (length_order) perm = order_by_length(dlen.data(), n, threads);
  // records are copied in file order (sequential reads of the file) to
  // their place in the output order
  Index slot;
  if (length_order) {
    slot.resize((size_t)n);
    parallel_chunks(n, 65536, threads, [&](int64_t lo, int64_t hi) {
      for (int64_t i = lo; i < hi; ++i) slot[(size_t)perm[(size_t)i]] = i;
    });
  }
  // every entry past the first is written below
  out.name_off.resize((size_t)n + 1);
  out.data_off.resize((size_t)n + 1);
  out.name_off[0] = 0;
  out.data_off[0] = 0;
  parallel_chunks(n, 8192, threads, [&](int64_t lo, int64_t hi) {
    for (int64_t r = lo; r < hi; ++r) {
      const int64_t i = length_order ? slot[(size_t)r] : r;
      out.name_off[(size_t)i + 1] = hdr[(size_t)r] - start[(size_t)r];
      out.data_off[(size_t)i + 1] = dlen[(size_t)r];
    }
  });
  for (int64_t i = 0; i < n; ++i) {
    out.name_off[(size_t)i + 1] += out.name_off[(size_t)i];
    out.data_off[(size_t)i + 1] += out.data_off[(size_t)i];
  }
  out.names.resize((size_t)out.name_off[(size_t)n]);
  out.data.resize((size_t)out.data_off[(size_t)n]);
  parallel_chunks(n, 1024, threads, [&](int64_t lo, int64_t hi) {
    for (int64_t r = lo; r < hi; ++r) {
      const int64_t i = length_order ? slot[(size_t)r] : r;
      const int64_t s = start[(size_t)r], e = start[(size_t)r + 1], h = hdr[(size_t)r];
      copy_bytes(out.names.data() + out.name_off[(size_t)i], buf + s, (size_t)(h - s));
      if (h + 1 < e) copy_lines(buf + h + 1, buf + e, out.data.data() + out.data_off[(size_t)i]);
    }
  });
  munmap((void*)buf, (size_t)size);
  if (length_order) {
    out.length_ordered = true;
    out.order = std::move(perm);
  }
  return GSDR_OK;
}

Records permute(const Records& r, const std::vector<int64_t>& perm, int threads) {
  const int64_t n = (int64_t)perm.size();
  Records o;
  o.name_off.assign((size_t)n + 1, 0);
  o.data_off.assign((size_t)n + 1, 0);
  for (int64_t i = 0; i < n; ++i) {
    const int64_t src = perm[(size_t)i];
    o.name_off[(size_t)i + 1] = o.name_off[(size_t)i] + (r.name_off[src + 1] - r.name_off[src]);
    o.data_off[(size_t)i + 1] = o.data_off[(size_t)i] + (r.data_off[src + 1] - r.data_off[src]);
  }
  o.names.resize((size_t)o.name_off[(size_t)n]);
  o.data.resize((size_t)o.data_off[(size_t)n]);
  if (threads < 1) threads = 1;
  std::atomic<int64_t> next{0};
  auto copy = [&]() {
    constexpr int64_t kChunk = 1024;
    for (int64_t lo; (lo = next.fetch_add(kChunk)) < n;) {
      const int64_t hi = std::min(n, lo + kChunk);
      for (int64_t i = lo; i < hi; ++i) {
        const int64_t src = perm[(size_t)i];
        copy_bytes(o.names.data() + o.name_off[(size_t)i], r.names.data() + r.name_off[src],
                    (size_t)(r.name_off[src + 1] - r.name_off[src]));
        copy_bytes(o.data.data() + o.data_off[(size_t)i], r.data.data() + r.data_off[src],
                    (size_t)(r.data_off[src + 1] - r.data_off[src]));
      }
    }
  };
  std::vector<std::thread> th;
  for (int t = 0; t < threads; ++t) th.emplace_back(copy);
  for (auto& x : th) x.join();
  return o;
}

// ------------------------------------------------------------------ writer
// func.cpp:785-796: for s ascending with cluster[s] == s, "name\n" then the
// raw data and "\n".
int64_t write_result(const Records& sorted, const int32_t* cluster, const char* path) {
  FILE* f = fopen(path, "wb");
  if (!f) return -1;
  std::vector<char> buf;
  buf.reserve(1 << 22);
  int64_t reps = 0;
  const int64_t n = sorted.count();
  for (int64_t s = 0; s < n; ++s) {
    if (cluster[s] != (int32_t)s) continue;
    ++reps;
    buf.insert(buf.end(), sorted.names.begin() + sorted.name_off[s], sorted.names.begin() + sorted.name_off[s + 1]);
    buf.push_back('\n');
    buf.insert(buf.end(), sorted.data.begin() + sorted.data_off[s], sorted.data.begin() + sorted.data_off[s + 1]);
    buf.push_back('\n');
    if (buf.size() > (1u << 22)) {
      if (fwrite(buf.data(), 1, buf.size(), f) != buf.size()) { fclose(f); return -1; }
      buf.clear();
    }
  }
  if (!buf.empty() && fwrite(buf.data(), 1, buf.size(), f) != buf.size()) { fclose(f); return -1; }
  if (fclose(f) != 0) return -1;
  return reps;
}

int write_fasta(const Records& r, const char* path, int width) {
  FILE* f = fopen(path, "wb");
  if (!f) return GSDR_E_IO;
  std::vector<char> buf;
  buf.reserve(1 << 22);
  for (int64_t i = 0; i < r.count(); ++i) {
    buf.insert(buf.end(), r.names.begin() + r.name_off[i], r.names.begin() + r.name_off[i + 1]);
    buf.push_back('\n');
    int64_t a = r.data_off[i], b = r.data_off[i + 1];
    for (int64_t p = a; p < b; p += width) {
      int64_t e = std::min<int64_t>(b, p + width);
      buf.insert(buf.end(), r.data.begin() + p, r.data.begin() + e);
      buf.push_back('\n');
    }
    if (buf.size() > (1u << 22)) {
      if (fwrite(buf.data(), 1, buf.size(), f) != buf.size()) { fclose(f); return GSDR_E_IO; }
      buf.clear();
    }
  }
  if (!buf.empty() && fwrite(buf.data(), 1, buf.size(), f) != buf.size()) { fclose(f); return GSDR_E_IO; }
  return fclose(f) == 0 ? GSDR_OK : GSDR_E_IO;
}

// ----------------------------------------------------------------- cutoffs
// func.cpp:36-45: the float threshold is compared with double literals.
int auto_word_length(float t) {
  if (t < 0.88) return 4;
  if (t < 0.94) return 5;
  if (t < 0.97) return 6;
  return 7;
}
// func.cpp:368-369: ceil((float)L * t) in float32.
int base_cutoff(int32_t raw_len, float t) {
  volatile float prod = (float)raw_len * t;
  return (int)std::ceil((float)prod);
}
// func.cpp:353-366 (the offset factor is overwritten with 1 at :364).
int word_cutoff(int32_t raw_len, float t, int k) {
  int required = raw_len - k + 1;
  volatile float one_minus = 1.0f - t;
  volatile float prod = (float)raw_len * one_minus;
  float ce = std::ceil((float)prod);
  volatile float scaled = ce * (float)k;
  int cutoff = (int)(float)scaled;
  required -= cutoff;
  if (required < 1) required = 1;
  return required;
}

// ------------------------------------------------------------- preprocess
// Builds, per sorted sequence: composition of codes 0..3 (func.cpp:207-232),
// the gap-free sequence packed as two bit planes per 32 bases with bit j of
// word i holding base 32i+j (func.cpp:375-446; lo plane = bit 0 of the code,
// hi plane = bit 1), and the multiset of k-mers over raw windows that contain
// no code-4 base (func.cpp:252-338) stored as distinct (word, count) entries.
// Counts are kept modulo 2^16 like the reference's unsigned short counters.
void preprocess(const Records& sorted, int k, int threads, Prep& p) {
  const int64_t n = sorted.count();
  p.n = (int32_t)n;
  p.k = k;
  p.raw_len.assign((size_t)n, 0);
  p.m.assign((size_t)n, 0);
  p.comp.assign((size_t)n * 4, 0);
  p.pk_off.assign((size_t)n + 1, 0);
  p.kp_off.assign((size_t)n + 1, 0);
  if (threads < 1) threads = 1;
  // Pass 1: lengths, composition and the distinct k-mers of every sequence,
  // kept in a buffer of the thread that made them. Sequences are handed out
  // in chunks from a shared counter: the input is sorted by length, so
  // fixed ranges would leave one thread all the long sequences.
  constexpr int64_t kChunk = 256;
  std::vector<std::vector<uint32_t>> tbuf((size_t)threads);
  std::vector<int32_t> owner((size_t)n, 0);
  std::vector<int64_t> at((size_t)n, 0);
  std::vector<int32_t> distinct((size_t)n, 0);
  auto pass1 = [&](int tid, std::atomic<int64_t>* next) {
    const uint32_t nwords = 1u << (2 * k);
    std::vector<uint32_t> cnt(nwords, 0);
    std::vector<uint32_t> touched(nwords);
    std::vector<uint32_t>& out = tbuf[(size_t)tid];
    // composition as two registers of two 32-bit counts (A|C, G|T)
    static const uint64_t kIncLo[5] = {1ull, 1ull << 32, 0, 0, 0};
    static const uint64_t kIncHi[5] = {0, 0, 1ull, 1ull << 32, 0};
    for (int64_t lo; (lo = next->fetch_add(kChunk)) < n;) {
      const int64_t hi = std::min(n, lo + kChunk);
      for (int64_t s = lo; s < hi; ++s) {
        const unsigned char* d = (const unsigned char*)sorted.data.data() + sorted.data_off[s];
        const int64_t L = sorted.len(s);
        uint64_t clo = 0, chi = 0;
        uint32_t nt = 0, word = 0;
        const uint32_t mask = nwords - 1;
        int valid = 0;
        for (int64_t i = 0; i < L; ++i) {
          const uint8_t c = base_code(d[i]);
          clo += kIncLo[c];
          chi += kIncHi[c];
          if (c == 4) { valid = 0; word = 0; continue; }
          word = ((word << 2) | c) & mask;
          if (++valid >= k) {
            if (cnt[word]++ == 0) touched[nt++] = word;
          }
        }
        const uint32_t ca = (uint32_t)clo, cc = (uint32_t)(clo >> 32), cg = (uint32_t)chi, ct = (uint32_t)(chi >> 32);
        p.raw_len[s] = (int32_t)L;
        p.m[s] = (int32_t)(ca + cc + cg + ct);
        p.comp[s * 4 + 0] = ca;
        p.comp[s * 4 + 1] = cc;
        p.comp[s * 4 + 2] = cg;
        p.comp[s * 4 + 3] = ct;
        owner[s] = tid;
        at[s] = (int64_t)out.size();
        distinct[s] = (int32_t)nt;
        for (uint32_t e = 0; e < nt; ++e) {
          const uint32_t w = touched[e];
          out.push_back(w | ((cnt[w] & 0xFFFFu) << 16));
          cnt[w] = 0;
        }
      }
    }
  };
  auto run = [&](auto fn) {
    std::atomic<int64_t> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < threads; ++t) th.emplace_back(fn, t, &next);
    for (auto& x : th) x.join();
  };
  run(pass1);
  p.max_m = 0;
  for (int64_t s = 0; s < n; ++s) {
    p.pk_off[s + 1] = p.pk_off[s] + (p.m[s] / 32 + 1);
    p.kp_off[s + 1] = p.kp_off[s] + distinct[s];
    p.max_m = std::max(p.max_m, p.m[s]);
  }
  p.pk.assign((size_t)p.pk_off[n] * 2, 0u);
  p.kp.resize((size_t)p.kp_off[n]);
  // Pass 2: the gap-free bases as two bit planes, and the profiles moved to
  // their place in sorted order.
  auto pass2 = [&](int, std::atomic<int64_t>* next) {
    for (int64_t lo; (lo = next->fetch_add(kChunk)) < n;) {
      const int64_t hi = std::min(n, lo + kChunk);
      for (int64_t s = lo; s < hi; ++s) {
        const char* d = sorted.data.data() + sorted.data_off[s];
        const int64_t L = sorted.len(s);
        uint32_t* w2 = p.pk.data() + p.pk_off[s] * 2;
        // both planes of 32 gap-free bases built in registers, stored once
        uint32_t plo = 0, phi = 0;
        int j = 0;
        for (int64_t i = 0; i < L; ++i) {
          const uint32_t c = base_code((unsigned char)d[i]);
          if (c == 4) continue;
          plo |= (c & 1u) << j;
          phi |= (c >> 1) << j;
          if (++j == 32) {
            w2[0] = plo;
            w2[1] = phi;
            w2 += 2;
            plo = phi = 0;
            j = 0;
          }
        }
        if (j > 0) {
          w2[0] = plo;
          w2[1] = phi;
        }
        if (distinct[s] > 0) {  // (no profile at all: kp may have no storage)
          const uint32_t* src = tbuf[(size_t)owner[s]].data() + at[s];
          copy_bytes(p.kp.data() + p.kp_off[s], src, sizeof(uint32_t) * (size_t)distinct[s]);
        }
      }
    }
  };
  run(pass2);
}

// -------------------------------------------------------------- generator
uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
uint64_t job_seed(uint64_t base, uint64_t j) {
  uint64_t x = base ^ (0xd1b54a32d192ed03ull * (j + 1));
  return splitmix64(x);
}

namespace {
// xoshiro256** seeded from splitmix64.
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) { uint64_t x = seed; for (auto& v : s) v = splitmix64(x); }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  double uni() { return (double)(next() >> 11) * 0x1.0p-53; }
  uint32_t below(uint32_t n) { return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32); }
  int64_t range(int64_t lo, int64_t hi) { return lo + (int64_t)(uni() * (double)(hi - lo + 1)); }
  double normal() {
    double u1 = uni(), u2 = uni();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};
const char kACGT[4] = {'A', 'C', 'G', 'T'};

void random_seq(Rng& g, int64_t L, std::string& out) {
  out.resize((size_t)L);
  for (int64_t i = 0; i < L; ++i) out[(size_t)i] = kACGT[g.below(4)];
}
// Per parent base: deleted with p_del; otherwise substituted with p_sub (to
// one of the three other bases); then a uniform base is inserted after it
// with p_ins.
void mutate(Rng& g, const std::string& parent, double p_sub, double p_ins, double p_del, std::string& out) {
  out.clear();
  out.reserve(parent.size() + parent.size() / 8 + 8);
  for (char b : parent) {
    if (g.uni() >= p_del) {
      if (g.uni() < p_sub) {
        int c = 0;
        while (kACGT[c] != b) ++c;
        out.push_back(kACGT[(c + 1 + g.below(3)) & 3]);
      } else {
        out.push_back(b);
      }
    }
    if (g.uni() < p_ins) out.push_back(kACGT[g.below(4)]);
  }
}
void shuffle_records(Rng& g, std::vector<std::string>& names, std::vector<std::string>& seqs) {
  for (size_t i = seqs.size(); i > 1; --i) {
    size_t j = (size_t)g.below((uint32_t)i);
    std::swap(seqs[i - 1], seqs[j]);
    std::swap(names[i - 1], names[j]);
  }
}
void to_records(const std::vector<std::string>& names, const std::vector<std::string>& seqs, Records& out) {
  out.clear();
  for (size_t i = 0; i < seqs.size(); ++i)
    out.push(names[i].data(), (int64_t)names[i].size(), seqs[i].data(), (int64_t)seqs[i].size());
}
int64_t lognormal_len(Rng& g, double median, double sigma, int64_t lo, int64_t hi) {
  double v = std::exp(std::log(median) + sigma * g.normal());
  int64_t L = (int64_t)std::llround(v);
  return std::min(hi, std::max(lo, L));
}
}  // namespace

void generate_family_job(uint64_t seed, int64_t n, double median, Records& out) {
  Rng g(seed);
  std::vector<std::string> names, seqs;
  names.reserve((size_t)n); seqs.reserve((size_t)n);
  std::string parent, child;
  int64_t parents = (n + 9) / 10;
  for (int64_t p = 0; p < parents && (int64_t)seqs.size() < n; ++p) {
    int64_t L = lognormal_len(g, median, 0.5, 150, 6000);
    random_seq(g, L, parent);
    for (int c = 0; c < 10 && (int64_t)seqs.size() < n; ++c) {
      double d = 0.12 * g.uni();
      mutate(g, parent, 0.8 * d, 0.1 * d, 0.1 * d, child);
      names.push_back(">f" + std::to_string(p) + "_c" + std::to_string(c));
      seqs.push_back(child);
    }
  }
  shuffle_records(g, names, seqs);
  to_records(names, seqs, out);
}

void generate(const GenSpec& spec, Records& out) {
  Rng g(spec.seed);
  std::vector<std::string> names, seqs;
  std::string parent, child;
  if (spec.law == 1) {
    // 500 parents, L ~ U{300..1500}; each parent plus 19 children
    // (sub 0.02, ins 0.005, del 0.005).
    int64_t total = spec.n_override > 0 ? spec.n_override : 10000;
    int64_t parents = (total + 19) / 20;
    for (int64_t p = 0; p < parents; ++p) {
      int64_t L = g.range(300, 1500);
      random_seq(g, L, parent);
      for (int c = 0; c < 20 && (int64_t)seqs.size() < total; ++c) {
        if (c == 0) child = parent;
        else mutate(g, parent, 0.02, 0.005, 0.005, child);
        names.push_back(">p" + std::to_string(p) + "_" + std::to_string(c));
        seqs.push_back(child);
      }
    }
    shuffle_records(g, names, seqs);
    to_records(names, seqs, out);
    return;
  }
  if (spec.law == 2) {
    int64_t total = spec.n_override > 0 ? spec.n_override : 200000;
    double median = spec.median > 0 ? spec.median : 1200.0;
    generate_family_job(spec.seed, total, median, out);
    return;
  }
  if (spec.law == 4) {
    // Pareto(1.5, min 100) lengths clipped at 30000; 30% singletons, 70% in
    // families of Geometric(1/8) size whose members are the (unemitted) root
    // mutated at d ~ U[0, 0.06] (80% substitutions, 20% indels); 5% of all
    // sequences get a tandem repeat (unit 1-6, run 50-500) written over a
    // window at a uniform position.
    int64_t total = spec.n_override > 0 ? spec.n_override : 2000000;
    int64_t singles = (int64_t)std::llround(0.3 * (double)total);
    auto pareto = [&]() {
      double u = g.uni();
      if (u < 1e-300) u = 1e-300;
      double v = 100.0 * std::pow(u, -1.0 / 1.5);
      int64_t L = v >= 30000.0 ? 30000 : (int64_t)std::floor(v);
      return std::max<int64_t>(100, L);
    };
    names.reserve((size_t)total); seqs.reserve((size_t)total);
    for (int64_t i = 0; i < singles; ++i) {
      random_seq(g, pareto(), child);
      names.push_back(">s" + std::to_string(i));
      seqs.push_back(child);
    }
    int64_t fam = 0;
    while ((int64_t)seqs.size() < total) {
      int64_t size = 1;
      while (g.uni() >= 1.0 / 8.0) ++size;
      random_seq(g, pareto(), parent);
      for (int64_t c = 0; c < size && (int64_t)seqs.size() < total; ++c) {
        double d = 0.06 * g.uni();
        mutate(g, parent, 0.8 * d, 0.1 * d, 0.1 * d, child);
        names.push_back(">g" + std::to_string(fam) + "_" + std::to_string(c));
        seqs.push_back(child);
      }
      ++fam;
    }
    for (auto& s : seqs) {
      if (g.uni() >= 0.05) continue;
      int unit_len = (int)g.range(1, 6);
      char unit[6];
      for (int u = 0; u < unit_len; ++u) unit[u] = kACGT[g.below(4)];
      int64_t run = std::min<int64_t>((int64_t)s.size(), g.range(50, 500));
      int64_t at = g.range(0, (int64_t)s.size() - run);
      for (int64_t x = 0; x < run; ++x) s[(size_t)(at + x)] = unit[x % unit_len];
    }
    shuffle_records(g, names, seqs);
    to_records(names, seqs, out);
    return;
  }
  // many-job workloads and stream requests are generated per job through
  // generate_family_job with their own seeds and medians.
  int64_t total = spec.n_override > 0 ? spec.n_override : 10000;
  generate_family_job(spec.seed, total, spec.median > 0 ? spec.median : 1200.0, out);
}

}  // namespace gsdr
