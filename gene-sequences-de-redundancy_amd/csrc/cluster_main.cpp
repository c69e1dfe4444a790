// `cluster` command line: drop-in for the reference's main.cpp / checkOption
// (main.cpp:2-27, func.cpp:2-57). Options are letter/value pairs
// (`t 0.9 i in.fasta o out.fasta w 6`); `-t/-i/-o/-w` and long options are
// accepted as well:
//   --mode plain|batched   launch form (default batched)
//   --spec K               batched: representatives per speculative step
//   --membership FILE      also write sorted index, original index, rep, name
//   --report FILE          JSON report (throughput, filter rates, cells)
//   --device N             HIP device index
//   --drop 0|1             1 = the cuda-version's drop mode (no composition or
//                          shared-word filter; cuda-version/func.cu:1068-1097);
//                          results then differ from the root reference's
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gsdr.h"
#include "host.hpp"

namespace {

// Options through the C ABI's parser (checkOption, func.cpp:6-57); on a
// usage or range error the reference prints its message and exits with
// status 0 (func.cpp:2-5).
gsdr_options parse(int argc, char** argv) {
  gsdr_options o;
  if (gsdr_parse_options(argc, argv, &o) != GSDR_OK) {
    std::cout << gsdr_last_error() << std::endl;
    std::exit(0);
  }
  return o;
}

int die(const char* what, int st) {
  std::cerr << "cluster: " << what << ": " << gsdr_strerror(st) << ": " << gsdr_last_error() << std::endl;
  return 1;
}

}  // namespace

int main(int argc, char** argv) {
  const auto t0 = std::chrono::steady_clock::now();
  const gsdr_options o = parse(argc, argv);
  // func.cpp:53-56
  std::cout << "input file:\t" << o.input << std::endl;
  std::cout << "output file:\t" << o.output << std::endl;
  std::cout << "similarity:\t" << o.threshold << std::endl;
  std::cout << "word length:\t" << o.word_length << std::endl;

  // Device start-up (runtime init, streams) overlaps reading and sorting
  // the input on the calling thread.
  gsdr_ctx* ctx = nullptr;
  int open_st = GSDR_OK;
  double open_s = 0.0;
  std::string open_err;  // the library's error text is per thread
  std::thread opener([&] {
    const auto a = std::chrono::steady_clock::now();
    open_st = gsdr_open(o.device, &ctx);
    if (open_st) open_err = gsdr_last_error();
    open_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
  });

  gsdr_records* recs = nullptr;
  // one pass puts the records straight into the length order
  // (func.cpp:70-107); sort then only reports the permutation
  int st = gsdr_records_read_fasta_sorted(o.input, &recs);
  const auto t_read = std::chrono::steady_clock::now();
  const int64_t n = st ? 0 : gsdr_records_count(recs);
  std::vector<int64_t> perm((size_t)n);
  if (!st) st = gsdr_records_sort(recs, perm.data());
  const auto t_sort = std::chrono::steady_clock::now();
  opener.join();
  if (open_st) {
    std::cerr << "cluster: open device: " << gsdr_strerror(open_st) << ": " << open_err << std::endl;
    return 1;
  }
  if (st) return die(recs ? "sort" : "read", st);
  char name[256];
  gsdr_device_info(ctx, name, sizeof(name), nullptr, nullptr);
  std::cout << "Using: " << name << std::endl;
  int64_t first_len = 0, last_len = 0;
  gsdr_records_get(recs, 0, nullptr, nullptr, nullptr, &first_len);
  gsdr_records_get(recs, n - 1, nullptr, nullptr, nullptr, &last_len);
  // func.cpp:108-111
  std::cout << "read file complete" << std::endl;
  std::cout << "longest/shortest\uff1a\t" << first_len << "/" << last_len << std::endl;
  std::cout << "reads count\uff1a\t" << n << std::endl;

  gsdr_store* store = nullptr;
  const auto t_store0 = std::chrono::steady_clock::now();
  if ((st = gsdr_load(ctx, recs, o.word_length, &store))) return die("load", st);
  const auto t1 = std::chrono::steady_clock::now();

  gsdr_params p;
  std::memset(&p, 0, sizeof(p));
  p.threshold = o.threshold;
  p.word_length = o.word_length;
  p.mode = o.mode;
  p.spec_reps = o.spec_reps;
  p.drop_filters = o.drop_filters;
  std::vector<int32_t> cl((size_t)n);
  gsdr_stats stats;
  std::memset(&stats, 0, sizeof(stats));
  if ((st = gsdr_cluster(ctx, store, &p, cl.data(), &stats))) return die("cluster", st);

  const auto t_clu = std::chrono::steady_clock::now();
  int64_t reps = 0;
  if ((st = gsdr_records_write_result(recs, cl.data(), o.output, &reps))) return die("write", st);
  if (o.membership &&
      (st = gsdr_records_write_membership(recs, perm.data(), cl.data(), o.membership)))
    return die("membership", st);
  const auto t2 = std::chrono::steady_clock::now();
  auto secs = [](auto a, auto b) { return std::chrono::duration<double>(b - a).count(); };
  const double wall = secs(t0, t2);
  const double dev = stats.ms_total / 1e3;
  // func.cpp:798 and main.cpp:26 (wall time here, not CPU time)
  std::cout << "cluster\uff1a" << reps << std::endl;
  std::cout << "time consuming\uff1a" << (long long)(wall * 1000.0) << "ms" << std::endl;
  if (o.report) {
    FILE* f = std::fopen(o.report, "w");
    if (f) {
      std::fprintf(f,
                   "{\"n_seqs\": %lld, \"n_bases\": %lld, \"reps\": %lld, \"rounds\": %lld, "
                   "\"wall_s\": %.6f, \"load_s\": %.6f, \"cluster_s\": %.6f, "
                   "\"open_s\": %.6f, \"read_s\": %.6f, \"sort_s\": %.6f, \"store_s\": %.6f, "
                   "\"cluster_call_s\": %.6f, \"write_s\": %.6f, "
                   "\"clustered_mbp_per_s\": %.3f, \"device_mbp_per_s\": %.3f, "
                   "\"pairs\": %lld, \"rej_length\": %lld, \"rej_comp\": %lld, \"rej_word\": %lld, "
                   "\"aligned\": %lld, \"accepted\": %lld, \"rej_align\": %lld, "
                   "\"cells_band\": %.0f, \"cells_exec\": %.0f, \"cells_full\": %.0f}\n",
                   (long long)stats.n_seqs, (long long)stats.n_bases, (long long)reps, (long long)stats.rounds,
                   wall, secs(t0, t1), dev, open_s, secs(t0, t_read), secs(t_read, t_sort), secs(t_store0, t1),
                   secs(t1, t_clu), secs(t_clu, t2),
                   (double)stats.n_bases / 1e6 / wall, (double)stats.n_bases / 1e6 / (dev > 0 ? dev : 1e-9),
                   (long long)stats.pairs, (long long)stats.rej_length, (long long)stats.rej_comp,
                   (long long)stats.rej_word, (long long)stats.aligned, (long long)stats.accepted,
                   (long long)stats.rej_align, stats.cells_band, stats.cells_exec, stats.cells_full);
      std::fclose(f);
    }
  }
  gsdr_store_free(store);
  gsdr_records_free(recs);
  gsdr_close(ctx);
  return 0;
}
