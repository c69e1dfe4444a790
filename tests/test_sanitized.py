"""The host code and the oracle under AddressSanitizer and UBSan (CPU).

The host side of the library (csrc/host.cpp: parallel FASTA reader into the
reference's length order, writers, cutoffs, preprocessing, generator) has no
device code, so g++ builds it here with -fsanitize=address,undefined beside a
small driver that runs it over awkward FASTA files (no trailing newline,
CRLF, blank lines, empty and header-only records, IUPAC codes, a trailing
header, one long line, more threads than records) and generated workloads,
and checks its results against the oracle's reader on the same files. The
oracle's command (oracle/oracle.cpp) is built the same way and clusters a
small generated file. Any sanitizer report fails the test.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "gene-sequences-de-redundancy_amd", "csrc")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")

DRIVER = r"""
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>
#include "host.hpp"
using namespace gsdr;
// driver <fasta> <threads>: prints per record (sorted order): original index,
// length, header; then checks preprocessing, cutoffs and writers
int main(int argc, char** argv) {
  if (argc < 3) return 2;
  const int threads = atoi(argv[2]);
  Records a, b;
  std::string err;
  if (read_fasta(argv[1], a, err, threads, true)) { printf("ERR %s\n", err.c_str()); return 0; }
  if (read_fasta(argv[1], b, err, 1, false)) { printf("ERR2 %s\n", err.c_str()); return 1; }
  // the one-pass length order equals sorting the file-order records
  const std::vector<int64_t> perm = length_order(b);
  Records c = permute(b, perm, threads);
  if (c.count() != a.count()) return 3;
  for (int64_t i = 0; i < a.count(); ++i) {
    if (a.order[i] != perm[i] || a.len(i) != c.len(i)) return 4;
    for (int64_t j = 0; j < a.len(i); ++j)
      if (a.data[a.data_off[i] + j] != c.data[c.data_off[i] + j]) return 5;
    printf("%lld %lld ", (long long)a.order[i], (long long)a.len(i));
    fwrite(a.names.data() + a.name_off[i], 1, (size_t)(a.name_off[i + 1] - a.name_off[i]), stdout);
    printf("\n");
  }
  for (int k = 4; k <= 8; ++k) {
    Prep p;
    preprocess(a, k, threads, p);
    if (p.n != (int32_t)a.count()) return 6;
  }
  long long sum = 0;
  for (int32_t L = 0; L <= 40000; L += 7)
    for (float t : {0.8f, 0.85f, 0.88f, 0.9f, 0.94f, 0.95f, 0.97f, 0.99f})
      sum += base_cutoff(L, t) + word_cutoff(L, t, auto_word_length(t));
  std::vector<int32_t> cl((size_t)a.count());
  for (int64_t i = 0; i < a.count(); ++i) cl[(size_t)i] = (int32_t)(i & ~1);
  const std::string out = std::string(argv[1]) + ".result";
  if (write_result(a, cl.data(), out.c_str()) < 0) return 7;
  if (write_fasta(a, (std::string(argv[1]) + ".w80").c_str(), 80)) return 8;
  printf("cutoffs %lld\n", sum);
  return 0;
}
"""

GEN = r"""
#include <cstdio>
#include "host.hpp"
using namespace gsdr;
int main() {
  for (int law : {1, 2, 4}) {
    GenSpec g; g.law = law; g.seed = 1000 + law; g.n_override = law == 4 ? 3000 : 1500;
    Records r; generate(g, r);
    Prep p; preprocess(r, 6, 4, p);
    printf("law %d n %lld\n", law, (long long)r.count());
  }
  Records r; generate_family_job(job_seed(1003, 5), 800, 400.0, r);
  printf("job n %lld\n", (long long)r.count());
  return 0;
}
"""

FILES = {
    "no_newline.fa": b">a\nACGT\nAC\n>b\nACGTACGT",
    "crlf.fa": b">a x\r\nACGTN\r\nRYACG\r\n>b\r\nAC\r\n\r\n>c\r\nACGTACGTAC\r\n",
    "blank.fa": b"\n\n>a\n\nACG\n\n>b\n>c\nA\n",
    "trailing_header.fa": b">a\nACGT\n>b\n",
    "header_only.fa": b">only\n",
    "long_line.fa": b">l\n" + b"ACGT" * 20000 + b"\n>s\nA\n",
    "iupac.fa": b">x\nNNNNRYKMSWBDHV\n>y\nacgtuACGTU\n",
    "many.fa": b"".join(b">r%d\n" % i + b"ACGT"[i % 4:i % 4 + 1] * (i % 37 + 1) + b"\n" for i in range(300)),
}


def build(tmp_path, name, src, extra):
    cpp = tmp_path / f"{name}.cpp"
    cpp.write_text(src)
    exe = tmp_path / name
    r = subprocess.run(["g++", "-std=c++17", *SAN, "-I", CSRC, str(cpp), *extra, "-o", str(exe), "-lpthread"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


@pytest.fixture(scope="module")
def host_driver(tmp_path_factory):
    return build(tmp_path_factory.mktemp("san"), "driver", DRIVER, [os.path.join(CSRC, "host.cpp")])


@pytest.mark.parametrize("name", sorted(FILES))
@pytest.mark.parametrize("threads", [1, 3, 16])
def test_reader_and_writers_sanitized(tmp_path, host_driver, name, threads):
    import oracle_ffi as O

    fa = tmp_path / name
    fa.write_bytes(FILES[name])
    r = subprocess.run([str(host_driver), str(fa), str(threads)], capture_output=True, timeout=300, env=ENV)
    assert r.returncode == 0, (r.returncode, r.stderr.decode(errors="replace")[-3000:])
    assert b"runtime error" not in r.stderr and b"AddressSanitizer" not in r.stderr
    lines = r.stdout.decode(errors="replace").splitlines()
    if lines and lines[0].startswith("ERR"):
        # only a file without any record is refused
        assert name == "header_only.fa" or not FILES[name].strip(), lines[0]
        return
    # the oracle reads the same records, and its length order is the same
    got = [(int(x.split()[0]), int(x.split()[1])) for x in lines[:-1]]
    recs = O.read_fasta(str(fa))
    assert len(got) == len(recs)
    assert [len(recs[i][1]) for i, _ in got] == [g[1] for g in got]
    assert [i for i, _ in got] == O.sort_order([len(d) for _, d in recs]).tolist()


def _big_fasta(path):
    """200,000 records of 1..53 bases: enough to split the length order over
    threads, with long runs of equal lengths."""
    path.write_bytes(b"".join(b">b%d\n" % i + b"ACGT"[i % 4:i % 4 + 1] * (i * 7919 % 53 + 1) + b"\n"
                              for i in range(200_000)))


@pytest.mark.parametrize("threads", [3, 16])
def test_threaded_length_order_sanitized(tmp_path, host_driver, threads):
    """The threaded length order (partitions from a shared task stack) under
    ASan/UBSan: the driver checks it against the one-thread std::sort order."""
    import oracle_ffi as O

    fa = tmp_path / "big.fa"
    _big_fasta(fa)
    r = subprocess.run([str(host_driver), str(fa), str(threads)], capture_output=True, timeout=600, env=ENV)
    assert r.returncode == 0, (r.returncode, r.stderr.decode(errors="replace")[-3000:])
    assert b"runtime error" not in r.stderr and b"AddressSanitizer" not in r.stderr
    got = [int(x.split()[0]) for x in r.stdout.decode().splitlines()[:-1]]
    lens = [i * 7919 % 53 + 1 for i in range(200_000)]
    assert got == O.sort_order(lens).tolist()


def test_threaded_length_order_thread_sanitizer(tmp_path):
    """The same reader and length order built with ThreadSanitizer: no data
    race between the threads that partition and finish the ranges."""
    exe = tmp_path / "tsan_driver"
    cpp = tmp_path / "tsan_driver.cpp"
    cpp.write_text(DRIVER)
    r = subprocess.run(["g++", "-std=c++17", "-fsanitize=thread", "-g", "-O1", "-I", CSRC, str(cpp),
                        os.path.join(CSRC, "host.cpp"), "-o", str(exe), "-lpthread"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    fa = tmp_path / "big.fa"
    _big_fasta(fa)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([str(exe), str(fa), "8"], capture_output=True, timeout=900, env=env)
    assert r.returncode == 0, (r.returncode, r.stderr.decode(errors="replace")[-3000:])
    assert b"ThreadSanitizer" not in r.stderr


def test_generator_and_preprocessing_sanitized(tmp_path):
    exe = build(tmp_path, "gen", GEN, [os.path.join(CSRC, "host.cpp")])
    r = subprocess.run([str(exe)], capture_output=True, timeout=600, env=ENV)
    assert r.returncode == 0, r.stderr.decode(errors="replace")[-3000:]
    assert b"law 4 n 3000" in r.stdout


def test_oracle_sanitized(tmp_path):
    exe = tmp_path / "orc"
    r = subprocess.run(["g++", "-std=c++17", *SAN, "-fopenmp", "-DORC_MAIN", os.path.join(ROOT, "oracle", "oracle.cpp"),
                        "-o", str(exe)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    fa = tmp_path / "in.fa"
    fa.write_bytes(FILES["many.fa"] + FILES["crlf.fa"] + FILES["iupac.fa"])
    env = dict(ENV, OMP_NUM_THREADS="2", ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    for t in ("0.8", "0.95", "0.99"):
        r = subprocess.run([str(exe), str(fa), str(tmp_path / "out.fa"), t], capture_output=True, timeout=600, env=env)
        assert r.returncode == 0, r.stderr.decode(errors="replace")[-3000:]
        assert b"runtime error" not in r.stderr
