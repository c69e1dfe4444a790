"""CPU tests of the product's host side and the C ABI boundary (no GPU calls).

- libgsdr.so loads and exports every function include/gsdr.h declares;
- records: generator determinism and sizes, FASTA round trip, writers;
- the `cluster` CLI keeps the reference's option handling (func.cpp:2-57);
- the reference's shipped result.fasta subset is a fixed point of the oracle
  (idempotence known answer).
"""
import os
import re
import subprocess

import numpy as np
import pytest

import gsdr
import oracle_ffi as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gsdr.h")
GOLDEN = os.path.join(ROOT, "tests", "golden")
HAS_GPU = os.path.exists("/dev/kfd")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gsdr_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 25
    out = subprocess.run(["nm", "-D", "--defined-only", gsdr.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (gsdr_[a-z_]+)\b", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    # and the Python binding covers them all
    bound = {name for name, _, _ in gsdr.SIGNATURES}
    assert set(names) <= bound


def test_library_loads_with_signatures():
    L = gsdr.lib()
    assert L.gsdr_strerror(0) == b"ok"
    assert L.gsdr_strerror(gsdr.E_LIMIT) == b"size limit exceeded"


@pytest.mark.skipif(HAS_GPU, reason="checks the no-device error path")
def test_open_fails_loudly_without_device():
    with pytest.raises(gsdr.GsdrError) as e:
        gsdr.Engine(0)
    assert e.value.status == gsdr.E_DEVICE


def test_generator_deterministic_and_sized():
    a = gsdr.Records.generate(1, 1001)
    b = gsdr.Records.generate(1, 1001)
    assert len(a) == 10000 and 8.5e6 < a.bases < 9.5e6
    assert [a.get(i) for i in range(0, 10000, 997)] == [b.get(i) for i in range(0, 10000, 997)]
    c = gsdr.Records.generate(2, 1002, 5000)
    assert len(c) == 5000
    lens = [len(c.get(i)[1]) for i in range(5000)]
    assert min(lens) >= 100 and max(lens) <= 7000
    d = gsdr.Records.generate(4, 1004, 20000)
    lens = np.array([len(d.get(i)[1]) for i in range(20000)])
    assert lens.min() >= 90 and lens.max() <= 31000
    assert abs(np.median(lens) - 160) < 40  # Pareto(1.5, 100): median 100 * 2^(2/3) = 159
    e = gsdr.Records.generate(1, 1002)
    assert a.get(0) != e.get(0)


def test_fasta_roundtrip_and_sort(tmp_path):
    r = gsdr.Records.generate(2, 7, 700)
    p = tmp_path / "x.fasta"
    r.write_fasta(str(p), 80)
    for line in p.read_bytes().split(b"\n"):
        assert len(line) <= 80 or line.startswith(b">")
    r2 = gsdr.Records.read_fasta(str(p))
    assert len(r2) == len(r)
    assert [r2.get(i) for i in range(len(r))] == [r.get(i) for i in range(len(r))]
    assert O.read_fasta(str(p)) == [r.get(i) for i in range(len(r))]
    perm = r2.sort()
    lens = np.array([len(r.get(i)[1]) for i in range(len(r))])
    assert np.array_equal(perm, O.sort_order(lens))


@pytest.mark.parametrize("pattern", ["ties", "pareto", "equal", "saw", "down", "pipe"])
def test_threaded_length_order_matches_std_sort(pattern):
    """The length order is std::sort's, unstable tie order included, with
    the top partitions and the per-range sorts on threads: 300,000 records
    (several levels of partitions) against the oracle's std::sort over
    std::string records."""
    n = 300_000
    rng = np.random.default_rng(11)
    i = np.arange(n)
    lens = {
        "ties": rng.integers(0, 41, n),
        "pareto": np.minimum((10 * rng.random(n) ** (-1 / 1.5)).astype(np.int64), 3000),
        "equal": np.full(n, 5),
        "saw": i % 97,
        "down": (n - i) % 50,
        "pipe": np.minimum(i, n - i) % 200,
    }[pattern].astype(np.int64)
    r = gsdr.Records.from_lists([b"A" * int(x) for x in lens])
    try:
        assert np.array_equal(r.sort(), O.sort_order(lens))
    finally:
        r.free()


def test_parallel_reader_awkward_file(tmp_path):
    """A file of several MiB read in parallel slices: ragged line widths,
    blank lines, CRLF, '>' inside data lines, a header right after a header,
    data before the first header line and no trailing newline; both readers
    against the oracle's reader and order."""
    rng = np.random.default_rng(11)
    parts = [b"first line is a name\n", b"ACGT\n"]
    alphabet = np.frombuffer(b"ACGTNacgt>", dtype=np.uint8)
    while sum(len(x) for x in parts) < (6 << 20):
        parts.append(b">r%d\n" % len(parts))
        kind = rng.integers(0, 8)
        if kind == 0:
            continue  # header followed by a header
        for _ in range(int(rng.integers(1, 12))):
            w = int(rng.integers(0, 200))
            line = alphabet[rng.integers(0, len(alphabet) - 1 if kind < 6 else len(alphabet), w)].tobytes()
            parts.append(line + (b"\r\n" if kind == 7 else b"\n"))
    parts.append(b">tail\nACGTACGT")
    p = tmp_path / "awkward.fasta"
    p.write_bytes(b"".join(parts))
    expect = O.read_fasta(str(p))
    assert len(expect) > 10000
    r = gsdr.Records.read_fasta(str(p))
    assert len(r) == len(expect)
    assert all(r.get(i) == expect[i] for i in range(len(expect)))
    order = O.sort_order([len(d) for _, d in expect])
    s = gsdr.Records.read_fasta(str(p), length_order=True)
    assert all(s.get(i) == expect[order[i]] for i in range(len(expect)))
    assert np.array_equal(s.sort(), order)
    assert np.array_equal(r.sort(), order)
    assert np.array_equal(r.sort(), order)  # already in order: same permutation again


def test_result_and_membership_writers(tmp_path):
    r = gsdr.Records.from_lists([b"ACGTAC", b"ACGA", b"AC"], names=[b">a", b">b", b">c"])
    cl = np.array([0, 0, 2], dtype=np.int32)
    assert r.write_result(cl, str(tmp_path / "r.fasta")) == 2
    assert (tmp_path / "r.fasta").read_bytes() == b">a\nACGTAC\n>c\nAC\n"
    r.write_membership(np.array([2, 0, 1]), cl, str(tmp_path / "m.tsv"))
    assert (tmp_path / "m.tsv").read_text() == "0\t2\t0\t>a\n1\t0\t0\t>b\n2\t1\t2\t>c\n"


def run_cli(*args, cwd=None):
    return subprocess.run([gsdr.CLI_PATH, *args], capture_output=True, text=True, cwd=cwd, timeout=60)


def test_cli_usage_errors_exit_zero(tmp_path):
    # odd argument count, unknown option, t out of range, w out of range: usage, exit 0
    for args in (("t",), ("x", "1"), ("t", "0.7"), ("t", "1.0"), ("w", "3"), ("w", "9"), ("t", "abc"),
                 ("--drop", "2"), ("d", "1"), ("p", "1")):
        p = run_cli(*args, cwd=tmp_path)
        assert p.returncode == 0, args
        assert "use like this" in p.stdout, args
    p = run_cli("t", "0.5", cwd=tmp_path)
    assert "similarity out of range" in p.stdout
    p = run_cli("w", "12", cwd=tmp_path)
    assert "word length out of range" in p.stdout


@pytest.mark.skipif(HAS_GPU, reason="checks the no-device error path")
def test_cli_without_device_fails_with_message(tmp_path):
    (tmp_path / "in.fasta").write_text(">a\nACGT\n")
    p = run_cli("-i", "in.fasta", "-t", "0.9", cwd=tmp_path)
    assert p.returncode == 1
    assert "input file:\tin.fasta" in p.stdout and "similarity:\t0.9" in p.stdout
    assert "word length:\t5" in p.stdout
    assert "HIP" in p.stderr or "device" in p.stderr


def test_reference_result_subset_is_fixed_point(tmp_path):
    fa = os.path.join(GOLDEN, "result_subset.fasta")
    recs = O.read_fasta(fa)
    assert len(recs) == 200
    cl, perm, st = O.cluster_file(fa, 0.95, out_path=str(tmp_path / "o.fasta"))
    assert (cl == np.arange(len(recs))).all()
    assert st["aligned"] > 0 and st["accepted"] == 0
    # output = the same records, in length order
    names = sorted(n for n, _ in recs)
    out_names = sorted(l for l in (tmp_path / "o.fasta").read_bytes().split(b"\n") if l.startswith(b">"))
    assert out_names == names


def test_reference_result_subset_full_lcs(tmp_path):
    # the same known answer with the oracle's full (unbanded) LCS
    fa = os.path.join(GOLDEN, "result_subset.fasta")
    cl, _, _ = O.cluster_file(fa, 0.95, fast=False)
    assert (cl == np.arange(len(cl))).all()


# thresholds of the workloads plus the two k-map boundaries (float vs double)
CUTOFF_THRESHOLDS = (0.80, 0.85, 0.88, 0.90, 0.94, 0.95, 0.99)


def test_product_cutoffs_match_oracle_table():
    """The product's float32 cutoffs (createCutoff, func.cpp:339-374) over
    every raw length 0..32,768 at every threshold and word length, against
    the oracle's independently written expressions (which tests/test_oracle.py
    pins to exact rational arithmetic and the float32 exception sets)."""
    lengths = np.arange(0, 32769, dtype=np.int32)
    for t in CUTOFF_THRESHOLDS:
        for k in (0, 4, 5, 6, 7, 8):
            bc, wc = gsdr.cutoffs(lengths, t, k)
            kk = k or O.kmap(t)
            want = np.array([O.cutoffs(int(L), t, kk) for L in lengths], dtype=np.int32)
            assert np.array_equal(bc, want[:, 0]), (t, k)
            assert np.array_equal(wc, want[:, 1]), (t, k)
    # the float32 quirk: ceilf(100 * (1 - 0.95f)) is 6, so wc(100) = 95 - 36
    bc, wc = gsdr.cutoffs([100], 0.95)
    assert (bc[0], wc[0]) == (95, 59)


def test_parse_options_follows_check_option():
    """checkOption (func.cpp:6-57) through the ABI: defaults, letters, the
    dashed forms, k from a float threshold compared with double literals,
    and the reference's usage/range messages instead of exit(0)."""
    d = gsdr.parse_options([])
    assert (d["input"], d["output"], d["word_length"]) == (b"testData.fasta", b"result.fasta", 6)
    assert abs(d["threshold"] - 0.95) < 1e-7 and d["mode"] == gsdr.MODE_BATCHED
    o = gsdr.parse_options(["t", "0.9", "i", "in.fa", "o", "out.fa", "w", "8"])
    assert (o["input"], o["output"], o["word_length"]) == (b"in.fa", b"out.fa", 8)
    assert gsdr.parse_options(["-t", "0.88"])["word_length"] == 4  # 0.88f < 0.88 (double)
    assert gsdr.parse_options(["t", "0.94"])["word_length"] == 5
    assert gsdr.parse_options(["t", "0.97"])["word_length"] == 7
    assert gsdr.parse_options(["input", "x"])["input"] == b"x"  # the first letter decides (func.cpp:13)
    o = gsdr.parse_options(["--mode", "plain", "--membership", "m.tsv", "--drop", "1", "t", "0.8"])
    assert (o["mode"], o["membership"], o["drop_filters"], o["word_length"]) == (gsdr.MODE_PLAIN, b"m.tsv", 1, 4)
    usage = "use like this: cluster i inputFile t threshold"
    for argv, first in ((["t"], usage), (["x", "1"], usage), (["t", "0.7"], "similarity out of range"),
                        (["t", "1.0"], "similarity out of range"), (["w", "9"], "word length out of range"),
                        (["w", "3"], "word length out of range"), (["t", "abc"], usage),
                        (["--mode", "fast"], usage)):
        with pytest.raises(gsdr.GsdrError) as e:
            gsdr.parse_options(argv)
        assert e.value.status == gsdr.E_INVALID
        msg = gsdr.lib().gsdr_last_error().decode()
        assert msg.startswith(first) and msg.endswith(usage), (argv, msg)


def integration_example():
    """The C++ block of section 2 of the integration guide (the
    reference-side binding), as the guide's file holds it."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    sec = text[text.index("## 2."):text.index("## 3.")]
    return sec[sec.index("```cpp") + len("```cpp"):sec.index("```", sec.index("```cpp") + 6)]


def build_integration_example(tmp_path):
    src = tmp_path / "main.cpp"
    src.write_text(integration_example())
    exe = tmp_path / "cluster_abi"
    libdir = os.path.dirname(gsdr.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", str(src), "-I", os.path.join(ROOT, "include"),
                    "-L", libdir, "-lgsdr", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    return exe


def test_integration_example_builds(tmp_path):
    """The maintainer's binding compiles and links against libgsdr.so with
    only include/gsdr.h (no func.h, no SYCL), and keeps the reference's
    usage behaviour: message on stdout, exit status 0."""
    exe = build_integration_example(tmp_path)
    for argv, want in ((["t", "0.5"], "similarity out of range"), (["t"], "use like this")):
        r = subprocess.run([str(exe)] + argv, capture_output=True, text=True, timeout=60)
        assert r.returncode == 0 and want in r.stdout, (argv, r.stdout, r.stderr)
