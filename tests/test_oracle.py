"""Pins the CPU oracle against checks that do not restate its code.

- cutoffs: float32 quirk counts measured independently against exact decimal
  arithmetic; the k boundaries at t = 0.88 / 0.94;
- shared k-mers: brute-force multiset intersection over raw windows;
- LCS: the textbook O(nm) dynamic programme;
- banded early-exit decision: the full LCS;
- length order: std::sort over records holding std::string data, against the
  product's (length, index) key sort;
- FASTA split: hand-built edge cases;
- greedy recursion: a brute-force Python restatement on small inputs.
"""
import math
from collections import Counter
from fractions import Fraction

import numpy as np
import pytest

import oracle_ffi as O

ACGT = np.frombuffer(b"ACGT", dtype=np.uint8)


def dp_lcs(a, b):
    """Textbook LCS DP over code strings (codes 4 removed first)."""
    a = [c for c in O.codes(a) if c < 4]
    b = np.array([c for c in O.codes(b) if c < 4], dtype=np.int16)
    prev = np.zeros(len(b) + 1, dtype=np.int32)
    for x in a:
        cur = np.zeros_like(prev)
        match = prev[:-1] + (b == x)
        # cur[j] = max(prev[j], cur[j-1], prev[j-1] + eq)
        best = np.maximum(prev[1:], match)
        cur[1:] = np.maximum.accumulate(best)
        prev = cur
    return int(prev[-1])


def rand_seq(rng, n, alphabet=b"ACGT"):
    al = np.frombuffer(alphabet, dtype=np.uint8)
    return al[rng.integers(0, len(al), n)].tobytes()


# ----------------------------------------------------------------- cutoffs
def quirk_set(t_str):
    """Lengths L <= 32768 where float32 ceil(L*(1-t)) differs from exact decimal ceil."""
    om = np.float32(1.0) - np.float32(t_str)
    out = set()
    for L in range(1, 32769):
        f = math.ceil(float(np.float32(L) * om))
        exact = math.ceil(Fraction(L) * (1 - Fraction(t_str)))
        assert f - exact in (0, 1)
        if f != exact:
            out.add(L)
    return out


@pytest.mark.parametrize("t_str,expected", [("0.95", 1638), ("0.90", 3276), ("0.88", 622), ("0.94", 309),
                                            ("0.80", 0), ("0.85", 0), ("0.99", 0)])
def test_word_cutoff_float_quirks(t_str, expected):
    # float32(1-t) lies above 1-t for 0.95/0.90/0.88/0.94 and below it for 0.80/0.85/0.99
    q = quirk_set(t_str)
    assert len(q) == expected
    if t_str == "0.95":
        assert q == set(range(20, 32769, 20))  # every L with L/20 integral lands just above it
    if t_str == "0.90":
        assert q == set(range(10, 32769, 10))
    # the oracle's shared-word cutoff is the decimal one, one step of k higher exactly on the quirk set
    t = float(np.float32(t_str))
    k = O.kmap(t)
    for L in range(1, 32769):
        exact = math.ceil(Fraction(L) * (1 - Fraction(t_str))) + (1 if L in q else 0)
        assert O.cutoffs(L, t, k)[1] == max(L - k + 1 - exact * k, 1), L


def test_word_cutoff_known_value():
    # L=100, t=0.95: float32 gives ceil(5.0000012) = 6 -> 100-6+1-36 = 59 (decimal would give 65)
    assert O.cutoffs(100, 0.95, 6)[1] == 59


@pytest.mark.parametrize("t_str", ["0.80", "0.85", "0.90", "0.95", "0.99", "0.88", "0.94"])
def test_base_cutoff_matches_decimal_ceil(t_str):
    t = float(np.float32(t_str))
    k = O.kmap(t)
    for L in range(1, 32769, 7):
        bc, _ = O.cutoffs(L, t, k)
        assert bc == math.ceil(Fraction(L) * Fraction(t_str)), (L, t_str)


def test_word_length_map():
    cases = {0.8: 4, 0.85: 4, 0.87: 4, 0.88: 4, 0.8800001: 5, 0.9: 5, 0.93: 5, 0.94: 5, 0.9400001: 6,
             0.95: 6, 0.96: 6, 0.97: 7, 0.99: 7}
    for t, k in cases.items():
        assert O.kmap(t) == k, t


def test_cutoff_floor_at_one():
    # wc is clamped at 1 (func.cpp:357)
    for L in (4, 10, 50):
        assert O.cutoffs(L, 0.8, 4)[1] == max(L - 3 - math.ceil(float(np.float32(L) * (np.float32(1) - np.float32(0.8)))) * 4, 1)
        assert O.cutoffs(L, 0.8, 4)[1] >= 1


# ----------------------------------------------------------- shared k-mers
def kmer_counter(seq, k):
    c = O.codes(seq)
    out = Counter()
    for i in range(k - 1, len(c)):
        w = c[i - k + 1:i + 1]
        if (w < 4).all():
            out[bytes(w)] += 1
    return out


@pytest.mark.parametrize("k", [4, 5, 6, 7, 8])
def test_summin_brute_force(k):
    rng = np.random.default_rng(k)
    for _ in range(25):
        a = rand_seq(rng, int(rng.integers(0, 400)), b"ACGTACGTACGTN")
        b = rand_seq(rng, int(rng.integers(0, 400)), b"ACGTACGTACGTRN")
        ca, cb = kmer_counter(a, k), kmer_counter(b, k)
        expect = sum(min(v, cb[w]) for w, v in ca.items())
        assert O.summin(a, b, k) == expect


def test_summin_lowercase_and_u():
    assert O.summin(b"acgtuACGU", b"ACGTTACGT", 4) == O.summin(b"ACGTTACGT", b"ACGTTACGT", 4)


@pytest.mark.parametrize("k", [4, 5, 6, 7, 8])
def test_shared_words_early_stop_matches_full_scan(k):
    """The exact deficit stop of the shared-word scan (fast mode) decides like
    the full scan (func.cpp:552-567) at the cutoffs around the true shared
    sum S: wc = S (pass) and S + 1 (fail) are the cases a wrong bound would
    flip; wc above the candidate's total (no scan at all) and far below it
    are covered too."""
    rng = np.random.default_rng(100 + k)
    checked = 0
    for i in range(60):
        n = int(rng.integers(1, 1500))
        a = rand_seq(rng, n, b"ACGTACGTACGTN" if i % 4 == 0 else b"ACGT")
        kind = i % 3
        if kind == 0:
            b = rand_seq(rng, int(rng.integers(1, n + 1)))
        else:
            b = bytearray(a[int(rng.integers(0, max(1, n // 4))):])
            for _ in range(int(len(b) * rng.uniform(0, 0.1))):
                b[int(rng.integers(0, len(b)))] = int(ACGT[rng.integers(0, 4)])
            b = bytes(b) or b"A"
        S = O.summin(a, b, k)
        total = sum(kmer_counter(b, k).values())
        for wc in {1, max(1, S - 1), max(1, S), S + 1, S + 5, total, total + 1}:
            full = O.shared_words_pass(a, b, k, wc, early=False)
            assert full == (S >= wc)
            assert O.shared_words_pass(a, b, k, wc, early=True) == full, (k, n, len(b), S, total, wc)
            checked += 1
    assert checked > 300


def test_fast_mode_clusters_like_reference_mode(tmp_path):
    """The oracle's fast mode (banded early-exit alignment, deficit stop in
    the shared-word scan) against its reference mode (full LCS, full scan) on
    whole clusterings, including a t=0.99 job where the stop fires often."""
    import gsdr

    for law, seed, n, t in ((2, 41, 1500, 0.99), (1, 42, 1200, 0.95), (2, 43, 1200, 0.90)):
        fa = str(tmp_path / f"in{seed}.fasta")
        gsdr.Records.generate(law, seed, n).write_fasta(fa)
        fast, _, sf = O.cluster_file(fa, t, threads=4, fast=True)
        ref, _, sr = O.cluster_file(fa, t, threads=4, fast=False)
        assert np.array_equal(fast, ref), (law, t)
        assert sf["rej_word"] == sr["rej_word"] and sf["aligned"] == sr["aligned"]


# --------------------------------------------------------------------- LCS
def test_lcs_random_small():
    rng = np.random.default_rng(1)
    for _ in range(300):
        a = rand_seq(rng, int(rng.integers(0, 150)))
        b = rand_seq(rng, int(rng.integers(0, 150)))
        assert O.lcs(a, b) == dp_lcs(a, b)


@pytest.mark.parametrize("n", [31, 32, 33, 63, 64, 65, 127, 128, 129, 255, 256, 257])
def test_lcs_word_boundaries(n):
    rng = np.random.default_rng(n)
    for m in (1, 31, 32, 33, 64, 65, n):
        a = rand_seq(rng, n)
        b = rand_seq(rng, m)
        assert O.lcs(a, b) == dp_lcs(a, b)
        assert O.lcs(b, a) == dp_lcs(a, b)


def test_lcs_adversarial():
    cases = [(b"A" * 200, b"A" * 150), (b"AC" * 100, b"CA" * 100), (b"ACGT" * 50, b"TGCA" * 50),
             (b"A" * 100, b"C" * 100), (b"", b"ACGT"), (b"NNNN", b"ACGT"), (b"ANCNGNT", b"ACGT"),
             (b"AAAAAAAAAC" * 20, b"CAAAAAAAAA" * 20)]
    for a, b in cases:
        assert O.lcs(a, b) == dp_lcs(a, b)


def test_lcs_medium():
    rng = np.random.default_rng(3)
    for n, m in ((1000, 1000), (1500, 300), (700, 1300)):
        a = rand_seq(rng, n)
        b = bytearray(a[: m] if m <= n else a + rand_seq(rng, m - n))
        for _ in range(m // 10):
            b[int(rng.integers(0, len(b)))] = int(ACGT[rng.integers(0, 4)])
        assert O.lcs(a, bytes(b)) == dp_lcs(a, bytes(b))


def test_decide_fast_matches_full():
    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(150):
        n = int(rng.integers(1, 3000))
        a = rand_seq(rng, n)
        kind = rng.integers(0, 3)
        if kind == 0:
            b = rand_seq(rng, int(rng.integers(1, n + 1)))
        else:
            b = bytearray(a[: int(rng.integers(1, n + 1))])
            for _ in range(int(len(b) * rng.uniform(0, 0.15))):
                b[int(rng.integers(0, len(b)))] = int(ACGT[rng.integers(0, 4)])
            b = bytes(b)
        L = O.lcs(a, b)
        for bc in (L - 2, L - 1, L, L + 1, int(math.ceil(len(b) * 0.9)), int(math.ceil(len(b) * 0.95))):
            if bc < 0:
                continue
            got, _ = O.decide_fast(a, b, bc)
            assert got == (L > bc), (n, len(b), L, bc)
            checked += 1
    assert checked > 600


# ------------------------------------------------------------ length order
def test_sort_order_matches_string_sort():
    import gsdr

    rng = np.random.default_rng(4)
    for trial in range(6):
        n = int(rng.integers(1, 3000))
        lens = rng.integers(0, 30 if trial % 2 else 2000, n)
        datas = [b"A" * int(L) for L in lens]
        r = gsdr.Records.from_lists(datas)
        perm = r.sort()
        assert np.array_equal(perm, O.sort_order(lens))
        assert all(lens[perm[i]] >= lens[perm[i + 1]] for i in range(n - 1))


# ------------------------------------------------------------- FASTA split
EDGE_FILES = {
    "crlf": b">a\r\nACGT\r\nAC\r\n>b\r\nGG\r\n",
    "blank": b">a\nAC\n\nGT\n>b\n\nTT\n",
    "no_trailing_newline": b">a\nACGT\n>b\nGGCC",
    "header_only_middle": b">a\n>b\nACGT\n",
    "header_last": b">a\nACGT\n>b\n",
    "first_line_data": b"ACGT\nGGGG\n>b\nTT\n",
    "iupac": b">x\nACGTNRYKMSWBDHV\nacgtu\n",
    "gt_inside": b">a\nAC>GT\n",
}
EDGE_EXPECT = {
    "crlf": [(b">a\r", b"ACGT\rAC\r"), (b">b\r", b"GG\r")],
    "blank": [(b">a", b"ACGT"), (b">b", b"TT")],
    "no_trailing_newline": [(b">a", b"ACGT"), (b">b", b"GGCC")],
    "header_only_middle": [(b">a", b""), (b">b", b"ACGT")],
    "header_last": [(b">a", b"ACGT"), (b">b", b"")],
    "first_line_data": [(b"ACGT", b"GGGG"), (b">b", b"TT")],
    "iupac": [(b">x", b"ACGTNRYKMSWBDHVacgtu")],
    "gt_inside": [(b">a", b"AC>GT")],
}


@pytest.mark.parametrize("name", sorted(EDGE_FILES))
def test_fasta_edge_cases(tmp_path, name):
    import gsdr

    p = tmp_path / (name + ".fasta")
    p.write_bytes(EDGE_FILES[name])
    assert O.read_fasta(str(p)) == EDGE_EXPECT[name]
    r = gsdr.Records.read_fasta(str(p))
    assert [r.get(i) for i in range(len(r))] == EDGE_EXPECT[name]
    expect = EDGE_EXPECT[name]
    order = O.sort_order([len(d) for _, d in expect])
    s = gsdr.Records.read_fasta(str(p), length_order=True)
    assert [s.get(i) for i in range(len(s))] == [expect[i] for i in order]
    assert np.array_equal(s.sort(), order)


# ---------------------------------------------------------- greedy driver
def brute_cluster(seqs, t, drop=False):
    """Greedy recursion of func.cpp:466-784 restated from the definitions
    (drop: without the composition and shared-word filters)."""
    t32 = np.float32(t)
    k = 4 if t32 < 0.88 else 5 if t32 < 0.94 else 6 if t32 < 0.97 else 7
    order = O.sort_order([len(s) for s in seqs])
    S = [seqs[i] for i in order]
    n = len(S)
    comp = [np.bincount(O.codes(s), minlength=5)[:4] for s in S]
    kc = [kmer_counter(s, k) for s in S]
    cut = []
    for s in S:
        L = len(s)
        bc = math.ceil(float(np.float32(L) * t32))
        wc = max(L - k + 1 - math.ceil(float(np.float32(L) * (np.float32(1) - t32))) * k, 1)
        cut.append((bc, wc))

    def P(r, q):
        bc, wc = cut[q]
        if drop:
            return dp_lcs(S[r], S[q]) > bc
        if int(np.minimum(comp[r], comp[q]).sum()) < bc:
            return False
        if sum(min(v, kc[r][w]) for w, v in kc[q].items()) < wc:
            return False
        return dp_lcs(S[r], S[q]) > bc

    cluster = [-1] * n
    for q in range(n):
        for r in range(q):
            if cluster[r] == r and P(r, q):
                cluster[q] = r
                break
        if cluster[q] == -1:
            cluster[q] = q
    return np.array(cluster), order


def family_set(seed, parents, kids, lo, hi, div):
    rng = np.random.default_rng(seed)
    seqs = []
    for _ in range(parents):
        p = bytearray(rand_seq(rng, int(rng.integers(lo, hi))))
        seqs.append(bytes(p))
        for _ in range(kids):
            c = bytearray(p)
            for _ in range(int(len(c) * rng.uniform(0, div))):
                op = rng.integers(0, 10)
                i = int(rng.integers(0, len(c)))
                if op < 8:
                    c[i] = int(ACGT[rng.integers(0, 4)])
                elif op == 8 and len(c) > 1:
                    del c[i]
                else:
                    c.insert(i, int(ACGT[rng.integers(0, 4)]))
            seqs.append(bytes(c))
    rng.shuffle(seqs)
    return seqs


@pytest.mark.parametrize("t", [0.95, 0.9, 0.8])
def test_greedy_brute_force(tmp_path, t):
    seqs = family_set(int(t * 100), 6, 5, 40, 160, 0.12)
    fa = tmp_path / "f.fasta"
    fa.write_bytes(b"".join(b">s%d\n%s\n" % (i, s) for i, s in enumerate(seqs)))
    expect, order = brute_cluster(seqs, t)
    for fast in (True, False):
        cl, perm, _ = O.cluster_file(str(fa), t, fast=fast)
        assert np.array_equal(perm, order)
        assert np.array_equal(cl, expect)


@pytest.mark.parametrize("t", [0.95, 0.8])
def test_greedy_brute_force_drop_mode(tmp_path, t):
    """Drop mode: the shared-word filter is not sound for LCS, so leaving it
    out can absorb more; the recursion with A alone pins the oracle's mode."""
    seqs = family_set(int(t * 100) + 7, 6, 5, 40, 160, 0.12)
    fa = tmp_path / "f.fasta"
    fa.write_bytes(b"".join(b">s%d\n%s\n" % (i, s) for i, s in enumerate(seqs)))
    expect, order = brute_cluster(seqs, t, drop=True)
    for fast in (True, False):
        cl, perm, st = O.cluster_file(str(fa), t, fast=fast, drop=True)
        assert np.array_equal(perm, order)
        assert np.array_equal(cl, expect)
        assert st["rej_comp"] == 0 and st["rej_word"] == 0
