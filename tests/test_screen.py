"""Soundness of the word-signature screen (CPU).

screen_kernel (csrc/kernels_impl.hpp) rejects a (representative r, candidate
q) pair when popcount(sig_q & ~sig_r) > T_q - wc_q, where sig_s has bit
sig_bucket(w) set for each distinct k-mer w of s (w itself for k <= 5, the top
11 bits of w * 0x9E3779B1 for longer words), T_q is the sum of q's k-mer
counts and wc_q its shared-word cutoff. That is sound when the bucket count
never exceeds the shared-word deficit T_q - sum_w min(c_r(w), c_q(w)), the
oracle's shared-word sum being orc_summin. Checked here over random, mutated
and near-identical pairs for k = 4..8, with the signature restated over
Python sets; a planted mistake (ignoring r's signature) must break it.
"""
import numpy as np

import oracle_ffi as O

ACGT = np.frombuffer(b"ACGT", dtype=np.uint8)
BITS = 2048


def words(seq, k):
    c = O.codes(seq).astype(np.int64)
    w = np.zeros(len(c) - k + 1, dtype=np.int64)
    for i in range(k):
        w = (w << 2) | c[i:len(c) - k + 1 + i]
    return set(int(x) for x in w)


def bucket(w, k):
    return w if k <= 5 else ((w * 0x9E3779B1) & 0xFFFFFFFF) >> 21


def signature(seq, k):
    return {bucket(w, k) for w in words(seq, k)}


def bound(q, r, k):
    return len(signature(q, k) - signature(r, k))


def mutate(rng, s, d):
    out = bytearray()
    for ch in s:
        if rng.random() < 0.1 * d:
            continue
        out.append(int(ACGT[rng.integers(0, 4)]) if rng.random() < 0.8 * d else ch)
        if rng.random() < 0.1 * d:
            out.append(int(ACGT[rng.integers(0, 4)]))
    return bytes(out) or b"A"


def pairs(rng, count):
    for i in range(count):
        n = int(rng.integers(40, 1600))
        r = ACGT[rng.integers(0, 4, n)].tobytes()
        kind = i % 4
        if kind == 0:
            q = ACGT[rng.integers(0, 4, int(rng.integers(20, n + 1)))].tobytes()
        elif kind == 1:
            q = mutate(rng, r, float(rng.choice([0.0, 0.01, 0.03, 0.08])))
        elif kind == 2:
            # a piece of r: every word of q is in r
            a = int(rng.integers(0, max(1, n - 20)))
            q = r[a:a + int(rng.integers(20, n - a + 1))]
        else:
            # low complexity: few distinct words, high counts
            unit = ACGT[rng.integers(0, 4, int(rng.integers(1, 7)))].tobytes()
            q = (unit * (n // len(unit) + 1))[: max(20, n // 2)]
        yield r, q


def test_signature_bound_below_shared_word_deficit():
    rng = np.random.default_rng(11)
    for k in (4, 5, 6, 7, 8):
        for r, q in pairs(rng, 60):
            if min(len(r), len(q)) < k:
                continue
            total = O.summin(q, q, k)  # T_q: the sum of q's counts
            deficit = total - O.summin(r, q, k)
            assert 0 <= bound(q, r, k) <= deficit, (k, len(r), len(q))


def test_screen_rejects_only_failing_pairs_and_rejects_most_random_pairs():
    rng = np.random.default_rng(12)
    rejected = random_rejected = random_pairs = 0
    for t in (0.94, 0.95, 0.97, 0.99):
        k = O.kmap(t)
        for i, (r, q) in enumerate(pairs(rng, 80)):
            if len(q) > len(r):
                r, q = q, r
            if len(q) < k:
                continue
            _, wc = O.cutoffs(len(q), t, k)
            slack = O.summin(q, q, k) - wc
            rej = bound(q, r, k) > slack
            if rej:
                rejected += 1
                assert O.summin(r, q, k) < wc, (t, len(r), len(q))
                assert not O.shared_words_pass(r, q, k, wc, False)
            if i % 4 == 0 and len(q) >= 300 and t >= 0.97:
                random_pairs += 1
                random_rejected += rej
    assert rejected > 50, rejected
    # unrelated sequences of a few hundred bases and more fail at t >= 0.97
    assert random_pairs > 10 and random_rejected >= 0.9 * random_pairs, (random_rejected, random_pairs)


def test_planted_mistake_is_caught():
    """Counting q's buckets without removing r's (a dropped term) exceeds the
    deficit of near-identical pairs: the inequality above would fail."""
    rng = np.random.default_rng(13)
    broken = 0
    for r, q in pairs(rng, 40):
        k = 7
        if min(len(r), len(q)) < 50:
            continue
        deficit = O.summin(q, q, k) - O.summin(r, q, k)
        broken += len(signature(q, k)) > deficit
    assert broken > 5, broken


def test_buckets_match_the_device_hash():
    """sig_bucket of the kernel header: identity below k = 6, else the top 11
    bits of the 32-bit product with 0x9E3779B1 (fixed values)."""
    assert bucket(1023, 5) == 1023
    assert bucket(0, 7) == 0
    assert bucket(1, 7) == 0x9E3779B1 >> 21
    assert bucket(65535, 8) == ((65535 * 0x9E3779B1) & 0xFFFFFFFF) >> 21
    assert all(0 <= bucket(w, 8) < BITS for w in range(0, 65536, 97))
