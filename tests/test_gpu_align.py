"""Device LCS decision (gsdr_align_pairs) against the oracle's full bit-vector LCS.

Every pair is tested at the exact decision boundary: with L = LCS(text, query)
from the oracle, bc = L-1 must accept and bc = L must reject (the reference
accepts iff LCS > bc, func.cpp:670-671), plus a few cutoffs off the boundary.
"""
import numpy as np
import pytest

import oracle_ffi as O

pytestmark = pytest.mark.gpu

ACGT = np.frombuffer(b"ACGT", dtype=np.uint8)


def rand_seq(rng, n):
    return ACGT[rng.integers(0, 4, n)].tobytes()


def mutate(rng, s, d):
    """Substitutions/indels at total rate d (80/20)."""
    out = bytearray()
    for ch in s:
        u = rng.random()
        if u < 0.1 * d:
            continue  # deletion
        if rng.random() < 0.8 * d:
            out.append(int(ACGT[(int(np.searchsorted(ACGT, ch)) + 1 + rng.integers(0, 3)) & 3]))
        else:
            out.append(ch)
        if rng.random() < 0.1 * d:
            out.append(int(ACGT[rng.integers(0, 4)]))
    return bytes(out)


def make_pairs(seed, count, max_len):
    rng = np.random.default_rng(seed)
    seqs, pairs = [], []
    for i in range(count):
        n = int(rng.choice([1, 2, 31, 32, 33, 63, 64, 65, 100, 257, 1000, 1500, 2048, 3000])) if i % 3 == 0 \
            else int(rng.integers(1, max_len))
        text = rand_seq(rng, n)
        kind = rng.integers(0, 4)
        if kind == 0:
            q = rand_seq(rng, int(rng.integers(1, max(2, n + 1))))  # unrelated, shorter
        elif kind == 1:
            q = mutate(rng, text, float(rng.uniform(0, 0.15)))  # similar length
        elif kind == 2:
            a = int(rng.integers(0, n))
            q = mutate(rng, text[a:a + int(rng.integers(1, n - a + 1))], float(rng.uniform(0, 0.1)))  # substring
        else:
            q = rand_seq(rng, int(rng.integers(1, max(2, n // 5 + 2))))  # much shorter
        if len(q) == 0:
            q = b"A"
        # sprinkle gap codes (N and IUPAC) that the loader removes
        if rng.random() < 0.2 and len(q) > 4:
            qb = bytearray(q)
            for _ in range(int(rng.integers(1, 4))):
                qb[int(rng.integers(0, len(qb)))] = int(rng.choice(list(b"NRY")))
            q = bytes(qb)
        seqs.append(text)
        seqs.append(q)
        pairs.append((2 * i, 2 * i + 1))
    return seqs, pairs


def boundary_cases(seqs, pairs):
    exp_pairs, bcs, expect = [], [], []
    for (t, q) in pairs:
        L = O.lcs(seqs[t], seqs[q])
        for bc in (L - 1, L, L - 3, L + 2):
            if bc < 0:
                continue
            exp_pairs.append((t, q))
            bcs.append(bc)
            expect.append(L > bc)
    return exp_pairs, np.array(bcs, dtype=np.int32), np.array(expect, dtype=bool)


@pytest.mark.parametrize("seed,count,max_len", [(1, 300, 600), (2, 200, 4000), (3, 40, 12000)])
def test_boundary_decisions(engine, seed, count, max_len):
    seqs, pairs = make_pairs(seed, count, max_len)
    ep, bcs, expect = boundary_cases(seqs, pairs)
    got, cells = engine.align_pairs(seqs, ep, bcs)
    bad = np.nonzero(got != expect)[0]
    assert len(bad) == 0, f"{len(bad)} mismatches, first: pair {ep[bad[0]]} bc {bcs[bad[0]]} expect {expect[bad[0]]}"
    assert cells > 0


@pytest.mark.parametrize("shape", list(range(12)))
def test_every_tile_shape(engine, shape):
    """Forces each tile shape (the smallest fitting shape at or above it)."""
    seqs, pairs = make_pairs(100 + shape, 60, 2500)
    ep, bcs, expect = boundary_cases(seqs, pairs)
    got, _ = engine.align_pairs(seqs, ep, bcs, shape=shape)
    assert np.array_equal(got, expect)


@pytest.mark.parametrize("shape", [4, 5, 7, 9, 11])
def test_single_base_runs(engine, shape):
    """Texts made of long single-base runs: whole words match nothing for most
    query bases, so row sums are all ones and carries ripple through every word
    of a lane and across lanes (the rarely taken ripple paths of the 64-lane
    tiles)."""
    rng = np.random.default_rng(40 + shape)
    seqs, pairs = [], []
    for i in range(10):
        runs, n = [], int(rng.integers(1500, 6000))
        while sum(len(r) for r in runs) < n:
            runs.append(bytes([int(rng.choice(ACGT))]) * int(rng.integers(33, 260)))
        text = b"".join(runs)
        if i % 3 == 0:
            q = rand_seq(rng, int(rng.integers(200, len(text) // 2)))
        elif i % 3 == 1:
            q = mutate(rng, text, 0.05)
        else:
            q = text[: len(text) // 3] + rand_seq(rng, 400) + text[len(text) // 2:]
        seqs += [text, q]
        pairs.append((2 * i, 2 * i + 1))
    ep, bcs, expect = boundary_cases(seqs, pairs)
    got, _ = engine.align_pairs(seqs, ep, bcs, shape=shape)
    assert np.array_equal(got, expect)


@pytest.mark.parametrize("shape,cap", [(0, 255), (1, 511), (2, 1023), (3, 2047), (6, 8000)])
def test_single_base_runs_narrow(engine, shape, cap):
    """The same single-base-run texts sized so that the one-word tiles (8, 16,
    32 and 64 lanes of one word) and the 64x4 tile hold them: all-ones sums
    and carry chains across lanes and segment boundaries in every row, with
    queries ending inside 32-row blocks."""
    rng = np.random.default_rng(70 + shape)
    seqs, pairs = [], []
    for i in range(16):
        runs, n = [], int(rng.integers(cap // 3, cap))
        while sum(len(r) for r in runs) < n:
            runs.append(bytes([int(rng.choice(ACGT))]) * int(rng.integers(33, 200)))
        text = b"".join(runs)[:n]
        if i % 4 == 0:
            q = rand_seq(rng, int(rng.integers(20, max(21, n // 2))))
        elif i % 4 == 1:
            q = mutate(rng, text, 0.05)
        elif i % 4 == 2:
            q = text[: n // 3] + rand_seq(rng, 40) + text[n // 2:]
        else:
            q = bytes(reversed(text))
        seqs += [text, q]
        pairs.append((2 * i, 2 * i + 1))
    ep, bcs, expect = boundary_cases(seqs, pairs)
    got, _ = engine.align_pairs(seqs, ep, bcs, shape=shape)
    assert np.array_equal(got, expect)


def test_near_threshold_long(engine):
    """Long similar pairs at t=0.95/0.90 cutoffs and the 32,767-base maximum."""
    rng = np.random.default_rng(7)
    seqs, ep, bcs, expect = [], [], [], []
    for n in (32767, 30000, 20000, 8191):
        text = rand_seq(rng, n)
        for d in (0.02, 0.06, 0.12):
            q = mutate(rng, text, d)
            L = O.lcs(text, q)
            seqs += [text, q]
            i = len(seqs) - 2
            for bc in (L - 1, L, int(np.ceil(np.float32(len(q)) * np.float32(0.95))),
                       int(np.ceil(np.float32(len(q)) * np.float32(0.9)))):
                if bc >= 0:
                    ep.append((i, i + 1))
                    bcs.append(bc)
                    expect.append(L > bc)
    got, _ = engine.align_pairs(seqs, ep, np.array(bcs, dtype=np.int32))
    assert np.array_equal(got, np.array(expect))


def test_empty_and_degenerate(engine):
    seqs = [b"ACGT", b"NNNN", b"A", b"ACGTACGT" * 8, b"T" * 64]
    pairs = [(0, 1), (1, 0), (2, 2), (3, 4), (4, 3), (3, 3)]
    expect_l = [O.lcs(seqs[a], seqs[b]) for a, b in pairs]
    for off in (-1, 0):
        bcs = np.array([max(0, L + off) for L in expect_l], dtype=np.int32)
        got, _ = engine.align_pairs(seqs, pairs, bcs)
        assert list(got) == [L > b for L, b in zip(expect_l, bcs)]


# split pairs (gsdr_align_pairs shape 12 + s): two waves, the first half of
# the rows forward and the second half backward on the reversed pair, joined
# at the middle row; pairs with fewer than 64 query bases run unsplit


@pytest.mark.parametrize("shape", [0, 1, 2, 3, 4, 5, 6, 7, 9, 11])
def test_split_boundary_decisions(engine, shape):
    """Starting shapes 0..3 (one word per lane) start the halves from two
    words per lane: the pairs whose bands a one-word tile holds."""
    seqs, pairs = make_pairs(200 + shape, 80, 5000)
    ep, bcs, expect = boundary_cases(seqs, pairs)
    got, _ = engine.align_pairs(seqs, ep, bcs, shape=12 + shape)
    bad = np.nonzero(got != expect)[0]
    assert len(bad) == 0, f"{len(bad)} mismatches, first: pair {ep[bad[0]]} bc {bcs[bad[0]]} expect {expect[bad[0]]}"


@pytest.mark.parametrize("shape", [0, 4, 8])
def test_split_single_base_runs(engine, shape):
    """Single-base-run texts through both halves (carries ripple through every
    word on the reversed pair too), queries of every length class."""
    rng = np.random.default_rng(90 + shape)
    seqs, pairs = [], []
    for i in range(12):
        runs, n = [], int(rng.integers(200, 6000))
        while sum(len(r) for r in runs) < n:
            runs.append(bytes([int(rng.choice(ACGT))]) * int(rng.integers(33, 260)))
        text = b"".join(runs)[:n]
        if i % 4 == 0:
            q = rand_seq(rng, int(rng.integers(64, max(65, n // 2))))
        elif i % 4 == 1:
            q = mutate(rng, text, 0.05)
        elif i % 4 == 2:
            q = text[: n // 3] + rand_seq(rng, 40) + text[n // 2:]
        else:
            q = bytes(reversed(text))
        seqs += [text, q]
        pairs.append((2 * i, 2 * i + 1))
    ep, bcs, expect = boundary_cases(seqs, pairs)
    got, _ = engine.align_pairs(seqs, ep, bcs, shape=12 + shape)
    assert np.array_equal(got, expect)


def test_split_near_threshold_long(engine):
    """Split pairs up to the 32,767-base maximum, at the LCS boundary and at
    the t=0.95 / t=0.90 cutoffs, with query lengths around 32-row blocks."""
    rng = np.random.default_rng(17)
    seqs, ep, bcs, expect = [], [], [], []
    for n in (32767, 30000, 12345, 4096, 97, 64):
        text = rand_seq(rng, n)
        for d in (0.0, 0.02, 0.06, 0.12):
            q = mutate(rng, text, d) or b"A"
            L = O.lcs(text, q)
            seqs += [text, q]
            i = len(seqs) - 2
            for bc in (L - 1, L, int(np.ceil(np.float32(len(q)) * np.float32(0.95))),
                       int(np.ceil(np.float32(len(q)) * np.float32(0.9)))):
                if bc >= 0:
                    ep.append((i, i + 1))
                    bcs.append(bc)
                    expect.append(L > bc)
    got, _ = engine.align_pairs(seqs, ep, np.array(bcs, dtype=np.int32), shape=12 + 4)
    assert np.array_equal(got, np.array(expect))


def test_split_list_longer_than_the_grid(engine):
    """More split pairs than blocks: each wave aligns whole pairs instead."""
    seqs, pairs = make_pairs(300, 240, 2000)
    ep, bcs, expect = boundary_cases(seqs, pairs)
    assert len(ep) > 2 * 256 + 100
    got, _ = engine.align_pairs(seqs, ep, bcs, shape=12 + 4)
    assert np.array_equal(got, expect)
