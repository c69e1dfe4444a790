"""Soundness of the exact pigeonhole prefilter (CPU).

pigeon_kernel (csrc/kernels_impl.hpp) rejects a pair when none of the E+1
disjoint segments of the shorter sequence (E = n + m - 2c, c = bc + 1) occurs
verbatim in the longer one at a diagonal inside [-(|x| - c), |y| - c]. This is
the same test written over Python strings with the kernel's segment length
(min(32, |x| // (E+1))) and gate (segments of at least 8 bases,
(E+1)^2 <= 4|x|); a reject must imply LCS < c by the oracle's LCS, over
random, mutated and near-threshold pairs, and the test must reject often
enough to matter.
"""
import numpy as np

import oracle_ffi as O

ACGT = np.frombuffer(b"ACGT", dtype=np.uint8)


def gap_free(seq):
    return bytes(ch for ch in seq if O.codes(bytes([ch]))[0] < 4)


def pigeon_rejects(a, b, c):
    """None when the pair is not examined (gate), else True = reject."""
    n, m = len(a), len(b)
    if c > min(n, m):
        return None
    x, y = (a, b) if n <= m else (b, a)
    mx, ny = len(x), len(y)
    segs = n + m - 2 * c + 1
    seg = min(32, mx // segs)
    if seg < 8 or segs * segs > 4 * mx:
        return None
    dlo = -(mx - c)
    for j in range(segs):
        ax = j * seg
        piece = x[ax:ax + seg]
        for dd in range(segs):
            by = ax + dlo + dd
            if 0 <= by and by + seg <= ny and y[by:by + seg] == piece:
                return False
    return True


def mutate(rng, s, d):
    out = bytearray()
    for ch in s:
        u = rng.random()
        if u < 0.1 * d:
            continue
        out.append(int(ACGT[rng.integers(0, 4)]) if rng.random() < 0.8 * d else ch)
        if rng.random() < 0.1 * d:
            out.append(int(ACGT[rng.integers(0, 4)]))
    return bytes(out) or b"A"


def test_pigeonhole_reject_implies_lcs_below_cutoff():
    rng = np.random.default_rng(2024)
    examined = rejected = 0
    for i in range(400):
        n = int(rng.integers(200, 1500))
        a = ACGT[rng.integers(0, 4, n)].tobytes()
        d = float(rng.choice([0.0, 0.01, 0.02, 0.04, 0.06, 0.1, 0.3]))
        b = mutate(rng, a, d) if i % 5 else ACGT[rng.integers(0, 4, n)].tobytes()
        ga, gb = gap_free(a), gap_free(b)
        L = O.lcs(ga, gb)
        # cutoffs around the true LCS and at the thresholds' values
        for c in {L - 2, L - 1, L, L + 1, L + 3, int(np.ceil(0.99 * len(gb))) + 1, int(np.ceil(0.95 * len(gb))) + 1}:
            if c < 1:
                continue
            r = pigeon_rejects(ga, gb, c)
            if r is None:
                continue
            examined += 1
            if r:
                rejected += 1
                assert L < c, (n, len(gb), L, c)
    # the gate lets enough pairs through, and the test rejects some of them
    assert examined > 500 and rejected > 100, (examined, rejected)


def test_pigeonhole_keeps_the_boundary():
    """At c = LCS the pair is accepted by the alignment, so the pigeonhole
    test must never reject it (a clean segment exists on the optimal path)."""
    rng = np.random.default_rng(7)
    kept = 0
    for _ in range(300):
        n = int(rng.integers(300, 1200))
        a = ACGT[rng.integers(0, 4, n)].tobytes()
        b = mutate(rng, a, float(rng.uniform(0, 0.03)))
        L = O.lcs(a, b)
        r = pigeon_rejects(a, b, L)
        if r is not None:
            assert r is False
            kept += 1
    assert kept > 100
