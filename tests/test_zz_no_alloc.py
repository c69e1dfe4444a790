"""gsdr_cluster_noalloc against gsdr_cluster through the C ABI.

The no-allocation entry runs the same clustering in the library's fixed
device scratch array. Two calls in a row (the second on a smaller input, so
stale contents of the shared array from the first call would show), in both
forms, must equal the allocating entry's cluster[] exactly; the small input
is also checked against the oracle. Inputs beyond the stated limit and
timed calls are refused with clear errors.
"""
import numpy as np
import pytest

import gsdr
import oracle_ffi as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [gsdr.MODE_BATCHED, gsdr.MODE_PLAIN], ids=["batched", "plain"])
def test_noalloc_equals_cluster_twice(engine, tmp_path, mode):
    inputs = [(2, 3101, 20000, 0.90), (1, 3102, 3000, 0.95)]  # the second call is smaller
    for law, seed, n, t in inputs:
        recs = gsdr.Records.generate(law, seed, n)
        fa = str(tmp_path / f"in_{seed}.fasta")
        recs.write_fasta(fa)
        recs.sort()
        k = gsdr.auto_word_length(t)
        store = engine.load(recs, k)
        try:
            print(f"\n# noalloc mode={mode} n={n} t={t}", flush=True)
            want, st_want = store.cluster(t, k, mode=mode)
            got, st_got = store.cluster(t, k, mode=mode, noalloc=True)
            assert np.array_equal(got, want), f"{np.count_nonzero(got != want)} assignments differ"
            assert st_got.reps == st_want.reps and st_got.aligned == st_want.aligned
            assert st_got.pairs == st_want.pairs and st_got.accepted == st_want.accepted
            if n <= 3000:
                ocl, _, _ = O.cluster_file(fa, t)
                assert np.array_equal(got, ocl)
        finally:
            store.free()


def test_noalloc_refuses_timing_and_oversized_inputs(engine):
    recs = gsdr.Records.generate(1, 3103, 800)
    recs.sort()
    store = engine.load(recs, 6)
    try:
        with pytest.raises(gsdr.GsdrError) as e:
            store.cluster(0.95, 6, mode=gsdr.MODE_BATCHED, timing=True, noalloc=True)
        assert e.value.status == gsdr.E_INVALID
        # the entry still works after a refused call
        a, _ = store.cluster(0.95, 6, mode=gsdr.MODE_BATCHED, noalloc=True)
        b, _ = store.cluster(0.95, 6, mode=gsdr.MODE_BATCHED)
        assert np.array_equal(a, b)
    finally:
        store.free()
    # one record more than the limit (short records keep the store small)
    n = 600001
    big = gsdr.Records.from_lists([b"ACGTACGTAC"] * n)
    big.sort()
    store = engine.load(big, 6)
    try:
        with pytest.raises(gsdr.GsdrError) as e:
            store.cluster(0.95, 6, mode=gsdr.MODE_PLAIN, noalloc=True)
        assert e.value.status == gsdr.E_LIMIT
        assert b"600,000" in gsdr.lib().gsdr_last_error()
    finally:
        store.free()
