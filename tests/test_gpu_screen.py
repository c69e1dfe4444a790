"""The word-signature screen on the GPU (word lengths >= 7): with the screen
on (the default) and off, both forms give the oracle's clusters and the same
stage counts (every pair the screen rejects is one the shared-word test
rejects), and the screen decides pairs when it is on and t >= 0.97."""
import numpy as np
import pytest

import gsdr
import oracle_ffi as O

pytestmark = pytest.mark.gpu


def with_screen(engine, on):
    t = engine.tuning()
    t.screen = 1 if on else 0
    engine.set_tuning(t)


STAGES = ("pairs", "rej_length", "rej_comp", "rej_word", "aligned", "accepted", "rej_align")


@pytest.mark.parametrize("law,seed,n,t,k", [(2, 81, 3000, 0.99, 0), (1, 82, 3000, 0.95, 7), (4, 83, 4000, 0.97, 0),
                                            (2, 84, 3000, 0.97, 0), (1, 85, 2500, 0.99, 8), (2, 86, 2000, 0.90, 7),
                                            (1, 87, 2000, 0.95, 0)])
def test_screen_on_off_match_oracle(engine, tmp_path, law, seed, n, t, k):
    recs = gsdr.Records.generate(law, seed, n)
    fa = str(tmp_path / "in.fasta")
    recs.write_fasta(fa)
    ocl, operm, _ = O.cluster_file(fa, t, k=k)
    assert np.array_equal(recs.sort(), operm)
    k = k or gsdr.auto_word_length(t)
    store = engine.load(recs, k)
    try:
        for mode in (gsdr.MODE_PLAIN, gsdr.MODE_BATCHED):
            seen = {}
            for on in (False, True):
                print(f"\n# screen={on} law={law} n={n} t={t} k={k} mode={mode}", flush=True)
                with_screen(engine, on)
                cl, st = store.cluster(t, k, mode=mode, tune=True)
                assert np.array_equal(cl, ocl), f"{np.count_nonzero(cl != ocl)} assignments differ"
                assert st.pairs == st.rej_length + st.rej_comp + st.rej_word + st.aligned
                assert st.screened <= st.rej_word
                if not on or k < 7:
                    assert st.screened == 0
                elif t >= 0.97:
                    # (at lower t the slack is often too wide for the bound)
                    assert st.screened > 0, st.screened
                seen[on] = st
            if mode == gsdr.MODE_PLAIN:
                # one representative per round: the pairs are the same with or
                # without the screen (the batched form's speculation may differ
                # only in the order the device finishes, never in pairs)
                for f in STAGES:
                    assert getattr(seen[True], f) == getattr(seen[False], f), f
    finally:
        engine.set_tuning(None)
        store.free()


def test_screen_in_a_mixed_batch(engine, tmp_path):
    """Jobs with and without signatures (k = 5 and k = 7) and a drop-mode
    job in one batched call: every job equals the oracle's."""
    specs = [(2, 91, 2000, 0.99, False), (1, 92, 1500, 0.90, False), (2, 93, 1800, 0.97, True),
             (4, 94, 2500, 0.95, False)]
    stores, ts, expect, drops = [], [], [], []
    try:
        for law, seed, n, t, drop in specs:
            recs = gsdr.Records.generate(law, seed, n)
            fa = str(tmp_path / f"j{seed}.fasta")
            recs.write_fasta(fa)
            ocl, _, _ = O.cluster_file(fa, t, drop=drop)
            recs.sort()
            stores.append(engine.load(recs, gsdr.auto_word_length(t)))
            ts.append(t)
            expect.append(ocl)
            drops.append(drop)
        print("\n# screen mixed batch", flush=True)
        outs, st = engine.cluster_jobs(stores, ts, drop_filters=drops)
        for j, (cl, ocl) in enumerate(zip(outs, expect)):
            assert np.array_equal(cl, ocl), f"job {j}: {np.count_nonzero(cl != ocl)} assignments differ"
        assert st.screened > 0
    finally:
        for s in stores:
            s.free()
