"""The pigeonhole prefilter on the GPU: with tuning.pigeon = 1 both forms give
the oracle's clusters, the prefilter examines and rejects pairs, and every
rejected pair is one the alignment would have rejected (the stage counts add
up)."""
import numpy as np
import pytest

import gsdr
import oracle_ffi as O

pytestmark = pytest.mark.gpu


@pytest.fixture
def pigeon_on(engine):
    t = engine.tuning()
    t.pigeon = 1
    engine.set_tuning(t)
    yield engine
    engine.set_tuning(None)


@pytest.mark.parametrize("law,seed,n,t", [(2, 61, 3000, 0.99), (1, 62, 3000, 0.95), (4, 63, 4000, 0.95),
                                          (2, 64, 3000, 0.97), (2, 65, 2500, 0.90)])
def test_pigeon_clusters_match_oracle(pigeon_on, tmp_path, law, seed, n, t):
    engine = pigeon_on
    recs = gsdr.Records.generate(law, seed, n)
    fa = str(tmp_path / "in.fasta")
    recs.write_fasta(fa)
    ocl, operm, _ = O.cluster_file(fa, t)
    assert np.array_equal(recs.sort(), operm)
    k = gsdr.auto_word_length(t)
    store = engine.load(recs, k)
    try:
        checked = rejected = 0
        for mode in (gsdr.MODE_PLAIN, gsdr.MODE_BATCHED):
            print(f"\n# pigeon law={law} n={n} t={t} mode={mode}", flush=True)
            cl, st = store.cluster(t, k, mode=mode, tune=True)
            assert np.array_equal(cl, ocl), f"{np.count_nonzero(cl != ocl)} assignments differ"
            assert st.pairs == st.rej_length + st.rej_comp + st.rej_word + st.aligned
            assert st.aligned == st.accepted + st.rej_align + int(st.spec_wasted) + int(st.rej_pigeon)
            assert st.rej_pigeon <= st.pigeon_checked
            checked += st.pigeon_checked
            rejected += st.rej_pigeon
        if t >= 0.97:
            assert checked > 0, checked
    finally:
        store.free()


@pytest.mark.parametrize("law,seed,n,t", [(2, 71, 2000, 0.99), (1, 72, 2000, 0.97)])
def test_pigeon_rejects_in_drop_mode(pigeon_on, tmp_path, law, seed, n, t):
    """Drop mode (no composition or shared-word filter) sends unrelated
    pairs of similar length to the alignment: the prefilter rejects many of
    them, and the clusters still equal the oracle's drop mode."""
    engine = pigeon_on
    recs = gsdr.Records.generate(law, seed, n)
    fa = str(tmp_path / "in.fasta")
    recs.write_fasta(fa)
    ocl, _, _ = O.cluster_file(fa, t, drop=True)
    recs.sort()
    k = gsdr.auto_word_length(t)
    store = engine.load(recs, k)
    try:
        for mode in (gsdr.MODE_PLAIN, gsdr.MODE_BATCHED):
            print(f"\n# pigeon drop law={law} n={n} t={t} mode={mode}", flush=True)
            cl, st = store.cluster(t, k, mode=mode, tune=True, drop_filters=True)
            assert np.array_equal(cl, ocl), f"{np.count_nonzero(cl != ocl)} assignments differ"
            assert st.aligned == st.accepted + st.rej_align + int(st.spec_wasted) + int(st.rej_pigeon)
            assert st.rej_pigeon > 0 and st.rej_pigeon <= st.pigeon_checked, (st.rej_pigeon, st.pigeon_checked)
    finally:
        store.free()
