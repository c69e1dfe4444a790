"""End-to-end parity: cluster[] and result.fasta of the HIP path against the oracle."""
import filecmp
import os

import numpy as np
import pytest

import gsdr
import oracle_ffi as O

pytestmark = pytest.mark.gpu


def run_both(engine, tmp_path, recs, t, mode=gsdr.MODE_PLAIN, k=0):
    fa = str(tmp_path / "in.fasta")
    recs.write_fasta(fa)
    ocl, operm, ost = O.cluster_file(fa, t, k=k, out_path=str(tmp_path / "orc.fasta"))
    r = gsdr.Records.read_fasta(fa)
    perm = r.sort()
    assert np.array_equal(perm, operm)
    cl, st = gsdr.cluster_records(engine, r, t, word_length=k, mode=mode)
    r.write_result(cl, str(tmp_path / "gpu.fasta"))
    return cl, ocl, st, ost, filecmp.cmp(str(tmp_path / "gpu.fasta"), str(tmp_path / "orc.fasta"), shallow=False)


@pytest.mark.parametrize("config,n,t", [(1, 3000, 0.95), (1, 10000, 0.95), (2, 4000, 0.90), (2, 3000, 0.80),
                                        (4, 4000, 0.95), (4, 4000, 0.80), (1, 2000, 0.99), (2, 2000, 0.85)])
def test_plain_matches_oracle(engine, tmp_path, config, n, t):
    recs = gsdr.Records.generate(config, 1000 + config, n)
    cl, ocl, st, ost, same_file = run_both(engine, tmp_path, recs, t)
    bad = np.nonzero(cl != ocl)[0]
    assert len(bad) == 0, f"{len(bad)} assignments differ, first at {bad[:5]}"
    assert same_file
    assert st.reps == ost["reps"]
    # stage accounting: every considered pair is rejected by exactly one stage or aligned
    assert st.pairs == st.rej_length + st.rej_comp + st.rej_word + st.aligned
    assert st.aligned == st.accepted + st.rej_align


@pytest.mark.parametrize("config,n,t", [(1, 3000, 0.95), (2, 2000, 0.90), (4, 4000, 0.80)])
def test_drop_mode_matches_oracle(engine, tmp_path, config, n, t):
    """Opt-in drop mode (no composition or shared-word filter) in both forms
    against the oracle's drop mode."""
    recs = gsdr.Records.generate(config, 1000 + config, n)
    fa = str(tmp_path / "in.fasta")
    recs.write_fasta(fa)
    ocl, operm, ost = O.cluster_file(fa, t, drop=True)
    r = gsdr.Records.read_fasta(fa, length_order=True)
    assert np.array_equal(r.sort(), operm)
    for mode in (gsdr.MODE_PLAIN, gsdr.MODE_BATCHED):
        cl, st = gsdr.cluster_records(engine, r, t, mode=mode, drop_filters=True)
        bad = np.nonzero(cl != ocl)[0]
        assert len(bad) == 0, f"{len(bad)} assignments differ, first at {bad[:5]}"
        assert st.rej_comp == 0 and st.rej_word == 0
        assert st.pairs == st.rej_length + st.aligned
        assert st.reps == ost["reps"]


@pytest.mark.parametrize("config,n,t,k", [(1, 2500, 0.95, 8), (2, 2000, 0.99, 7), (4, 3000, 0.97, 0),
                                          (1, 2000, 0.90, 7)])
def test_long_words_match_oracle(engine, tmp_path, config, n, t, k):
    """Word lengths 7 and 8 keep their shared-word tables in global memory,
    one per representative slot of a step; both forms against the oracle."""
    recs = gsdr.Records.generate(config, 2000 + config, n)
    for mode in (gsdr.MODE_PLAIN, gsdr.MODE_BATCHED):
        cl, ocl, st, ost, same_file = run_both(engine, tmp_path, recs, t, mode=mode, k=k)
        bad = np.nonzero(cl != ocl)[0]
        assert len(bad) == 0, f"{len(bad)} assignments differ, first at {bad[:5]}"
        assert same_file
        assert st.pairs == st.rej_length + st.rej_comp + st.rej_word + st.aligned
        assert st.reps == ost["reps"]


def test_iupac_crlf_edge_set(engine, tmp_path):
    rng = np.random.default_rng(5)
    lines = []
    base = "".join(rng.choice(list("ACGT"), 700))
    for i in range(300):
        s = list(base[: int(rng.integers(200, 700))])
        for _ in range(int(rng.integers(0, 12))):
            s[int(rng.integers(0, len(s)))] = rng.choice(list("NRYSWKMacgtu"))
        for _ in range(int(rng.integers(0, 10))):
            s[int(rng.integers(0, len(s)))] = rng.choice(list("ACGT"))
        seq = "".join(s)
        lines.append(f">r{i}\r\n")
        for p in range(0, len(seq), 70):
            lines.append(seq[p:p + 70] + ("\r\n" if i % 2 else "\n"))
        if i % 17 == 0:
            lines.append("\n")
    fa = tmp_path / "edge.fasta"
    fa.write_text("".join(lines))
    for t in (0.95, 0.85):
        ocl, operm, _ = O.cluster_file(str(fa), t, out_path=str(tmp_path / "o.fasta"))
        r = gsdr.Records.read_fasta(str(fa))
        perm = r.sort()
        assert np.array_equal(perm, operm)
        cl, _ = gsdr.cluster_records(engine, r, t)
        assert np.array_equal(cl, ocl)
        r.write_result(cl, str(tmp_path / "g.fasta"))
        assert filecmp.cmp(str(tmp_path / "g.fasta"), str(tmp_path / "o.fasta"), shallow=False)


def test_single_record_and_ties(engine, tmp_path):
    fa = tmp_path / "one.fasta"
    fa.write_text(">only\nACGTACGTAC\n")
    r = gsdr.Records.read_fasta(str(fa))
    r.sort()
    cl, st = gsdr.cluster_records(engine, r, 0.95)
    assert list(cl) == [0] and st.reps == 1
    # many equal lengths: the tie order must follow libstdc++ std::sort
    rng = np.random.default_rng(9)
    recs = [("".join(rng.choice(list("ACGT"), 400))) for _ in range(40)]
    fa2 = tmp_path / "ties.fasta"
    fa2.write_text("".join(f">t{i}\n{s}\n" for i, s in enumerate(recs * 5)))
    ocl, operm, _ = O.cluster_file(str(fa2), 0.95)
    r2 = gsdr.Records.read_fasta(str(fa2))
    assert np.array_equal(r2.sort(), operm)
    cl2, _ = gsdr.cluster_records(engine, r2, 0.95)
    assert np.array_equal(cl2, ocl)


@pytest.mark.parametrize("config,n,t", [(1, 10000, 0.95), (2, 6000, 0.90), (4, 6000, 0.95), (4, 6000, 0.80),
                                        (2, 3000, 0.99)])
def test_batched_matches_oracle(engine, tmp_path, config, n, t):
    recs = gsdr.Records.generate(config, 2000 + config, n)
    cl, ocl, st, ost, same_file = run_both(engine, tmp_path, recs, t, mode=gsdr.MODE_BATCHED)
    bad = np.nonzero(cl != ocl)[0]
    assert len(bad) == 0, f"{len(bad)} assignments differ, first at {bad[:5]}"
    assert same_file
    assert st.reps == ost["reps"]
    assert st.rounds < ost["reps"] or ost["reps"] < 8  # speculation fixes several representatives per step


@pytest.mark.parametrize("spec", [1, 2, 7, 64])
def test_speculation_width_invariant(engine, spec):
    recs = gsdr.Records.generate(2, 77, 5000)
    recs.sort()
    store = engine.load(recs, 5)
    try:
        base, _ = store.cluster(0.9, 5, mode=gsdr.MODE_PLAIN)
        got, st = store.cluster(0.9, 5, mode=gsdr.MODE_BATCHED, spec_reps=spec)
        assert np.array_equal(base, got)
        timed, _ = store.cluster(0.9, 5, mode=gsdr.MODE_BATCHED, spec_reps=spec, timing=True)
        assert np.array_equal(base, timed)
    finally:
        store.free()


def test_jobs_alone_equals_in_mix(engine):
    specs = [(2, 11, 3000, 0.90), (1, 12, 2500, 0.95), (4, 13, 4000, 0.80), (2, 14, 1500, 0.85), (4, 15, 3000, 0.99),
             (1, 16, 800, 0.88)]
    stores, alone, ts = [], [], []
    try:
        for cfg, seed, n, t in specs:
            r = gsdr.Records.generate(cfg, seed, n)
            r.sort()
            k = gsdr.auto_word_length(t)
            s = engine.load(r, k)
            stores.append(s)
            ts.append(t)
            cl, _ = s.cluster(t, k, mode=gsdr.MODE_PLAIN)
            alone.append(cl)
        mixed, st = engine.cluster_jobs(stores, ts, mode=gsdr.MODE_BATCHED)
        for a, b in zip(alone, mixed):
            assert np.array_equal(a, b)
        assert st.n_seqs == sum(s.n for s in stores)
    finally:
        for s in stores:
            s.free()


@pytest.mark.parametrize("mode", [gsdr.MODE_PLAIN, gsdr.MODE_BATCHED])
def test_reference_result_subset_fixed_point(engine, tmp_path, mode):
    """The reference's own output re-clustered at t=0.95 keeps every record
    (closest pairs sit exactly at LCS == ceil(L*t), rejected by the strict >)."""
    fa = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "result_subset.fasta")
    r = gsdr.Records.read_fasta(fa)
    r.sort()
    cl, st = gsdr.cluster_records(engine, r, 0.95, mode=mode)
    assert (cl == np.arange(len(r))).all()
    assert st.aligned > 0 and st.accepted == 0
