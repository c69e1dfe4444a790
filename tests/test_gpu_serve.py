"""Serving: jobs admitted while others run share launches; each result must be
the same as the job clustered alone, and (staggered admission) the oracle's."""
import numpy as np
import pytest

import gsdr
import oracle_ffi as O
from gsdr import serve

pytestmark = pytest.mark.gpu

SPECS = [(2, 31, 3000, 0.90), (1, 32, 2500, 0.95), (4, 33, 4000, 0.80), (2, 34, 1500, 0.85), (4, 35, 3000, 0.99),
         (2, 36, 2000, 0.80)]


def make(cfg, seed, n):
    r = gsdr.Records.generate(cfg, seed, n)
    r.sort()
    return r


def alone(engine, recs, t):
    cl, _ = gsdr.cluster_records(engine, recs, t, mode=gsdr.MODE_PLAIN)
    return cl


def test_staggered_admission_matches_alone(engine, tmp_path):
    recs, ref = [], []
    for c, s, n, t in SPECS:
        r = gsdr.Records.generate(c, s, n)
        fa = str(tmp_path / f"job_{s}.fasta")
        r.write_fasta(fa)
        r.sort()
        recs.append(r)
        # the solo run is checked against the oracle on the same file
        cl = alone(engine, r, t)
        ocl, _, _ = O.cluster_file(fa, t)
        assert np.array_equal(cl, ocl), f"job seed {s} t={t}: solo run differs from the oracle"
        ref.append(cl)
    srv = serve.Server(engine, hbm_budget=8 << 30, max_slots=8)
    try:
        ids = {}
        order = [[0, 1], [2, 3], [4, 5]]
        for wave in order:
            for i in wave:
                jid = srv.submit(recs[i], SPECS[i][3])
                assert jid is not None
                ids[i] = jid
            srv.step()  # later jobs join while earlier ones are mid-run
        got = {}
        for _ in range(100000):
            running = srv.step()
            for i, jid in ids.items():
                if i not in got:
                    r = srv.poll(jid, len(recs[i]))
                    if r is not None:
                        got[i] = r[0]
            if running == 0 and len(got) == len(ids):
                break
        assert len(got) == len(ids)
        for i in ids:
            assert np.array_equal(got[i], ref[i]), f"job {i} differs from its solo run"
    finally:
        srv.close()


def test_budget_admission_and_errors(engine):
    a = make(2, 41, 3000)
    b = make(2, 42, 3000)
    # a budget that holds one job of this size only
    srv = serve.Server(engine, hbm_budget=int(a.bases * 4.25 + 3000 * 120 + (1 << 16)) + 1000, max_slots=4)
    try:
        ja = srv.submit(a, 0.9)
        assert ja is not None
        assert srv.submit(b, 0.9) is None  # must wait: budget is taken
        got = None
        while got is None:
            srv.step()
            got = srv.poll(ja, len(a))
        jb = srv.submit(b, 0.9)  # memory is free again
        assert jb is not None
        done = None
        while done is None:
            srv.step()
            done = srv.poll(jb, len(b))
        assert np.array_equal(done[0], alone(engine, b, 0.9))
        big = make(2, 43, 20000)
        with pytest.raises(gsdr.GsdrError) as e:
            srv.submit(big, 0.9)
        assert e.value.status == gsdr.E_LIMIT
        with pytest.raises(gsdr.GsdrError) as e:
            srv.poll(123456, 10)
        assert e.value.status == gsdr.E_NOTFOUND
    finally:
        srv.close()


def test_slot_limit(engine):
    srv = serve.Server(engine, hbm_budget=8 << 30, max_slots=2)
    try:
        rs = [make(1, 50 + i, 800) for i in range(3)]
        ids = [srv.submit(r, 0.95) for r in rs]
        assert ids[0] is not None and ids[1] is not None and ids[2] is None
    finally:
        srv.close()


def test_poisson_stream(engine):
    specs = serve.make_jobs(7, 8, size_range=(500, 4000))
    recs = [serve.generate_job(s) for s in specs]
    out, summary = serve.run_stream(engine, specs, recs, rate=40.0, hbm_budget=4 << 30, max_slots=6, seed=3)
    for r, s, rec in zip(out, specs, recs):
        assert np.array_equal(r["cluster"], alone(engine, rec, s["t"]))
    assert summary["p99_latency_s"] >= summary["p50_latency_s"] > 0
    assert summary["sustained_mbp_per_s"] > 0
