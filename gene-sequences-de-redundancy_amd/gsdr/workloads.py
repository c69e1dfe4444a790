"""Seeded synthetic workloads of the de-redundancy path.

Each single-job workload is (generator law, seed, record count, thresholds);
the many-job and stream workloads are lists of per-job specs drawn from a
seeded numpy generator. The records themselves come from the library's C++
generator (gsdr_records_generate), so a spec names its input exactly on any
machine with the same build.

  law 1  mutated families: 500 parents of 300..1500 bases, 19 children each
  law 2  log-normal lengths (median 1,200 or the job's median, sigma 0.5,
         clipped to 150..6,000), 10 children per parent with divergence
         uniform in [0, 0.12], 80 % substitutions / 20 % indels
  law 4  Pareto(1.5) lengths from 100 bases clipped at 30,000; 30 %
         singletons, 70 % in geometric families (mean 8) with divergence
         uniform in [0, 0.06]; 5 % carry a tandem repeat
"""
import numpy as np

# name -> (law, seed, records, default threshold, description)
SINGLE = {
    "c1": (1, 1001, 10000, 0.95, "mutated families: 10,000 records from 500 parents"),
    "c2": (2, 1002, 200000, 0.90, "mutated families: 200,000 records, log-normal median 1,200"),
    "c4": (4, 1004, 2000000, 0.95, "skewed single job: 2,000,000 records, Pareto(1.5) lengths"),
    # a high-threshold job of the many-job and stream laws: nearly every record
    # is its own representative (n/64 speculative steps at most)
    "t99": (2, 1002, 177000, 0.99, "high-threshold job: 177,000 records, log-normal median 1,200"),
}

MANY_MEDIANS = (150.0, 400.0, 1500.0, 5000.0)
MANY_THRESHOLDS = (0.80, 0.85, 0.90, 0.95, 0.99)
STREAM_THRESHOLDS = (0.80, 0.90, 0.95, 0.99)


def many_jobs(count, seed=1003, size_range=(500, 50000)):
    """Independent jobs of the many-job workload: sizes log-uniform over
    size_range, per-job median length and threshold drawn from the sets,
    each job generated with law 2 at its median."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(size_range[0]), np.log(size_range[1])
    specs = []
    for j in range(count):
        n = int(round(np.exp(rng.uniform(lo, hi))))
        median = float(rng.choice(MANY_MEDIANS))
        t = float(rng.choice(MANY_THRESHOLDS))
        specs.append({"id": j, "law": 2, "n": n, "median": median, "t": t, "seed": int(seed * 1000003 + j)})
    return specs


def stream_jobs(count, seed=1005, size_range=(1000, 200000)):
    """Jobs of the serving stream: sizes log-uniform over size_range,
    thresholds drawn from the set, law 2 at median 1,200."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(size_range[0]), np.log(size_range[1])
    specs = []
    for j in range(count):
        n = int(round(np.exp(rng.uniform(lo, hi))))
        t = float(rng.choice(STREAM_THRESHOLDS))
        specs.append({"id": j, "law": 2, "n": n, "median": 1200.0, "t": t, "seed": int(seed * 1000003 + j)})
    return specs


def generate(spec, sort=False):
    """Records of one job spec (dict with law, seed, n and optional median)."""
    from . import Records

    r = Records.generate(spec["law"], spec["seed"], spec["n"], spec.get("median", 0.0))
    if sort:
        r.sort()
    return r
