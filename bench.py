#!/usr/bin/env python3
"""Throughput benchmark of the de-redundancy path on one MI355X.

A step is one complete clustering (every round) of a synthetic mutated-family
workload whose packed store is already resident in HBM. Prints one JSON line:
clustered Mbp/s of the measured form (`value`), banded GCUPS, the alignment
kernels' VALU roofline on the cells they execute, the filter's HBM figure,
per-stage rejection rates, the plain form beside the batched one, the
`cluster` command timed end to end (file in, result.fasta out) in both forms,
the check of cluster[] and result.fasta against the oracle's digests at this
size, and the oracle timed on the host cores (a reported baseline).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c4|c2|c1]
                  [--mode batched|plain] [--no-cpu]
"""
import argparse
import ctypes as C
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "gene-sequences-de-redundancy_amd")
sys.path.insert(0, PKG)

# gfx950 (MI355X) peaks used for the roofline
PEAK_HBM_GBS = 8000.0
PEAK_VALU_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12  # int32 lane-ops/s: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz

from gsdr import workloads as W  # noqa: E402

# workloads: (generator law, seed, record count, threshold, description)
CONFIGS = {name: (law, seed, n, t, f"{desc}, t={t}") for name, (law, seed, n, t, desc) in W.SINGLE.items()}
FIXTURE = os.path.join(ROOT, "tests", "golden", "full_hashes.json")
CLI = os.path.join(PKG, "build", "cluster")


def sha_file(path):
    import hashlib

    h = hashlib.sha256()
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 24), b""):
            h.update(chunk)
    return h.hexdigest()


def sha_cluster(cl):
    import hashlib

    import numpy as np

    return hashlib.sha256(np.ascontiguousarray(cl, dtype="<i4").tobytes()).hexdigest()


def oracle_digests(cfg_name, n, t):
    """The oracle's digests for this exact workload (tests/golden/full_hashes.json), or None."""
    if not os.path.exists(FIXTURE):
        return None
    e = json.load(open(FIXTURE))["single"].get(f"{cfg_name}_{t:.2f}")
    if e is None or e["n"] != n or abs(e["t"] - t) > 1e-9:
        return None
    return e


def host_cpu():
    """CPU model and logical CPU count of this host (lscpu, nproc)."""
    import subprocess

    model = ""
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def cpu_baseline(cfg_name, threads):
    """Oracle (reference algorithm: full bit-vector LCS per pair, OpenMP over
    each round's candidates) on a bounded sample of the same workload law."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import gsdr
    import oracle_ffi as O

    gen, seed, n, t, _ = CONFIGS[cfg_name]
    sample_n = {"c1": 10000, "c2": 20000, "c4": 150000, "t99": 20000}[cfg_name]
    recs = gsdr.Records.generate(gen, seed, sample_n)
    with tempfile.TemporaryDirectory() as d:
        fa = os.path.join(d, "sample.fasta")
        recs.write_fasta(fa)
        t0 = time.perf_counter()
        _, _, st = O.cluster_file(fa, t, threads=threads, fast=False)
        dt = time.perf_counter() - t0
    out = {"value": round(recs.bases / 1e6 / dt, 3), "unit": "Mbp/s", "cores": threads, "kind": "port",
           "sample": f"a smaller instance of the same generator law: {sample_n} records ({recs.bases} bases, "
                     f"seed {seed}) clustered completely at t={t} by the oracle in the reference's mode (full "
                     f"bit-vector LCS per aligned pair, OpenMP over each round's candidates, {threads} threads), "
                     f"{dt:.2f} s; not the benchmarked {n}-record workload, whose full-LCS run takes hours",
           "reps": st["reps"]}
    out.update(host_cpu())
    return out


def committed_traffic(args, nrec, t):
    """HBM bytes per launch measured by rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    of this same default workload (scaled by a plain copy of known size), and
    the align kernels' VALU wave instructions per clustering (SQ_INSTS_VALU),
    read from the committed summary; empty for any other workload."""
    default = CONFIGS["c4"]
    if args.config != "c4" or nrec != default[2] or abs(t - default[3]) > 1e-9:
        return {}
    path = os.path.join(ROOT, "profiles", "align_filter_c4_t095_pmc.md")
    out = {}
    try:
        for line in open(path):
            for key, name in (("align", "traffic_per_align_launch_bytes"),
                              ("filter", "traffic_per_filter_launch_bytes"),
                              ("valu_insts_per_clustering", "align_valu_insts_per_clustering")):
                if line.startswith(name + " = "):
                    out[key] = int(float(line.split("=")[1]))
    except OSError:
        return {}
    return out


def parallel_map(fn, items, threads=16):
    """Runs fn over items on host threads (the generator releases the GIL)."""
    import concurrent.futures as cf

    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(fn, items))


def apply_knobs(eng, args, tuning=None):
    """Sets each --knob NAME=VALUE on the engine's tuned knobs; returns the
    dict of knobs set (None when none)."""
    import gsdr

    for kv in args.knob:
        # any scalar knob by name (the values are checked by gsdr_set_tuning)
        name, _, val = kv.partition("=")
        if name not in dict(gsdr.Tuning._fields_) or name in ("shape_mask", "shape_bias", "ms_default", "ms_tuned"):
            sys.exit(f"bench.py: --knob {kv}: not a scalar launch knob")
        tu = eng.tuning()
        setattr(tu, name, int(val))
        eng.set_tuning(tu)
        tuning = dict(tuning or {}, **{name: int(val)})
    return tuning


def bench_jobs(args):
    """Many independent jobs (log-uniform sizes, per-job median length and
    threshold) clustered together in shared batched launches."""
    import numpy as np

    import gsdr

    nj = args.jobs_count
    specs = W.many_jobs(nj, size_range=(500, int(50000 * args.jobs_scale)))

    def gen(s):
        return W.generate(s, sort=True)

    t0 = time.perf_counter()
    recs = parallel_map(gen, specs)
    print(f"# generated {nj} jobs, {sum(len(r) for r in recs)} records in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    eng = gsdr.Engine(0)
    t0 = time.perf_counter()
    stores = []
    for s, r in zip(specs, recs):
        st = eng.load(r, gsdr.auto_word_length(s["t"]))
        stores.append(st)
    load_s = time.perf_counter() - t0
    bases = sum(r.bases for r in recs)
    print(f"# loaded {bases / 1e6:.1f} Mbp in {load_s:.1f} s, HBM {sum(s.footprint for s in stores) / 2**30:.2f} GiB",
          file=sys.stderr, flush=True)
    ts = [s["t"] for s in specs]
    tuning = apply_knobs(eng, args)
    tune = tuning is not None
    eng.cluster_jobs(stores, ts, tune=tune)  # warmup
    times = []
    for i in range(args.steps):
        t0 = time.perf_counter()
        outs, st = eng.cluster_jobs(stores, ts, tune=tune)
        times.append(time.perf_counter() - t0)
        print(f"# step {i}: {times[-1] * 1e3:.1f} ms, steps {st.rounds}, reps {st.reps}", file=sys.stderr, flush=True)
    # each job alone must give the same clusters (spot check)
    same = all(np.array_equal(stores[j].cluster(ts[j], stores[j].k, mode=gsdr.MODE_BATCHED)[0], outs[j])
               for j in range(0, nj, max(1, nj // 8)))
    plain_s = None
    if not args.no_compare:
        t0 = time.perf_counter()
        for j in range(nj):
            stores[j].cluster(ts[j], stores[j].k, mode=gsdr.MODE_PLAIN)
        plain_s = time.perf_counter() - t0
    ms = 1e3 * sum(times) / len(times)
    out = {"metric": "clustered Mbp/s, many independent jobs in shared launches", "value": round(bases / 1e6 / (ms / 1e3), 3),
           "unit": "Mbp/s", "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": round(ms, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
           "data": "synthetic (seeded generator of the stated law)",
           "config": {"workload": f"many jobs: {nj} jobs, sizes log-uniform [500, {int(50000 * args.jobs_scale)}], "
                                  "medians {150,400,1500,5000}, t in {.80,.85,.90,.95,.99}",
                      "records": int(sum(len(r) for r in recs)), "bases": int(bases)},
           "plain_sequential_s": round(plain_s, 3) if plain_s else None,
           "batched_over_plain": round(plain_s / (ms / 1e3), 3) if plain_s else None,
           "alone_equals_mix_sample": bool(same), "reps": int(st.reps), "tuned": tuning}
    for s in stores:
        s.free()
    eng.close()
    print(json.dumps(out), flush=True)


def bench_serve(args):
    """Poisson arrivals of clustering jobs at several offered loads,
    admitted by HBM footprint into the continuous-batching server. Jobs are
    generated by loader threads ahead of their arrival."""
    import numpy as np

    import gsdr
    from gsdr import serve

    rates = [float(x) for x in args.rates.split(",")]
    eng = gsdr.Engine(0)
    results = []
    for ri, rate in enumerate(rates):
        count = max(4, int(round(rate * args.duration)))
        specs = W.stream_jobs(count, seed=1005 + ri, size_range=(1000, int(200000 * args.jobs_scale)))
        print(f"# rate {rate}: {count} jobs over {args.duration:.0f} s", file=sys.stderr, flush=True)
        out, summary = serve.run_stream(eng, specs, None, rate, hbm_budget=int(args.budget_gb * 2**30),
                                        max_slots=args.slots, seed=ri, loaders=args.loaders,
                                        make=lambda sp: W.generate(sp, sort=True), progress_s=30.0,
                                        drain_s=args.drain)
        # the first jobs again, each alone: the same clusters as in the mix
        check = []
        for r in out[:3]:
            rec = W.generate(specs[r["job"]], sort=True)
            cl, _ = gsdr.cluster_records(eng, rec, specs[r["job"]]["t"], mode=gsdr.MODE_BATCHED)
            check.append(bool(np.array_equal(cl, r["cluster"])))
            rec.free()
        summary["alone_equals_mix_first3"] = all(check)
        summary["offered_mbp_per_s"] = summary["offered_mbp"] / max(1e-9, count / rate)
        print(f"# rate {rate}: {json.dumps(summary)}", file=sys.stderr, flush=True)
        results.append(summary)
    eng.close()
    best = max(results, key=lambda s: s["sustained_mbp_per_s"])
    print(json.dumps({"metric": "serving: sustained Mbp/s and p50/p99 job latency under Poisson arrivals",
                      "value": round(best["sustained_mbp_per_s"], 3), "unit": "Mbp/s", "n_gpus": 1,
                      "higher_is_better": True, "data": "synthetic (seeded generator of the stated law)",
                      "config": {"workload": f"Poisson job stream: sizes log-uniform [1000, {int(200000 * args.jobs_scale)}], "
                                             f"t in {{.80,.90,.95,.99}}, {args.duration:.0f} s of arrivals per rate",
                                 "budget_gb": args.budget_gb, "slots": args.slots, "drain_s": args.drain},
                      "rates": results}), flush=True)


def copy_calibration():
    """One 1 GiB device-to-device hipMemcpy: the copy kernel's FETCH_SIZE and
    WRITE_SIZE per dispatch calibrate the counters against known bytes."""
    hip = C.CDLL("libamdhip64.so")
    nbytes = 1 << 30
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes": nbytes, "copies": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default=os.environ.get("GSDR_BENCH_CONFIG", "c4"), choices=sorted(CONFIGS))
    ap.add_argument("--n", type=int, default=0, help="override the record count")
    ap.add_argument("--mode", default="batched", choices=["plain", "batched"])
    ap.add_argument("--no-compare", action="store_true", help="skip timing the other form")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cli", action="store_true", help="skip timing the cluster command end to end")
    ap.add_argument("--dump", default="", help="write cluster[] (int32, sorted order) to this file")
    ap.add_argument("--threshold", type=float, default=0.0, help="override the workload's threshold")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy (reference kernel for HBM counters)")
    ap.add_argument("--tune", action="store_true", help="autotune the launch knobs on the workload first")
    ap.add_argument("--tune-trials", type=int, default=40)
    ap.add_argument("--pigeon", action="store_true", help="enable the exact pigeonhole prefilter")
    ap.add_argument("--no-screen", action="store_true", help="turn the word-signature screen off")
    ap.add_argument("--split-rows", type=int, default=-1,
                    help="pairs with this many rows or more on two waves (0 = never; default: built in)")
    ap.add_argument("--knob", action="append", default=[], metavar="NAME=VALUE",
                    help="set one scalar launch knob of gsdr_tuning (repeatable), e.g. align_blocks=8")
    ap.add_argument("--serial-shapes", action="store_true",
                    help="run the tile-shape kernels of a step phase one after another (profiling)")
    ap.add_argument("--jobs", action="store_true", help="many independent jobs in shared launches")
    ap.add_argument("--jobs-count", type=int, default=512)
    ap.add_argument("--jobs-scale", type=float, default=1.0, help="scale the upper job size")
    ap.add_argument("--serve", action="store_true", help="Poisson stream of jobs into the serving loop")
    ap.add_argument("--rates", default="0.5,1,2,4")
    ap.add_argument("--duration", type=float, default=20.0, help="seconds of arrivals per rate")
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--slots", type=int, default=64)
    ap.add_argument("--loaders", type=int, default=8, help="serving: host threads generating and submitting jobs")
    ap.add_argument("--drain", type=float, default=300.0, help="serving: seconds after the last arrival to finish")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.jobs:
        bench_jobs(args)
        return
    if args.serve:
        bench_serve(args)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (every box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import gsdr

    gen, seed, n, t, desc = CONFIGS[args.config]
    if args.n:
        n = args.n
    if args.threshold:
        t = args.threshold
        desc = desc.rsplit(", t=", 1)[0] + f", t={t}"
    want = oracle_digests(args.config, n, t)
    tmpdir = tempfile.TemporaryDirectory(prefix="gsdr_bench_")
    fasta = os.path.join(tmpdir.name, "input.fasta")
    recs = gsdr.Records.generate(gen, seed, n)
    cli_ok = not args.no_cli
    if cli_ok or want:
        # the input as a file, in generator order: what the drop-in command reads
        recs.write_fasta(fasta)
        if want:
            assert sha_file(fasta) == want["fasta"], "generated input differs from the oracle fixture's"
    recs.sort()
    k = gsdr.auto_word_length(t)
    eng = gsdr.Engine(0)
    store = eng.load(recs, k)
    mode = gsdr.MODE_BATCHED if args.mode == "batched" else gsdr.MODE_PLAIN
    print(f"# {desc}; n={len(recs)} bases={recs.bases} k={k} store={store.footprint / 2**20:.1f} MiB "
          f"device={eng.info()['name']}", file=sys.stderr, flush=True)
    tuning = None
    if args.tune:
        # untimed: the autotuner clusters the store under several launch
        # settings (all give the same clusters) and keeps the fastest
        tu = store.autotune(t, k, max_trials=args.tune_trials)
        tuning = tu.as_dict()
        print(f"# autotune: {tuning}", file=sys.stderr, flush=True)
    if args.pigeon:
        # the exact pigeonhole prefilter between the filter and the alignment
        tu = eng.tuning()
        tu.pigeon = 1
        eng.set_tuning(tu)
        tuning = dict(tuning or {}, pigeon=1)
    if args.no_screen:
        # the phase-B filter without the word-signature screen in front
        tu = eng.tuning()
        tu.screen = 0
        eng.set_tuning(tu)
        tuning = dict(tuning or {}, screen=0)
    if args.split_rows >= 0:
        # pairs with this many rows on two waves (forward and backward halves)
        tu = eng.tuning()
        tu.split_rows = args.split_rows
        eng.set_tuning(tu)
        tuning = dict(tuning or {}, split_rows=args.split_rows)
    if args.serial_shapes:
        # the tile-shape kernels of a step phase one after another on one
        # stream (profiling: kernel durations then do not overlap)
        tu = eng.tuning()
        tu.fork_shapes = 0
        eng.set_tuning(tu)
        tuning = dict(tuning or {}, fork_shapes=0)
    tuning = apply_knobs(eng, args, tuning)
    tune = tuning is not None
    ref = None
    for i in range(args.warmup):
        cl, st = store.cluster(t, k, mode=mode, tune=tune)
        ref = cl if ref is None else ref
        print(f"# warmup {i}: {st.ms_total:.1f} ms, reps {st.reps}, rounds {st.rounds}", file=sys.stderr, flush=True)
    times, stats = [], []
    for i in range(args.steps):
        t0 = time.perf_counter()
        cl, st = store.cluster(t, k, mode=mode, tune=tune)
        times.append(time.perf_counter() - t0)
        stats.append(st)
        if ref is not None:
            assert (cl == ref).all(), "result changed between steps"
        ref = cl
        print(f"# step {i}: {times[-1] * 1e3:.1f} ms, reps {st.reps}, rounds {st.rounds}, aligned {st.aligned}",
              file=sys.stderr, flush=True)
    ms = 1e3 * sum(times) / len(times)
    # one more step with HIP events around every filter and alignment launch
    # (on the engine's stream) for the kernel-level figures
    clt, st = store.cluster(t, k, mode=mode, timing=True, tune=tune)
    assert (clt == ref).all(), "timed step differs"
    if args.dump:
        ref.astype("<i4").tofile(args.dump)
    print(f"# timing step: {st.ms_total:.1f} ms, align {st.ms_align:.1f} ms, filter {st.ms_filter:.1f} ms",
          file=sys.stderr, flush=True)
    parity = None
    if want:
        # the oracle's answer for this exact workload: cluster[] and result.fasta
        res = os.path.join(tmpdir.name, "device_result.fasta")
        recs.write_result(ref, res)
        parity = {"fixture": f"{args.config}_{t:.2f}", "cluster_sha256_equal": sha_cluster(ref) == want["cluster"],
                  "result_sha256_equal": sha_file(res) == want["result"], "reps": int(st.reps),
                  "oracle_reps": want["reps"]}
        assert parity["cluster_sha256_equal"] and parity["result_sha256_equal"], f"differs from the oracle: {parity}"
    bases = recs.bases
    align_s = st.ms_align / 1e3
    filt_s = st.ms_filter / 1e3
    out = {
        # the headline metric's own name on its workload; other workloads and
        # thresholds name their threshold
        "metric": f"clustered Mbp/s + banded GCUPS at t={t:g}, batched vs plain, bit-exact, 1 GPU",
        "value": round(bases / 1e6 / (ms / 1e3), 3),
        "unit": "Mbp/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded generator of the stated law)",
        "config": {"workload": desc, "records": len(recs), "bases": bases, "threshold": t, "word_length": k,
                   "form": args.mode, "tuned": tuning,
                   "value_is": "device pipeline: one whole clustering with the packed store resident in HBM"},
        "oracle_parity": parity,
        "reps": int(st.reps),
        "rounds": int(st.rounds),
        "stage": {"pairs": st.pairs, "rej_length": st.rej_length, "rej_comp": st.rej_comp,
                  "rej_word": st.rej_word, "aligned": st.aligned, "accepted": st.accepted,
                  "rej_align": st.rej_align, "align_skipped": int(st.spec_wasted),
                  "pigeon_checked": int(st.pigeon_checked), "rej_pigeon": int(st.rej_pigeon),
                  "screened": int(st.screened)},
        "cells": {"band": st.cells_band, "decided": st.cells_decided, "exec": st.cells_exec,
                  "full": st.cells_full},
        "kernel_ms": {"align": round(st.ms_align, 3), "filter": round(st.ms_filter, 3),
                      "align_launches": st.align_launches, "filter_launches": st.filter_launches},
    }
    # one align launch = the tile-shape kernels of one step phase (launched
    # together, one filter launch each); HIP events bracket each such launch
    groups = max(1, st.filter_launches)
    traffic = committed_traffic(args, len(recs), t) if args.mode == "batched" else {}
    if align_s > 0:
        # work executed: 4 int32 ops per 32-cell word step (AND, ADD, carry
        # ADD, AND-OR) over the bit-cells the kernels computed (cells_exec / 8)
        # per align time. The exact decision band of every aligned pair
        # (cells_band) is larger, since exact early decisions and the elastic
        # window skip part of it; its rate is the separate `effective` figure
        # (work the reference's band would need, per second) and can pass the
        # peak.
        ach_exec = st.cells_exec / 8.0 / align_s / 1e12
        eff = st.cells_band / 8.0 / align_s / 1e12
        out["banded_gcups"] = round(st.cells_exec / align_s / 1e9, 3)
        out["banded_gcups_effective"] = round(st.cells_band / align_s / 1e9, 3)
        out["roofline"] = {"bound": "valu", "kernel": "align_kernel (all tile shapes of a step phase)",
                           "achieved": round(ach_exec, 4), "peak": round(PEAK_VALU_TOPS, 2), "unit": "TOP/s",
                           "frac": round(ach_exec / PEAK_VALU_TOPS, 5), "traffic": traffic.get("align"),
                           "work": "cells_exec / 8 int32 lane-ops",
                           "per_launch_ms": round(st.ms_align / groups, 5),
                           "effective_band": {"achieved": round(eff, 4), "frac": round(eff / PEAK_VALU_TOPS, 5),
                                              "work": "cells_band / 8"},
                           "achieved_decided_cells": round(st.cells_decided / 8.0 / align_s / 1e12, 4)}
        if traffic.get("valu_insts_per_clustering"):
            # instruction-level cross-check from the committed rocprof run of
            # this workload: VALU wave instructions x 64 lanes per align time
            vi = traffic["valu_insts_per_clustering"] * 64 / align_s / 1e12
            out["roofline"]["valu_issued"] = {"achieved": round(vi, 4), "frac": round(vi / PEAK_VALU_TOPS, 5),
                                              "source": "profiles/align_filter_c4_t095_pmc.md SQ_INSTS_VALU"}
    if filt_s > 0:
        gbs = st.filter_bytes / filt_s / 1e9
        out["filter_roofline"] = {"bound": "hbm", "achieved": round(gbs, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                  "frac": round(gbs / PEAK_HBM_GBS, 5), "traffic": traffic.get("filter"),
                                  "per_launch_ms": round(st.ms_filter / groups, 5)}
    if not args.no_compare:
        other = gsdr.MODE_PLAIN if mode == gsdr.MODE_BATCHED else gsdr.MODE_BATCHED
        t0 = time.perf_counter()
        cl2, st2 = store.cluster(t, k, mode=other)
        dt = time.perf_counter() - t0
        assert (cl2 == ref).all(), "plain and batched forms disagree"
        out["other_form"] = {"form": "plain" if other == gsdr.MODE_PLAIN else "batched",
                             "ms": round(dt * 1e3, 3), "mbp_per_s": round(bases / 1e6 / dt, 3),
                             "rounds": int(st2.rounds)}
        plain_ms = dt * 1e3 if other == gsdr.MODE_PLAIN else ms
        batched_ms = ms if other == gsdr.MODE_PLAIN else dt * 1e3
        out["batched_over_plain"] = round(plain_ms / batched_ms, 3)
    store.free()
    eng.close()
    if cli_ok:
        out["end_to_end"] = cli_end_to_end(fasta, t, bases, tmpdir.name, want, args.no_compare)
    recs.free()
    tmpdir.cleanup()
    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0))
        out["cpu_baseline"] = cpu_baseline(args.config, threads)
    print(json.dumps(out), flush=True)


def cli_end_to_end(fasta, t, bases, tmp, want, batched_only):
    """The drop-in command timed from process start to exit (input file in
    the page cache, result.fasta written and closed), batched and plain."""
    import subprocess

    res = {}
    forms = ["batched"] if batched_only else ["batched", "plain"]
    for form in forms:
        out_fa = os.path.join(tmp, f"cli_{form}.fasta")
        report = os.path.join(tmp, f"cli_{form}.json")
        cmd = [CLI, "t", f"{t}", "i", fasta, "o", out_fa, "--mode", form, "--report", report]
        runs = []
        for rep in range(2 if form == "batched" else 1):  # the first batched run warms the page cache
            t0 = time.perf_counter()
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
            runs.append(time.perf_counter() - t0)
            assert r.returncode == 0, r.stderr
        wall = runs[-1]
        inner = json.load(open(report))
        entry = {"wall_s": round(wall, 4), "mbp_per_s": round(bases / 1e6 / wall, 3),
                 "inside_main_s": round(inner["wall_s"], 4),
                 "stages_s": {kk: round(inner[kk], 4) for kk in ("read_s", "sort_s", "open_s", "store_s",
                                                                 "cluster_call_s", "write_s")}}
        if want:
            entry["result_sha256_equal"] = sha_file(out_fa) == want["result"]
            assert entry["result_sha256_equal"], f"{form} command's result.fasta differs from the oracle's"
        res[form] = entry
        print(f"# cluster command ({form}): {wall:.3f} s wall, {entry['mbp_per_s']} Mbp/s", file=sys.stderr,
              flush=True)
    if "plain" in res:
        res["batched_over_plain"] = round(res["plain"]["wall_s"] / res["batched"]["wall_s"], 3)
    res["command"] = "cluster t T i input.fasta o result.fasta --mode FORM (wall from process start to exit)"
    return res


if __name__ == "__main__":
    main()
