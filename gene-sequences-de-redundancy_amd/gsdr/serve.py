"""Continuous-batching server on top of gsdr_serve_* (include/gsdr.h).

Server wraps one device server; run_stream replays a seeded Poisson arrival
stream of clustering jobs: a loader thread submits each job at its arrival
time (host preprocessing and upload overlap the device steps), jobs that do
not fit the HBM budget or a slot wait and retry, and the main thread keeps
stepping every admitted job in shared launches until all results are in.
"""
import ctypes as C
import threading
import time

import numpy as np

from . import (E_BUSY, OK, Params, Records, Stats, _check, _i32p, auto_word_length, lib)


class Server:
    def __init__(self, engine, hbm_budget, max_slots):
        self.h = C.c_void_p()
        _check(lib().gsdr_serve_open(engine.h, int(hbm_budget), int(max_slots), C.byref(self.h)), "serve_open")

    def submit(self, sorted_records, threshold, word_length=0, spec_reps=0):
        """Returns a job id, or None when the job must wait (no slot or budget)."""
        p = Params()
        p.threshold = threshold
        p.word_length = word_length or auto_word_length(threshold)
        p.spec_reps = spec_reps
        jid = C.c_int64()
        st = lib().gsdr_serve_submit(self.h, sorted_records.h, C.byref(p), C.byref(jid))
        if st == E_BUSY:
            return None
        _check(st, "serve_submit")
        return jid.value

    def step(self):
        running = C.c_int32()
        _check(lib().gsdr_serve_step(self.h, C.byref(running)), "serve_step")
        return running.value

    def poll(self, job_id, n):
        out = np.zeros(n, dtype=np.int32)
        st = Stats()
        done = C.c_int32()
        _check(lib().gsdr_serve_poll(self.h, job_id, _i32p(out), C.byref(st), C.byref(done)), "serve_poll")
        return (out, st) if done.value else None

    def close(self):
        if self.h:
            lib().gsdr_serve_close(self.h)
            self.h = C.c_void_p()


def make_jobs(seed, count, size_range=(1000, 200000), thresholds=(0.80, 0.90, 0.95, 0.99), median=1200.0):
    """Job specs of the serving workload: log-uniform sizes, thresholds drawn
    from the set, sequences generated like the 200,000-record mutated families."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(size_range[0]), np.log(size_range[1])
    specs = []
    for j in range(count):
        n = int(round(np.exp(rng.uniform(lo, hi))))
        t = float(rng.choice(thresholds))
        specs.append({"id": j, "n": n, "t": t, "seed": int(seed * 1000003 + j)})
    return specs


def generate_job(spec, median=1200.0):
    r = Records.generate(2, spec["seed"], spec["n"], median)
    r.sort()
    return r


def run_stream(engine, specs, records, rate, hbm_budget, max_slots, seed=0, step_sleep=0.0, loaders=1, make=None,
               progress_s=0.0, drain_s=None):
    """Replays a Poisson stream (rate jobs/s) of jobs.

    `records` holds the jobs' sorted Records, or is None: then the `loaders`
    host threads call make(spec) for each job ahead of its arrival and free
    the records once the job is admitted (the server keeps its own copy), so
    long streams need not be held in host memory. The loaders take the jobs
    in arrival order and submit each at its arrival time; jobs that do not
    fit the HBM budget or a slot wait and retry. progress_s > 0 prints a
    progress line to stderr that often; drain_s bounds the time after the
    last arrival (jobs not finished by then are reported as unfinished).
    Returns per-job dicts with latency (arrival to result) and clusters, and
    a summary."""
    import sys

    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, len(specs))
    arrivals = np.cumsum(gaps) - gaps[0]
    srv = Server(engine, hbm_budget, max_slots)
    lock = threading.Lock()
    submitted = {}   # job id -> (spec index, arrival time)
    sizes = {}       # spec index -> (records, bases)
    results = {}
    error = []
    nxt = [0]
    max_running = 0
    stop = [False]
    t0 = time.perf_counter()

    def loader():
        try:
            while not stop[0]:
                with lock:
                    i = nxt[0]
                    nxt[0] += 1
                if i >= len(specs):
                    return
                rec = records[i] if records is not None else make(specs[i])
                with lock:
                    sizes[i] = (len(rec), rec.bases)
                wait = t0 + arrivals[i] - time.perf_counter()
                if wait > 0:
                    time.sleep(wait)
                while not stop[0]:
                    jid = srv.submit(rec, specs[i]["t"])
                    if jid is not None:
                        with lock:
                            submitted[jid] = (i, t0 + arrivals[i])
                        break
                    time.sleep(0.005)
                if records is None:
                    rec.free()
        except Exception as e:  # surfaced by the stepping thread
            error.append(e)

    threads = [threading.Thread(target=loader, daemon=True) for _ in range(max(1, loaders))]
    for th in threads:
        th.start()
    last_report = t0
    deadline = None if drain_s is None else t0 + arrivals[-1] + drain_s
    while len(results) < len(specs):
        if error:
            raise error[0]
        now = time.perf_counter()
        if deadline is not None and now > deadline:
            break
        running = srv.step()
        max_running = max(max_running, running)
        with lock:
            pending = list(submitted.items())
        for jid, (i, t_arr) in pending:
            got = srv.poll(jid, sizes[i][0])
            if got is not None:
                t_done = time.perf_counter()
                results[i] = {"job": i, "n": specs[i]["n"], "t": specs[i]["t"], "bases": sizes[i][1],
                              "latency_s": t_done - t_arr, "cluster": got[0], "reps": int(got[1].reps),
                              "steps": int(got[1].rounds)}
                with lock:
                    del submitted[jid]
        if progress_s and now - last_report >= progress_s:
            last_report = now
            print(f"# serve rate {rate}: {now - t0:.0f} s, {nxt[0]} arrived or loading, {len(results)} done, "
                  f"{running} running", file=sys.stderr, flush=True)
        if running == 0 and step_sleep:
            time.sleep(step_sleep)
        elif running == 0:
            time.sleep(0.001)
    stop[0] = True
    for th in threads:
        th.join()
    wall = time.perf_counter() - t0
    srv.close()
    out = [results[i] for i in sorted(results)]
    lat = np.array([r["latency_s"] for r in out]) if out else np.zeros(1)
    summary = {"jobs": len(out), "jobs_offered": len(specs), "unfinished": len(specs) - len(out),
               "rate_jobs_per_s": rate, "wall_s": wall,
               "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
               "sustained_mbp_per_s": sum(r["bases"] for r in out) / 1e6 / wall, "max_concurrent": max_running,
               "offered_mbp": sum(b for _, b in sizes.values()) / 1e6}
    return out, summary
