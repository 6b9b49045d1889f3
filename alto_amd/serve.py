"""Serving: a Python front end to the library's continuous-batching CP-ALS
server (alto_server_*), the seeded job-stream workload (Poisson arrivals,
3-5 modes, log-uniform dims and nonzero counts, Zipf(1.1) indices,
lognormal values, rank from {8, 16, 32, 64}), and a stream runner that
reports jobs/s and p50 / p99 arrival-to-result latency."""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np

from . import PREC_F64, PTR_DEVICE, PTR_HOST, AltoTensor, _ptr_array, check, lib


class JobDesc(C.Structure):
    _fields_ = [("nmodes", C.c_int), ("dims", C.c_void_p), ("nnz", C.c_uint64), ("cidx", C.c_void_p),
                ("vals", C.c_void_p), ("ptr_kind", C.c_int), ("rank", C.c_uint64), ("max_iters", C.c_int),
                ("epsilon", C.c_double), ("seed", C.c_uint), ("precision", C.c_int), ("want_factors", C.c_int)]


class JobResult(C.Structure):
    _fields_ = [("status", C.c_int), ("iters", C.c_int), ("fit", C.c_double), ("t_submit", C.c_double),
                ("t_start", C.c_double), ("t_end", C.c_double), ("footprint", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [("queued", C.c_int), ("running", C.c_int), ("hbm_in_use", C.c_uint64), ("hbm_peak", C.c_uint64),
                ("hbm_budget", C.c_uint64)]


def _bind():
    L = lib()
    vp, u64, i32 = C.c_void_p, C.c_uint64, C.c_int
    L.alto_server_create.argtypes = [u64, i32, C.POINTER(vp)]
    L.alto_server_create.restype = i32
    L.alto_server_submit.argtypes = [vp, C.POINTER(JobDesc), C.POINTER(u64)]
    L.alto_server_submit.restype = i32
    L.alto_server_result.argtypes = [vp, u64, i32, C.POINTER(JobResult), vp, vp]
    L.alto_server_result.restype = i32
    L.alto_server_release.argtypes = [vp, u64]
    L.alto_server_release.restype = i32
    L.alto_server_stats.argtypes = [vp, C.POINTER(ServerStats)]
    L.alto_server_stats.restype = i32
    L.alto_server_destroy.argtypes = [vp]
    L.alto_server_destroy.restype = i32
    return L


class Server:
    """Continuous-batching server: admission by HBM footprint, one stream per
    slot; with lockstep_nnz > 0, jobs of at most that many nonzeros (rank <=
    64) are iterated in lockstep groups, at most lockstep_sessions at once."""

    def __init__(self, hbm_budget_bytes=200 << 30, slots=8, lockstep_nnz=0, lockstep_sessions=64):
        self.L = _bind()
        h = C.c_void_p()
        # the server reads these settings when it is created
        check(self.L.alto_tune_set(22, int(lockstep_nnz)), "tune_set")
        check(self.L.alto_tune_set(23, int(lockstep_sessions)), "tune_set")
        try:
            check(self.L.alto_server_create(int(hbm_budget_bytes), int(slots), C.byref(h)), "server_create")
        finally:
            self.L.alto_tune_set(22, 0)
            self.L.alto_tune_set(23, 64)
        self.h = h.value
        self._keep = {}

    def submit(self, dims, cidx, vals, rank, max_iters=20, epsilon=1e-5, seed=1, precision=PREC_F64, device=False,
               nnz=None, want_factors=False):
        """cidx: list of per-mode coordinate arrays (numpy uint64) or device pointers
        (device=True, with nnz given); vals likewise."""
        dims_a = np.ascontiguousarray(dims, dtype=np.uint64)
        if device:
            rows = _ptr_array(list(cidx))
            vptr = vals
            n = int(nnz)
            keep = (dims_a, rows)
        else:
            arrs = [np.ascontiguousarray(c, dtype=np.uint64) for c in cidx]
            v = np.ascontiguousarray(vals, dtype=np.float64)
            rows = _ptr_array([a.ctypes.data for a in arrs])
            vptr = v.ctypes.data
            n = v.shape[0]
            keep = (dims_a, rows, arrs, v)
        d = JobDesc(len(dims_a), dims_a.ctypes.data, n, C.cast(rows, C.c_void_p), vptr,
                    PTR_DEVICE if device else PTR_HOST, rank, max_iters, epsilon, seed, precision, int(want_factors))
        jid = C.c_uint64()
        check(self.L.alto_server_submit(self.h, C.byref(d), C.byref(jid)), "server_submit")
        self._keep[jid.value] = keep
        return jid.value

    def result(self, job, wait=True, dims=None, rank=None):
        r = JobResult()
        U = lam = None
        if dims is not None:
            U = [np.zeros((int(d), rank), np.float64) for d in dims]
            lam = np.zeros(rank, np.float64)
            up = _ptr_array([u.ctypes.data for u in U])
            check(self.L.alto_server_result(self.h, job, int(wait), C.byref(r), C.cast(up, C.c_void_p),
                                            lam.ctypes.data), "server_result")
        else:
            check(self.L.alto_server_result(self.h, job, int(wait), C.byref(r), None, None), "server_result")
        return r, U, lam

    def release(self, job):
        self.L.alto_server_release(self.h, job)
        self._keep.pop(job, None)

    def stats(self):
        s = ServerStats()
        check(self.L.alto_server_stats(self.h, C.byref(s)), "server_stats")
        return s

    def close(self):
        if self.h:
            check(self.L.alto_server_destroy(self.h), "server_destroy")
            self.h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def job_stream(n_jobs, seed=5, rate=50.0, nnz_range=(1e4, 2e7), dim_range=(64, 2 ** 20), ranks=(8, 16, 32, 64)):
    """Seeded job parameters: arrival time, dims, nnz, rank, generator seed."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for k in range(n_jobs):
        t += rng.exponential(1.0 / rate)
        nm = int(rng.integers(3, 6))
        dims = [int(round(math.exp(rng.uniform(math.log(dim_range[0]), math.log(dim_range[1]))))) for _ in range(nm)]
        cells = float(np.prod(np.array(dims, dtype=np.float64)))
        nnz = int(round(math.exp(rng.uniform(math.log(nnz_range[0]), math.log(nnz_range[1])))))
        nnz = max(1, min(nnz, int(0.3 * cells)))
        jobs.append(dict(arrival=t, dims=dims, nnz=nnz, rank=int(rng.choice(ranks)), seed=5_000_000 + k))
    return jobs


def _generate(j, device="cuda"):
    """One job's tensor on the GPU: nnz Zipf(1.1) draws per mode over a
    seeded permutation of the mode's range (duplicate coordinates kept; in
    COO they add up) and lognormal values: (coords, vals, nnz)."""
    from . import gen
    coords, vals = gen.sparse_draws(j["dims"], j["nnz"], j["seed"], s=1.1, values="lognormal", device=device)
    return coords, vals, int(vals.numel())


def materialize(jobs, device="cuda"):
    """Generates every job's tensor on the GPU at once (see _generate)."""
    for j in jobs:
        j["coords"], j["vals"], j["nnz"] = _generate(j, device)
    return jobs


def run_stream(jobs, precision, slots, budget=200 << 30, max_iters=20, epsilon=1e-5, time_scale=1.0, window=48,
               lockstep_nnz=0):
    """Submits the jobs at their (scaled) arrival times and returns throughput
    and latency.  A job without a materialised tensor ("coords") has it
    generated on the GPU just before its submission and freed as soon as the
    job finished, and at most `window` generated jobs hold their inputs at
    once (submission waits for any of them to finish), so a stream of any
    length runs in bounded memory.  Latency is submission to result; the
    generation, before submission, is not part of it.  lockstep_nnz: see
    Server."""
    import collections

    import torch

    srv = Server(budget, slots, lockstep_nnz=lockstep_nnz)
    ids, nnzs = [], []
    live = collections.deque()  # (job id, generated inputs) in submission order
    done = {}

    def retire(block):
        # frees the inputs of every finished job; with block, waits until at
        # least one has finished (whichever it is: a long job submitted
        # early does not hold the window)
        while live:
            keep = collections.deque()
            for jid, inputs in live:
                r, _, _ = srv.result(jid, wait=False)
                if r.status == 0 or r.status == 1:
                    keep.append((jid, inputs))
                else:
                    done[jid] = r
            freed = len(keep) < len(live)
            live.clear()
            live.extend(keep)
            if freed or not block:
                return
            time.sleep(0.0002)

    t0 = time.perf_counter()
    for j in jobs:
        if "coords" in j:
            coords, vals, nnz = j["coords"], j["vals"], j["nnz"]
            gen_keep = None
        else:
            while len(live) >= window:
                retire(True)
            coords, vals, nnz = _generate(j)
            torch.cuda.current_stream().synchronize()  # the server reads them on its own streams
            gen_keep = (coords, vals)
        wait = j["arrival"] * time_scale - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        jid = srv.submit(j["dims"], [c.data_ptr() for c in coords], vals.data_ptr(), j["rank"], max_iters, epsilon,
                         j["seed"] % (2 ** 31), precision, device=True, nnz=nnz)
        ids.append(jid)
        nnzs.append(nnz)
        if gen_keep is not None:
            live.append((jid, gen_keep))
            retire(False)
    while live:
        retire(True)
    res = [done[i] if i in done else srv.result(i)[0] for i in ids]
    peak = srv.stats().hbm_peak
    srv.close()
    bad = [r.status for r in res if r.status != 2]
    if bad:
        raise RuntimeError(f"{len(bad)} jobs failed (status {bad[:5]})")
    lat = np.array([r.t_end - r.t_submit for r in res])
    first = min(r.t_submit for r in res)
    last = max(r.t_end for r in res)
    return {"jobs": len(res), "jobs_per_s": len(res) / (last - first), "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)), "mean_iters": float(np.mean([r.iters for r in res])),
            "fits": [r.fit for r in res], "hbm_peak_gb": peak / 2 ** 30, "makespan_s": last - first,
            "total_nnz": int(sum(nnzs))}


def run_sequential(jobs, precision, max_iters=20, epsilon=1e-5):
    """The jobs one after another on the calling thread: create_alto and a
    device-seeded CP-ALS session per job, as a server worker runs them, with
    no admission or overlap (the path to profile kernel by kernel).  A job
    holds either host arrays (cidx, vals) or device tensors (coords, vals)."""
    L = lib()
    fits, iters = [], 0
    t0 = time.perf_counter()
    for j in jobs:
        if "cidx" in j:  # host COO (numpy)
            g = AltoTensor(j["dims"], j["cidx"], j["vals"], nprtn=1)
        else:  # device tensors from materialize()
            g = AltoTensor.from_device(j["dims"], [c.data_ptr() for c in j["coords"]], j["vals"].data_ptr(), j["nnz"],
                                       nprtn=1)
        h = C.c_void_p()
        check(L.alto_cpd_session_create_seeded(g.h, j["rank"], precision, j["seed"] % (2 ** 31), None, C.byref(h)),
              "cpd_session_create_seeded")
        fit = np.zeros(max_iters, np.float64)
        ran, conv = C.c_int(0), C.c_int(0)
        try:
            check(L.alto_cpd_session_iterate(h, max_iters, epsilon, fit.ctypes.data, C.byref(ran), C.byref(conv), 0),
                  "cpd_session_iterate")
        finally:
            L.alto_cpd_session_destroy(h)
        fits.append(float(fit[ran.value - 1]))
        iters += ran.value
        del g
    dt = time.perf_counter() - t0
    return {"jobs": len(jobs), "jobs_per_s": len(jobs) / dt, "mean_iters": iters / len(jobs), "fits": fits,
            "makespan_s": dt}


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description="CP-ALS job stream through the server or one job at a time")
    ap.add_argument("--jobs", type=int, default=100)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--precision", type=int, default=PREC_F64, help="0 FP64, 1 FP32, 2 BF16")
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--sequential", action="store_true", help="no server: jobs one after another")
    ap.add_argument("--time-scale", type=float, default=0.0, help="arrival-time scale (0: all jobs at once)")
    ap.add_argument("--lockstep-nnz", type=int, default=0,
                    help="jobs of at most this many nonzeros iterated in lockstep groups (0: none)")
    args = ap.parse_args(argv)
    jobs = materialize(job_stream(args.jobs, seed=args.seed))
    if args.sequential:
        r = run_sequential(jobs, args.precision)
    else:
        r = run_stream(jobs, args.precision, args.slots, time_scale=args.time_scale, lockstep_nnz=args.lockstep_nnz)
    r.pop("fits", None)
    print(json.dumps({"precision": args.precision, "sequential": args.sequential, "slots": args.slots, **r}),
          flush=True)


if __name__ == "__main__":
    main()
