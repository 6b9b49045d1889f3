#!/usr/bin/env python3
"""CP-ALS throughput of the MI355X ALTO engine on one GPU, with every
benchmark workload measured in the same run.

Headline workload: the 3-mode 12092 x 9184 x 28818 tensor with 76.9M unique
nonzeros, Zipf(1.1) indices over a seeded permutation of each mode and
lognormal(0,1) values (synthetic, generated on the GPU), FP64 values and
factors, 64-bit ALTO keys, rank 32.  A step is one CP-ALS iteration (the
MTTKRP, pseudo-inverse, normalisation and Gram of every mode plus the fit)
with the tensor and model resident in HBM.  Printed: one JSON line with
iterations/s, the all-mode MTTKRP roofline (algorithmic bytes / measured
kernel time vs 8 TB/s, and the HBM traffic of the same launches read from
the FETCH_SIZE / WRITE_SIZE counters in two rocprofv3 runs of this script),
FP32 / BF16 variants, the first iteration's fit against the oracle's, and
the host-core baseline on the same tensor.  Beside it, each other workload:
the 1000^3 1M-nonzero rank-16 CPU configuration on the GPU and on the host
cores, CP-APR on 120M Poisson counts (128-bit keys), the 5-mode 200M-nonzero
rank-64 precision study, the 2000-job serving stream and the batched MTTKRP.
"""
from __future__ import annotations

import argparse
import json

import numpy as np
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the serving leg runs 8 worker streams; HIP maps a process's streams onto
# this many hardware queues (4 by default), so give each stream its own.
# It must be set before the runtime initialises.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "8")

DIMS = [12092, 9184, 28818]
NNZ = 76_900_000
RANK = 32
SEED = 2
HBM_PEAK_GBS = 8000.0
COPY_PROBE_BYTES = 256 << 20


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """(threads this process may use, physical cores of the host, CPU model).
    The thread count is the process's CPU share: its affinity set, capped by
    OMP_NUM_THREADS when that is set."""
    share = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.split(",")[0].isdigit() and int(env.split(",")[0]) > 0:
        share = min(share, int(env.split(",")[0]))
    model, cores = "unknown", set()
    try:
        phys = core = None
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name":
                model = v
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
            elif not k and phys is not None:
                cores.add((phys, core))
                phys = core = None
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    return share, len(cores) or os.cpu_count(), model


def traffic_probe():
    """Profiling driver for the counter runs: a COPY_PROBE_BYTES device copy
    (the calibration kernel), then the headline tensor and three all-mode
    FP64 MTTKRP passes (the first also builds the mode streams)."""
    import torch

    import alto_amd as A
    from alto_amd import gen

    a = torch.ones(COPY_PROBE_BYTES // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    coords, vals = gen.sparse_coo(DIMS, NNZ, SEED, dist="zipf", s=1.1, values="lognormal")
    T = A.AltoTensor.from_device(DIMS, [c.data_ptr() for c in coords], vals.data_ptr(), NNZ, nprtn=16)
    del coords, vals
    U0, _ = A.kruskal_random_init(DIMS, RANK, SEED)
    Uf = [torch.from_numpy(u).cuda() for u in U0]
    outs = [torch.zeros(d, RANK, dtype=torch.float64, device="cuda") for d in DIMS]
    uptr = A._ptr_array([u.data_ptr() for u in Uf])
    stream = torch.cuda.current_stream().cuda_stream
    for _ in range(3):
        for m in range(3):
            A.check(A.lib().alto_mttkrp(T.h, m, RANK, uptr, outs[m].data_ptr(), A.PREC_F64, 1, stream), "mttkrp")
    torch.cuda.synchronize()


def measure_traffic():
    """HBM bytes per all-mode-pass MTTKRP call (mean over the three modes),
    read (FETCH_SIZE) and written (WRITE_SIZE), from two rocprofv3 counter
    runs of traffic_probe() in child processes.  Each counter is scaled by
    what it reads for the device copy of known size (FETCH_SIZE counts a
    copy's reads at half: scale 2.0 on this GPU).  The MTTKRP dispatches of
    the last pass (its kernels' names contain "mttkrp") are summed.  Returns
    a dict, or a dict with "error" when the profiler is missing or fails."""
    import csv
    import shutil
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return {"error": "rocprofv3 not found"}
    out = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="alto_traffic_", dir="/tmp")
        cmd = ["timeout", "-s", "KILL", "240", prof, "--pmc", counter, "-d", d, "-o", "run", "--output-format", "csv",
               "--", sys.executable, os.path.join(ROOT, "bench.py"), "--traffic-probe"]
        t0 = time.time()
        r = subprocess.run(cmd, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"), stdout=subprocess.DEVNULL,
                           stderr=subprocess.PIPE, text=True)
        files = [os.path.join(root, f) for root, _, fs in os.walk(d) for f in fs if f.endswith("counter_collection.csv")]
        log(f"traffic: rocprofv3 --pmc {counter}: exit {r.returncode} in {time.time() - t0:.1f}s")
        if not files:
            shutil.rmtree(d, ignore_errors=True)
            return {"error": f"rocprofv3 --pmc {counter} wrote no counters (exit {r.returncode}): "
                             f"{r.stderr.strip().splitlines()[-1:] if r.stderr else ''}"}
        rows = sorted(csv.DictReader(open(files[0])), key=lambda x: int(x["Dispatch_Id"]))
        shutil.rmtree(d, ignore_errors=True)
        copies = [float(x["Counter_Value"]) for x in rows if "copyBuffer" in x["Kernel_Name"]]
        mtt = [float(x["Counter_Value"]) for x in rows if "mttkrp" in x["Kernel_Name"]]
        if not copies or not mtt or len(mtt) % 3:
            return {"error": f"{counter}: {len(copies)} copy and {len(mtt)} MTTKRP dispatches"}
        scale = (COPY_PROBE_BYTES / 1024.0) / copies[-1]
        last = mtt[-len(mtt) // 3:]  # the last all-mode pass
        out[counter] = {"kib_last_pass": sum(last), "copy_kib": copies[-1], "scale": scale,
                        "bytes_per_call": sum(last) * 1024.0 * scale / 3, "dispatches_per_pass": len(last)}
    rd, wr = out["FETCH_SIZE"]["bytes_per_call"], out["WRITE_SIZE"]["bytes_per_call"]
    return {"read_bytes_per_call": rd, "write_bytes_per_call": wr, "bytes_per_call": rd + wr,
            "fetch_scale": out["FETCH_SIZE"]["scale"], "write_scale": out["WRITE_SIZE"]["scale"],
            "dispatches_per_pass": out["FETCH_SIZE"]["dispatches_per_pass"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nnz", type=int, default=NNZ)
    ap.add_argument("--no-cpu", action="store_true", help="skip the host-core baselines")
    ap.add_argument("--no-variants", action="store_true", help="skip the FP32/BF16 runs")
    ap.add_argument("--no-c1", action="store_true", help="skip the 1000^3 1M-nonzero rank-16 workload")
    ap.add_argument("--no-apr", action="store_true",
                    help="skip CP-APR on the 2^20 x 2^18 x 2^16 x 1024 Poisson count tensor (rank 10, 30 x 10)")
    ap.add_argument("--no-precision-study", action="store_true",
                    help="skip the 5-mode 50K x 50K x 10K x 2K x 512 Zipf(1.2) FP64 / FP32 / BF16 study (rank 64)")
    ap.add_argument("--no-traffic", action="store_true", help="skip the two rocprofv3 counter runs")
    ap.add_argument("--apr", action="store_true", help=argparse.SUPPRESS)  # on by default
    ap.add_argument("--precision-study", action="store_true", help=argparse.SUPPRESS)  # on by default
    ap.add_argument("--no-autotune", action="store_true", help="use the default MTTKRP launch shapes")
    ap.add_argument("--mttkrp-only", type=int, default=0, metavar="REPS",
                    help="profiling driver: build the tensor, run REPS all-mode MTTKRPs, print nothing else")
    ap.add_argument("--copy-only", action="store_true",
                    help="profiling calibration: a 256 MiB device copy (for FETCH_SIZE / WRITE_SIZE scaling)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="profiling driver of the traffic measurement: the calibration copy, then 3 all-mode MTTKRPs")
    ap.add_argument("--batched-only", action="store_true",
                    help="profiling driver: only the 16-job batched MTTKRP comparison, its JSON on stdout")
    ap.add_argument("--apr-only", action="store_true",
                    help="profiling driver: only the CP-APR workload, its JSON on stdout")
    ap.add_argument("--c1-only", action="store_true",
                    help="profiling driver: only the 1000^3 1M-nonzero rank-16 GPU runs, their JSON on stdout")
    ap.add_argument("--serve-jobs", type=int, default=2000,
                    help="jobs of the Poisson(50/s) serving stream to run (0 = skip; the full stream is 2000)")
    ap.add_argument("--tune", action="append", default=[], metavar="PARAM=VALUE",
                    help="alto_tune_set(PARAM, VALUE) before the run (alto_hip.h ALTO_TUNE_*), e.g. 13=0 for "
                         "streams without the fiber layout; repeatable")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported (the multi-GPU "
              "path is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    import torch

    import alto_amd as A
    from alto_amd import gen

    A.lib()
    for kv in args.tune:
        k, v = kv.split("=")
        A.check(A.lib().alto_tune_set(int(k), int(v)), f"alto_tune_set({kv})")
    stream = torch.cuda.current_stream().cuda_stream
    nnz = args.nnz
    if args.copy_only:
        a = torch.ones(32 << 20, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(4):
            b.copy_(a)
        torch.cuda.synchronize()
        return
    if args.traffic_probe:
        traffic_probe()
        return
    if args.batched_only:
        print(json.dumps(batched_leg()), flush=True)
        return
    if args.apr_only:
        print(json.dumps(apr_leg()), flush=True)
        return
    if args.c1_only:
        print(json.dumps(c1_leg(False)), flush=True)
        return

    t0 = time.time()
    coords, vals = gen.sparse_coo(DIMS, nnz, SEED, dist="zipf", s=1.1, values="lognormal")
    torch.cuda.synchronize()
    log(f"generated {nnz} nonzeros in {time.time() - t0:.2f}s")
    t0 = time.time()
    T = A.AltoTensor.from_device(DIMS, [c.data_ptr() for c in coords], vals.data_ptr(), nnz, nprtn=16)
    create_s = time.time() - t0
    del coords, vals
    torch.cuda.empty_cache()
    log(f"ALTO creation time:   {create_s:f}")

    U0, _ = A.kruskal_random_init(DIMS, RANK, SEED)
    Uf = [torch.from_numpy(u).cuda() for u in U0]
    outs = [torch.zeros(d, RANK, dtype=torch.float64, device="cuda") for d in DIMS]
    uptr = A._ptr_array([u.data_ptr() for u in Uf])

    def allmode():
        for m in range(3):
            A.check(A.lib().alto_mttkrp(T.h, m, RANK, uptr, outs[m].data_ptr(), A.PREC_F64, 1, stream), "mttkrp")

    if args.mttkrp_only:
        for _ in range(args.mttkrp_only):
            allmode()
        torch.cuda.synchronize()
        return
    # the first MTTKRP of each mode also builds that mode's stream (sorted
    # copy of the nonzeros); set-up = format + streams, the steady all-mode
    # time taken out
    bytes_format = T.info().device_bytes
    t0 = time.time()
    allmode()
    torch.cuda.synchronize()
    first_pass_s = time.time() - t0
    t0 = time.time()
    allmode()
    torch.cuda.synchronize()
    steady_pass_s = time.time() - t0
    streams_s = max(first_pass_s - steady_pass_s, 0.0)
    bytes_streams = T.info().device_bytes
    log(f"mode streams built in {streams_s:.3f}s; device bytes {bytes_format / 1e9:.2f} GB (format) -> "
        f"{bytes_streams / 1e9:.2f} GB (with streams)")

    def run_als(prec, steps, warmup):
        sess = A.CpdSession(T, U0, precision=prec, stream=stream)
        sess.iterate(warmup)
        torch.cuda.synchronize()
        t = time.perf_counter()
        fits, _ = sess.iterate(steps, epsilon=-1.0)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        _, _, fit, tm = sess.get()
        del sess
        return dt, fits, tm

    # the library's autotuner picks the MTTKRP launch shape per mode (kept on
    # the tensor; the CP-ALS sessions below use it)
    tuned = {}
    if not args.no_autotune:
        t0 = time.time()
        tuned["fp64"] = T.mttkrp_autotune(RANK, A.PREC_F64, reps=3, stream=stream)
        log(f"autotune fp64 ({time.time() - t0:.1f}s): " +
            ", ".join(f"mode {e['mode']}: v{e['variant']} seg {e['seg']} {e['ms']:.3f} ms (default {e['default_ms']:.3f})"
                      for e in tuned["fp64"]))
    # iteration 1 from the initial model, for the check against the oracle
    s1 = A.CpdSession(T, U0, precision=A.PREC_F64, stream=stream)
    fit1_gpu = float(s1.iterate(1)[0][0])
    del s1
    dt, fits, tm = run_als(A.PREC_F64, args.steps, args.warmup)
    ms_per_step = dt / args.steps * 1e3
    log(f"FP64 CP-ALS: {args.steps} iterations in {dt:.3f}s ({ms_per_step:.3f} ms/iter), fit {fits[-1]:.6f}")

    # all-mode MTTKRP roofline: algorithmic bytes per launch / kernel time;
    # each launch bracketed by its own events on the launch stream (torch's
    # current stream): the average gives `achieved`, the median is reported
    reps = 20
    mode_ms, mode_med_ms, mode_bytes = [], [], []
    for m in range(3):
        A.check(A.lib().alto_mttkrp(T.h, m, RANK, uptr, outs[m].data_ptr(), A.PREC_F64, 1, stream), "mttkrp")
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
        ev[0].record()
        for k in range(reps):
            A.check(A.lib().alto_mttkrp(T.h, m, RANK, uptr, outs[m].data_ptr(), A.PREC_F64, 1, stream), "mttkrp")
            ev[k + 1].record()
        torch.cuda.synchronize()
        per = [ev[k].elapsed_time(ev[k + 1]) for k in range(reps)]
        mode_ms.append(ev[0].elapsed_time(ev[reps]) / reps)
        mode_med_ms.append(float(np.median(per)))
        b = nnz * (8 + 8) + sum(DIMS[k] * RANK * 8 for k in range(3) if k != m) + DIMS[m] * RANK * 8
        mode_bytes.append(b)
    allmode_ms = sum(mode_ms)
    achieved = sum(mode_bytes) / (allmode_ms * 1e-3) / 1e9
    log(f"all-mode MTTKRP {allmode_ms:.3f} ms ({', '.join(f'{x:.3f}' for x in mode_ms)}), {achieved:.1f} GB/s")
    del Uf, outs
    copy_gbs = device_copy_gbs()
    log(f"device copy (1 GiB): {copy_gbs:.0f} GB/s")

    variants = {}
    if not args.no_variants:
        for name, prec in (("fp32", A.PREC_F32), ("bf16", A.PREC_BF16)):
            if not args.no_autotune:
                tuned[name] = T.mttkrp_autotune(RANK, prec, reps=3, stream=stream)
            vdt, vfits, _ = run_als(prec, args.steps, args.warmup)
            variants[name] = {"iters_per_s": args.steps / vdt, "ms_per_step": vdt / args.steps * 1e3,
                              "fit": float(vfits[-1]), "fit_delta_vs_fp64_gpu": float(abs(vfits[-1] - fits[-1]))}
            log(f"{name} CP-ALS: {vdt / args.steps * 1e3:.3f} ms/iter, fit {vfits[-1]:.6f}")

    cpu = fit1 = fitn = None
    if not args.no_cpu:
        n_total = args.warmup + args.steps
        cpu, fit1_oracle, ofits = cpu_baseline(T, U0, n_total)
        fit1 = {"gpu": fit1_gpu, "oracle": fit1_oracle, "abs_diff": abs(fit1_gpu - fit1_oracle)}
        log(f"iteration-1 fit: GPU {fit1_gpu:.15f} oracle {fit1_oracle:.15f} (|diff| {fit1['abs_diff']:.1e})")
        # the fit after the runs' warmup + timed iterations against the
        # oracle's FP64 run from the same model
        fitn = {"iterations": n_total, "oracle": float(ofits[-1]), "fp64_gpu": float(fits[-1]),
                "fp64_abs_diff": abs(float(fits[-1]) - float(ofits[-1]))}
        for name, v in variants.items():
            v["fit_delta_vs_oracle"] = abs(v["fit"] - float(ofits[-1]))
            fitn[name + "_abs_diff"] = v["fit_delta_vs_oracle"]
        log(f"fit after {n_total} iterations vs the oracle: " + ", ".join(
            f"{k} {v:.1e}" for k, v in fitn.items() if k.endswith("abs_diff")))
    del T
    torch.cuda.empty_cache()

    c1 = c1_leg(not args.no_cpu) if not args.no_c1 else None
    apr = apr_leg() if not args.no_apr else None
    prec_study = precision_leg() if not args.no_precision_study else None
    serving = batched = None
    if args.serve_jobs > 0:
        serving = serve_leg(args.serve_jobs, not args.no_cpu)
        batched = batched_leg()
    traffic = None
    if not args.no_traffic:
        traffic = measure_traffic()
        if "error" in traffic:
            log(f"traffic: {traffic['error']}")
        else:
            log(f"traffic: {traffic['read_bytes_per_call'] / 1e9:.3f} GB read + "
                f"{traffic['write_bytes_per_call'] / 1e9:.3f} GB written per MTTKRP call "
                f"(algorithmic {np.mean(mode_bytes) / 1e9:.3f} GB)")

    line = {
        "metric": "CP-ALS iters/s + MTTKRP HBM GB/s (frac peak); batched jobs/s, p99 latency",
        "value": args.steps / dt,
        "unit": "CP-ALS iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Zipf(1.1) indices over seeded permutations, lognormal values, generated on the GPU)",
        "config": {"workload": "cp-als nell2-shape zipf", "dims": DIMS, "nnz": nnz, "rank": RANK,
                   "alto_key_bits": 64, "modes": 3, "precision": "fp64"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic["bytes_per_call"] if traffic and "error" not in traffic else None,
                     "traffic_detail": traffic,
                     "kernel": "alto_mttkrp FP64 (k_mttkrp_fib long-fiber part + k_mttkrp16 rest), all modes",
                     "ms_per_launch": mode_ms, "median_ms_per_launch": mode_med_ms, "launches_timed": reps,
                     "algorithmic_bytes_per_launch": mode_bytes, "measured_copy_GBps": copy_gbs},
        "mttkrp_allmode_ms": allmode_ms,
        "autotune": tuned,
        "alto_create_s": create_s,
        "alto_create_with_streams_s": create_s + streams_s,
        "device_bytes": {"format": bytes_format, "with_mode_streams": bytes_streams},
        "cpd_timers_s": tm,
        "iteration1_fit_vs_oracle": fit1,
        "final_fit_vs_oracle": fitn,
        "precision_variants": variants,
        "c1_cpu_config": c1,
        "cp_apr": apr,
        "precision_study": prec_study,
        "serving": serving,
        "batched_mttkrp": batched,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def c1_leg(with_cpu):
    """The 1000 x 1000 x 1000 tensor with 1M unique uniform coordinates and
    uniform (0,1] values, 64-bit keys, rank 16: 20 CP-ALS iterations and 100
    mode-0 MTTKRPs (the command line's -p -t 0) on the GPU; on the host
    cores, the same MTTKRPs through the reference's own mttkrp_alto_par
    (oracle/_ref, built from the reference sources) and the 20 iterations
    through the oracle's C restatement."""
    import torch

    import alto_amd as A
    from alto_amd import gen

    dims, nnz, R = [1000, 1000, 1000], 1_000_000, 16
    coords, vals = gen.sparse_coo(dims, nnz, 1, dist="uniform", values="uniform")
    T = A.AltoTensor.from_device(dims, [c.data_ptr() for c in coords], vals.data_ptr(), nnz, nprtn=16)
    del coords, vals
    U0, _ = A.kruskal_random_init(dims, R, 1)
    stream = torch.cuda.current_stream().cuda_stream
    sess = A.CpdSession(T, U0, precision=A.PREC_F64, stream=stream)
    sess.iterate(1)
    del sess
    sess = A.CpdSession(T, U0, precision=A.PREC_F64, stream=stream)
    torch.cuda.synchronize()
    t = time.perf_counter()
    fits, _ = sess.iterate(20, epsilon=-1.0)
    torch.cuda.synchronize()
    als_s = time.perf_counter() - t
    del sess
    # the same 20 iterations, each replayed as a HIP graph (ALTO_TUNE_GRAPHS;
    # the first and the later iterations are captured once each, outside the
    # timed region, on a warm-up session)
    A.check(A.lib().alto_tune_set(5, 1), "tune_set")
    try:
        sess = A.CpdSession(T, U0, precision=A.PREC_F64, stream=stream)
        sess.iterate(2)
        del sess
        sess = A.CpdSession(T, U0, precision=A.PREC_F64, stream=stream)
        torch.cuda.synchronize()
        t = time.perf_counter()
        gfits, _ = sess.iterate(20, epsilon=-1.0)
        torch.cuda.synchronize()
        graph_s = time.perf_counter() - t
        del sess
    finally:
        A.check(A.lib().alto_tune_set(5, 0), "tune_set")
    Uf = [torch.from_numpy(u).cuda() for u in U0]
    out = torch.zeros(dims[0], R, dtype=torch.float64, device="cuda")
    uptr = A._ptr_array([u.data_ptr() for u in Uf])
    A.check(A.lib().alto_mttkrp(T.h, 0, R, uptr, out.data_ptr(), A.PREC_F64, 1, stream), "mttkrp")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(100):
        A.check(A.lib().alto_mttkrp(T.h, 0, R, uptr, out.data_ptr(), A.PREC_F64, 1, stream), "mttkrp")
    e1.record()
    torch.cuda.synchronize()
    mtt_s = e0.elapsed_time(e1) * 1e-3
    res = {"dims": dims, "nnz": nnz, "rank": R,
           "gpu": {"cpals_20_iters_s": als_s, "iters_per_s": 20 / als_s, "final_fit": float(fits[-1]),
                   "mttkrp_mode0_100_s": mtt_s, "mttkrp_mode0_ms": mtt_s * 10,
                   "graph_replay": {"cpals_20_iters_s": graph_s, "iters_per_s": 20 / graph_s,
                                    "final_fit_abs_diff_vs_eager": abs(float(gfits[-1]) - float(fits[-1]))}}}
    log(f"c1: GPU 20 CP-ALS iterations {als_s * 1e3:.1f} ms (replayed as graphs {graph_s * 1e3:.1f} ms), "
        f"100 mode-0 MTTKRPs {mtt_s * 1e3:.1f} ms")
    if with_cpu:
        import oracle
        threads, cores, model = host_cpu()
        oracle.lib().or_set_threads(threads)
        st = T.export()
        o = oracle.OracleAlto.wrap(dims, st["idx"].shape[1], T.nprtn, st["masks"], st["idx"], st["vals"],
                                   st["prtn_ptr"], st["intervals"], st["cr_masks"])
        U = [u.copy() for u in U0]
        lam = np.ones(R)
        t = time.time()
        _, ofits = o.cpd_als(U, lam, 20, -1.0)
        c_als = time.time() - t
        cpu = {"threads": threads, "host_physical_cores": cores, "cpu_model": model,
               "oracle_cpals_20_iters_s": c_als, "oracle_final_fit": float(ofits[-1]),
               "final_fit_abs_diff_gpu_vs_oracle": abs(float(ofits[-1]) - float(fits[-1]))}
        W = st["idx"].shape[1]
        inf = T.info()
        if oracle.ref_available(W):
            with _HostStdoutToNull():
                r = oracle.RefAlto.wrap(dims, W, T.nprtn, st["masks"], np.array(inf.alto_mask[:W], np.uint64),
                                        st["idx"], st["vals"], st["prtn_ptr"], st["intervals"], st["cr_masks"],
                                        np.array(inf.alto_cr_mask[:W], np.uint64))
                r.mttkrp(0, U0)
                t = time.time()
                for _ in range(100):
                    r.mttkrp(0, U0)
                cpu["reference_mttkrp_mode0_100_s"] = time.time() - t
                del r
        else:
            t = time.time()
            for _ in range(100):
                o.mttkrp(0, U0)
            cpu["oracle_mttkrp_mode0_100_s"] = time.time() - t
        res["cpu"] = cpu
        log(f"c1: host {threads} threads: oracle 20 iterations {c_als:.2f}s, 100 mode-0 MTTKRPs "
            f"{cpu.get('reference_mttkrp_mode0_100_s', cpu.get('oracle_mttkrp_mode0_100_s')):.2f}s")
    del T
    torch.cuda.empty_cache()
    return res


def apr_leg():
    """CP-APR on counts from a rank-10 Gamma(0.5) Kruskal model (about 120M
    nonzeros), 128-bit ALTO keys, integer values, rank 10, 30 outer x 10 inner
    iterations in bench mode (every inner iteration runs, as with -p)."""
    import torch

    import alto_amd as A
    from alto_amd import gen

    dims = [2 ** 20, 2 ** 18, 2 ** 16, 1024]
    t0 = time.time()
    coords, vals, _, _ = gen.poisson_kruskal_counts(dims, 10, 120_000_000, seed=3)
    nnz = int(vals.numel())
    T = A.AltoTensor.from_device(dims, [c.data_ptr() for c in coords], vals.data_ptr(), nnz, nprtn=16, lit_bits=128,
                                 val_kind=A.VAL_U64)
    del coords, vals
    torch.cuda.empty_cache()
    log(f"cp-apr: {nnz} nonzeros generated + ALTO in {time.time() - t0:.1f}s")
    U, lam = A.kruskal_random_init(dims, 10, 3)
    t = time.perf_counter()
    r = T.cp_apr(U, lam, 30, 10, epsilon=0.0, eps_div_zero=0.0, dobench=True)
    dt = time.perf_counter() - t
    # algorithmic bytes per Phi: the stream as stored (packed coordinates of
    # 8 B per 64 index bits, 8 B value) + every factor once + Phi once
    # (the output averaged over the modes)
    s_pk = 8 if T.info().alto_bits <= 64 else 16
    phi_calls = r["total_inner"]
    per_phi = r["t_phi"] / max(phi_calls, 1)
    avg_out = sum(d * 10 * 8 for d in dims) / len(dims)
    phi_gbs = (nnz * (s_pk + 8) + sum(d * 10 * 8 for d in dims) + avg_out) / per_phi / 1e9
    log(f"cp-apr: {dt:.2f}s total, {phi_calls} Phi calls, {per_phi * 1e3:.3f} ms/Phi ({phi_gbs:.0f} GB/s)")
    del T
    torch.cuda.empty_cache()
    return {"dims": dims, "nnz": nnz, "rank": 10, "outer": r["outer_iters"], "inner_total": phi_calls,
            "time_s": dt, "phi_time_s": r["t_phi"], "phi_ms_per_call": per_phi * 1e3, "phi_algorithmic_GBps": phi_gbs,
            "loglik": r["loglik"], "ls_fit": r["ls_fit"], "kkt": r["kkt"]}


def precision_leg():
    """5-mode 50K x 50K x 10K x 2K x 512 tensor, 200M nonzeros, Zipf(1.2) per
    mode, uniform (0,1) values, rank 64: per-mode MTTKRP time and relative error
    of FP32 and BF16-storage (FP32 accumulation) against FP64, then 25 CP-ALS
    iterations in each precision and the final-fit delta against FP64."""
    import torch

    import alto_amd as A
    from alto_amd import gen

    dims = [50000, 50000, 10000, 2000, 512]
    R = 64
    t0 = time.time()
    coords, vals = gen.sparse_coo(dims, 200_000_000, 4, dist="zipf", s=1.2, values="uniform")
    nnz = int(vals.numel())
    T = A.AltoTensor.from_device(dims, [c.data_ptr() for c in coords], vals.data_ptr(), nnz, nprtn=16, lit_bits=128)
    del coords, vals
    torch.cuda.empty_cache()
    log(f"precision: {nnz} nonzeros generated + ALTO in {time.time() - t0:.1f}s")
    stream = torch.cuda.current_stream().cuda_stream
    U0, _ = A.kruskal_random_init(dims, R, 4)
    res = {"dims": dims, "nnz": nnz, "rank": R, "mttkrp": {}}
    ref = {}
    for name, prec in (("fp64", A.PREC_F64), ("fp32", A.PREC_F32), ("bf16", A.PREC_BF16)):
        if prec == A.PREC_F64:
            F = [torch.from_numpy(u).cuda() for u in U0]
        elif prec == A.PREC_F32:
            F = [torch.from_numpy(u).float().cuda() for u in U0]
        else:
            F = [torch.from_numpy(u).float().cuda().to(torch.bfloat16) for u in U0]
        ptr = A._ptr_array([f.data_ptr() for f in F])
        per = []
        for m in range(len(dims)):
            out = torch.zeros(dims[m], R, dtype=torch.float64, device="cuda")
            A.check(A.lib().alto_mttkrp(T.h, m, R, ptr, out.data_ptr(), prec, 0, stream), "mttkrp")
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(3):
                A.check(A.lib().alto_mttkrp(T.h, m, R, ptr, out.data_ptr(), prec, 0, stream), "mttkrp")
            e1.record()
            torch.cuda.synchronize()
            o = out.double()
            if prec == A.PREC_F64:
                ref[m] = o
                err = 0.0
            else:
                err = float(torch.linalg.norm(o - ref[m]) / torch.linalg.norm(ref[m]))
            per.append({"ms": e0.elapsed_time(e1) / 3, "rel_err_vs_fp64": err})
        res["mttkrp"][name] = per
        del F
    # the fit of this workload is not a usable measure (rank 64 cannot fit the
    # 200M-nonzero tensor: the residual formula goes negative and the fit far
    # above 1), so the precisions are compared on the model itself: per mode,
    # ||U_p - U_fp64||_F / ||U_fp64||_F after the same k iterations (k = 1, 5,
    # 25; the iteration is not contracting on this tensor, so differences
    # grow with k), and the same for lambda
    checkpoints = (1, 5, 25)
    models = {}
    for name, prec in (("fp64", A.PREC_F64), ("fp32", A.PREC_F32), ("bf16", A.PREC_BF16)):
        sess = A.CpdSession(T, U0, precision=prec, stream=stream)
        dt, done, fits, models[name] = 0.0, 0, [], {}
        for k in checkpoints:
            torch.cuda.synchronize()
            t = time.perf_counter()
            f, _ = sess.iterate(k - done, epsilon=-1.0)
            torch.cuda.synchronize()
            dt += time.perf_counter() - t
            fits += list(f)
            done = k
            models[name][k] = sess.get()[:2]
        res[name + "_cpals"] = {"iters_per_s": done / dt, "final_fit": float(fits[-1])}
        del sess
    for name in ("fp32", "bf16"):
        r = res[name + "_cpals"]
        for k in checkpoints:
            U64, l64 = models["fp64"][k]
            Up, lp = models[name][k]
            r[f"factor_rel_diff_vs_fp64_after_{k}"] = [
                float(np.linalg.norm(Up[m] - U64[m]) / np.linalg.norm(U64[m])) for m in range(len(dims))]
            r[f"lambda_rel_diff_vs_fp64_after_{k}"] = float(np.linalg.norm(lp - l64) / np.linalg.norm(l64))
        r["final_fit_delta_vs_fp64"] = abs(r["final_fit"] - res["fp64_cpals"]["final_fit"])
    log(f"precision: {json.dumps(res)}")
    del T
    torch.cuda.empty_cache()
    return res


def device_copy_gbs(nbytes=1 << 30, reps=10):
    """Measured device-to-device copy rate (read + write bytes / time), the
    practical HBM ceiling beside the nominal 8 TB/s."""
    import torch
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    torch.cuda.empty_cache()
    return 2 * nbytes / (ms * 1e-3) / 1e9


# batched serving: jobs of up to this many nonzeros iterate in lockstep
# groups (400-job saturated stream, FP32: 87.1 jobs/s without, 91.2 with 200K,
# 88.3 with 1M, 84.8 with 4M)
SERVE_LOCKSTEP_NNZ = 200_000


def serve_leg(n_jobs, with_cpu):
    """The job stream (Poisson 50/s, 3-5 modes, dims log-uniform [64, 2^20],
    nnz log-uniform [1e4, 2e7], Zipf(1.1), rank in {8,16,32,64}, 20 iterations
    or fit change < 1e-5) under a 200 GB HBM budget: the plain FP64 per-job path
    (one job at a time) against the batched server (8 concurrent jobs, jobs of
    up to 200K nonzeros iterated in lockstep groups) in FP32 and BF16, at the
    stream's arrival times and saturated (every job available at once).  Each job's tensor is generated on the GPU just
    before its submission and freed when it finished.  With with_cpu, the
    first jobs of the stream also run through the oracle on the host cores
    (a bounded sample)."""
    import torch

    from alto_amd import PREC_BF16, PREC_F32, PREC_F64
    from alto_amd.serve import job_stream, run_stream

    jobs = job_stream(n_jobs, seed=5)
    out = {"jobs": n_jobs, "arrival_rate_per_s": 50.0, "hbm_budget_gb": 200,
           "stream_nnz_requested": int(sum(j["nnz"] for j in jobs))}
    out["lockstep_max_nnz"] = SERVE_LOCKSTEP_NNZ
    for name, prec, slots, scale in (("fp64_per_job", PREC_F64, 1, 1.0), ("fp32_batched", PREC_F32, 8, 1.0),
                                     ("bf16_batched", PREC_BF16, 8, 1.0),
                                     ("fp64_per_job_saturated", PREC_F64, 1, 0.0),
                                     ("fp32_batched_saturated", PREC_F32, 8, 0.0),
                                     ("bf16_batched_saturated", PREC_BF16, 8, 0.0)):
        t0 = time.time()
        r = run_stream(jobs, prec, slots, time_scale=scale, lockstep_nnz=SERVE_LOCKSTEP_NNZ if slots > 1 else 0)
        fits = np.array(r.pop("fits"), dtype=np.float64)
        out[name] = r
        fin = np.isfinite(fits)
        out[name + "_mean_fit"] = float(fits[fin].mean()) if fin.any() else None
        if prec == PREC_F64 and scale == 1.0:
            ref_fits = fits
        else:
            # each job's final fit against its plain FP64 run, relative to
            # max(1, |fit|) (these tensors' fits reach 1e5: the fit formula's
            # residual goes negative on them)
            rel = np.abs(fits - ref_fits) / np.maximum(1.0, np.abs(ref_fits))
            ok = np.isfinite(rel) & (rel <= 1e-3)
            out[name + "_fit_vs_fp64"] = {
                "jobs_within_1e-3": int(ok.sum()), "jobs": int(len(fits)), "nonfinite": int((~fin).sum()),
                "max_rel_delta_finite": float(rel[np.isfinite(rel)].max()) if np.isfinite(rel).any() else None,
                "jobs_outside": [int(i) for i in np.nonzero(~ok)[0][:10]]}
        chk = out.get(name + "_fit_vs_fp64")
        log(f"serving {name}: {r['jobs']} jobs ({r['total_nnz'] / 1e9:.2f}G nonzeros) {r['jobs_per_s']:.1f} jobs/s "
            f"p50 {r['p50_latency_s'] * 1e3:.1f} ms p99 {r['p99_latency_s'] * 1e3:.1f} ms ({time.time() - t0:.0f}s)" +
            (f"; fit within 1e-3 of FP64: {chk['jobs_within_1e-3']}/{chk['jobs']} (outside: {chk['jobs_outside']})"
             if chk else ""))
    for name in ("fp32", "bf16"):
        out[name + "_batched_speedup_vs_fp64"] = (out[name + "_batched"]["jobs_per_s"] /
                                                  out["fp64_per_job"]["jobs_per_s"])
        out[name + "_saturated_speedup_vs_fp64"] = (out[name + "_batched_saturated"]["jobs_per_s"] /
                                                    out["fp64_per_job_saturated"]["jobs_per_s"])
    if with_cpu:
        out["cpu_sample"] = cpu_serving_sample(jobs)
    torch.cuda.empty_cache()
    return out


def cpu_serving_sample(jobs, budget_s=20.0):
    """The stream's first jobs through the oracle's CP-ALS on the host cores
    (same tensor, generated on the GPU and copied over; same rank, 20
    iterations or fit change < 1e-5), until budget_s of host time is spent;
    jobs/s of that sample."""
    import oracle
    from alto_amd.serve import _generate

    threads, cores, model = host_cpu()
    oracle.lib().or_set_threads(threads)
    done, spent, nnz = 0, 0.0, 0
    for j in jobs:
        coords, vals, n = _generate(j)
        cidx = np.stack([c.cpu().numpy().astype(np.uint64) for c in coords])
        v = vals.cpu().numpy()
        del coords, vals
        t = time.time()
        o = oracle.OracleAlto(j["dims"], cidx, v, W=1 if sum(max((d - 1).bit_length(), 1) for d in j["dims"]) <= 64
                              else 2, nprtn=threads)
        U, lam = oracle.random_init(j["dims"], j["rank"], j["seed"] % (2 ** 31))
        o.cpd_als(U, lam, 20, 1e-5)
        spent += time.time() - t
        done += 1
        nnz += n
        del o
        if spent >= budget_s:
            break
    log(f"serving CPU sample: {done} jobs ({nnz / 1e6:.1f}M nonzeros) in {spent:.1f}s on {threads} threads")
    return {"jobs": done, "nnz": nnz, "host_s": spent, "jobs_per_s": done / spent if spent else None,
            "threads": threads, "host_physical_cores": cores, "cpu_model": model, "kind": "port",
            "sample": f"the stream's first {done} jobs (ALTO build + CP-ALS, 20 iterations or fit change < 1e-5) "
                      f"through the oracle, no arrival pacing"}


def batched_leg(n_jobs=16, reps=50):
    """Many small MTTKRPs: one alto_mttkrp_batched call against one
    alto_mttkrp call per job, on one stream, timed with events on that stream
    over `reps` back-to-back calls (host submission included).  16 tensors of
    3-5 modes, dims log-uniform in [64, 2^16], nnz log-uniform in [1e4, 2e5],
    Zipf(1.1), lognormal values; mode-0 MTTKRP accumulated into the output.
    Then CP-ALS on the same tensors, rank 16: each session one after another
    against all in lockstep, wall clock."""
    import ctypes as C

    import torch

    import alto_amd as A
    from alto_amd.gen import sparse_coo

    rng = np.random.default_rng(11)
    L = A.lib()
    s = torch.cuda.current_stream()
    tensors, keep = [], []
    for k in range(n_jobs):
        dims = [int(np.exp(rng.uniform(np.log(64), np.log(2 ** 16)))) for _ in range(3 + k % 3)]
        nnz = int(np.exp(rng.uniform(np.log(1e4), np.log(2e5))))
        coords, vals = sparse_coo(dims, nnz, seed=300 + k)
        nnz = int(vals.numel())  # fewer when the coordinate space is small
        coords = [c.to(torch.int64).contiguous() for c in coords]
        t = A.AltoTensor.from_device(dims, [c.data_ptr() for c in coords], vals.data_ptr(), nnz)
        torch.cuda.synchronize()
        del coords, vals
        tensors.append((t, dims, nnz))
    out = {"jobs": n_jobs, "total_nnz": int(sum(x[2] for x in tensors)), "calls_timed": reps}
    for R, prec, name in ((16, A.PREC_F64, "r16_fp64"), (32, A.PREC_F64, "r32_fp64"), (16, A.PREC_F32, "r16_fp32"),
                          (32, A.PREC_F32, "r32_fp32")):
        dt = torch.float64 if prec == A.PREC_F64 else torch.float32
        fptrs, outs = [], []
        for t, dims, _ in tensors:
            fs = [torch.rand((d, R), dtype=dt, device="cuda") for d in dims]
            arr = A._ptr_array([f.data_ptr() for f in fs])
            o = torch.zeros((dims[0], R), dtype=torch.float64, device="cuda")
            keep.append((fs, arr, o))
            fptrs.append(arr)
            outs.append(o)
        th = A._ptr_array([x[0].h for x in tensors])
        md = (C.c_int * n_jobs)(*([0] * n_jobs))
        fp = A._ptr_array([C.cast(a, C.c_void_p).value for a in fptrs])
        op = A._ptr_array([o.data_ptr() for o in outs])
        sp = s.cuda_stream

        def single():
            for (t, _, _), a, o in zip(tensors, fptrs, outs):
                A.check(L.alto_mttkrp(t.h, 0, R, a, o.data_ptr(), prec, 1, sp), "alto_mttkrp")

        def batched():
            A.check(L.alto_mttkrp_batched(n_jobs, th, md, R, fp, op, prec, 1, sp), "alto_mttkrp_batched")

        res = {}
        for fn_name, fn in (("per_job_calls_ms", single), ("batched_call_ms", batched)):
            for _ in range(3):
                fn()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            for _ in range(reps):
                fn()
            e1.record(s)
            torch.cuda.synchronize()
            res[fn_name] = e0.elapsed_time(e1) / reps
        res["speedup"] = res["per_job_calls_ms"] / res["batched_call_ms"]
        out[name] = res
        log(f"batched MTTKRP {name}: {n_jobs} jobs, per-job calls {res['per_job_calls_ms']:.3f} ms, "
            f"one batched call {res['batched_call_ms']:.3f} ms (x{res['speedup']:.2f})")
    del keep
    # CP-ALS on the same tensors: 20 iterations of each session one after
    # another against all sessions in lockstep (alto_cpd_sessions_iterate),
    # with the dense steps batched (the default) and as per-session chains
    for R, prec, name in ((16, A.PREC_F64, "cpals_r16_fp64"), (16, A.PREC_F32, "cpals_r16_fp32"),
                          (32, A.PREC_F64, "cpals_r32_fp64"), (64, A.PREC_F64, "cpals_r64_fp64")):
        inits = [A.kruskal_random_init(dims, R, 50 + k)[0] for k, (_, dims, _) in enumerate(tensors)]
        # first launches of this rank's kernels (code loading) outside the timing
        warm = [A.CpdSession(t, U, precision=prec) for (t, _, _), U in zip(tensors, inits)]
        for x in warm:
            x.iterate(1)
        A.cpd_sessions_iterate(warm, 1)
        del warm
        alone = [A.CpdSession(t, U, precision=prec) for (t, _, _), U in zip(tensors, inits)]
        lock = [A.CpdSession(t, U, precision=prec) for (t, _, _), U in zip(tensors, inits)]
        chains = [A.CpdSession(t, U, precision=prec) for (t, _, _), U in zip(tensors, inits)]
        graphed = [A.CpdSession(t, U, precision=prec) for (t, _, _), U in zip(tensors, inits)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for x in alone:
            x.iterate(20)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        A.cpd_sessions_iterate(lock, 20)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        A.check(L.alto_tune_set(20, 0), "tune_set")  # ALTO_TUNE_LOCKSTEP_DENSE_BATCH
        try:
            t3 = time.perf_counter()
            A.cpd_sessions_iterate(chains, 20)
            torch.cuda.synchronize()
            t4 = time.perf_counter()
        finally:
            A.check(L.alto_tune_set(20, 1), "tune_set")
        A.check(L.alto_tune_set(21, 1), "tune_set")  # ALTO_TUNE_LOCKSTEP_GRAPHS
        try:
            t5 = time.perf_counter()
            A.cpd_sessions_iterate(graphed, 20)
            torch.cuda.synchronize()
            t6 = time.perf_counter()
        finally:
            A.check(L.alto_tune_set(21, 0), "tune_set")
        dfit = max(max(abs(x.get()[2] - y.get()[2]), abs(x.get()[2] - z.get()[2]), abs(x.get()[2] - w.get()[2]))
                   for x, y, z, w in zip(alone, lock, chains, graphed))
        out[name] = {"iterations": 20, "one_after_another_ms": (t1 - t0) * 1e3, "lockstep_ms": (t2 - t1) * 1e3,
                     "lockstep_per_session_chains_ms": (t4 - t3) * 1e3, "lockstep_graph_replay_ms": (t6 - t5) * 1e3,
                     "speedup": (t1 - t0) / (t2 - t1), "max_fit_delta": dfit}
        log(f"lockstep CP-ALS {name}: {n_jobs} sessions x 20 iterations, one after another {(t1 - t0) * 1e3:.1f} ms, "
            f"lockstep {(t2 - t1) * 1e3:.1f} ms (x{(t1 - t0) / (t2 - t1):.2f}; with per-session dense chains "
            f"{(t4 - t3) * 1e3:.1f} ms; replayed as a graph {(t6 - t5) * 1e3:.1f} ms), max fit delta {dfit:.1e}")
        del alone, lock, chains, graphed
    del tensors
    torch.cuda.empty_cache()
    return out


class _HostStdoutToNull:
    """Sends what C code prints to stdout (the reference harness reports its
    partition buffers) to /dev/null, so that stdout carries only the JSON line."""

    def __enter__(self):
        import ctypes
        self.libc = ctypes.CDLL(None)
        sys.stdout.flush()
        self.libc.fflush(None)
        self.saved = os.dup(1)
        self.null = os.open(os.devnull, os.O_WRONLY)
        os.dup2(self.null, 1)
        return self

    def __exit__(self, *exc):
        self.libc.fflush(None)
        os.dup2(self.saved, 1)
        os.close(self.saved)
        os.close(self.null)
        return False


def cpu_baseline(T, U0, n_total):
    """One CP-ALS iteration on the host cores (this process's CPU share), on
    the same tensor exported from the GPU; returns (baseline, the
    iteration's fit, the oracle's fit trace over n_total iterations from the
    same model -- a parity check, outside the timing).  With the reference harness built (oracle/_ref: the
    reference's include/alto.hpp compiled from its sources), the iteration's
    MTTKRPs are the reference's own mttkrp_alto_par and the rest (R x R
    solves, norms, Grams, fit: the reference's driver needs BLAS / LAPACK,
    absent here) the oracle's C restatement; without it, the oracle's whole
    iteration."""
    import numpy as np

    import oracle

    threads, cores, model = host_cpu()
    oracle.lib().or_set_threads(threads)
    st = T.export()
    inf = T.info()
    t0 = time.time()
    o = oracle.OracleAlto.wrap(DIMS, st["idx"].shape[1], T.nprtn, st["masks"], st["idx"], st["vals"],
                               st["prtn_ptr"], st["intervals"], st["cr_masks"])
    U = [u.copy() for u in U0]
    lam = np.ones(RANK)
    t1 = time.time()
    it, fit = o.cpd_als(U, lam, 1, 0.0)
    dt1 = time.time() - t1
    # n_total iterations from the same model: the fit trace for the parity
    # check, and the oracle's iteration time averaged over them
    t4 = time.time()
    _, ofits = o.cpd_als([u.copy() for u in U0], np.ones(RANK), n_total, 0.0)
    per_iter = (time.time() - t4) / n_total
    log(f"CPU oracle: 1 CP-ALS iteration in {dt1:.2f}s, {n_total} iterations in {per_iter * n_total:.1f}s "
        f"({per_iter:.3f}s each) on {threads} threads (wrap {t1 - t0:.1f}s)")
    port = {"value": 1.0 / per_iter, "unit": "CP-ALS iterations/s", "cores": threads, "kind": "port",
            "host_physical_cores": cores, "cpu_model": model,
            "sample": f"{n_total} CP-ALS iterations (rank {RANK}) on the full {inf.nnz}-nonzero tensor, averaged, "
                      f"OpenMP over {threads} threads (this process's CPU share), nprtn {T.nprtn}"}
    W = st["idx"].shape[1]
    if not oracle.ref_available(W):
        return port, float(fit[0]), ofits

    def best_of(reps, fn):
        best = float("inf")
        for _ in range(reps):
            t = time.time()
            fn()
            best = min(best, time.time() - t)
        return best

    t_oracle_mttkrp = best_of(2, lambda: [o.mttkrp(m, U) for m in range(3)])
    with _HostStdoutToNull():
        r = oracle.RefAlto.wrap(DIMS, W, T.nprtn, st["masks"], np.array(inf.alto_mask[:W], np.uint64), st["idx"],
                                st["vals"], st["prtn_ptr"], st["intervals"], st["cr_masks"],
                                np.array(inf.alto_cr_mask[:W], np.uint64))
        r.mttkrp(0, U)  # first touch of its buffers
        t_ref_mttkrp = best_of(2, lambda: [r.mttkrp(m, U) for m in range(3)])
        del r
    dense = max(per_iter - t_oracle_mttkrp, 0.0)
    t_iter = dense + t_ref_mttkrp
    log(f"CPU reference: all-mode mttkrp_alto_par {t_ref_mttkrp:.3f}s (oracle {t_oracle_mttkrp:.3f}s, best of 2); "
        f"oracle dense steps {dense:.3f}s; iteration {t_iter:.3f}s")
    return {"value": 1.0 / t_iter, "unit": "CP-ALS iterations/s", "cores": threads, "kind": "reference",
            "host_physical_cores": cores, "cpu_model": model,
            "sample": f"1 CP-ALS iteration (rank {RANK}) on the full {inf.nnz}-nonzero tensor, OpenMP over {threads} "
                      f"threads, nprtn {T.nprtn}: the reference's mttkrp_alto_par for the three MTTKRPs "
                      f"({t_ref_mttkrp:.3f} s, best of 2) plus the oracle's solves, norms, Grams and fit "
                      f"({dense:.3f} s: its iteration averaged over {n_total} less its three MTTKRPs, best of 2)",
            "reference_allmode_mttkrp_s": t_ref_mttkrp, "port": port}, float(fit[0]), ofits


if __name__ == "__main__":
    main()
