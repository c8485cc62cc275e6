#!/usr/bin/env python3
"""Headline benchmark: SpMV on the power-law matrix (n = 2,097,152, row lengths ~ L^-2.1 on
[3, 8192], columns uniform without replacement, values N(0,1)) on one MI355X.

A step is one y = A*x of that matrix with A, x, y resident in HBM.  The headline line is the
FP32-value SELL-C-sigma (C = 32, sigma = 1024) SpMV, in achieved GB/s of algorithmic bytes
((n+1)*4 + nnz*(4+4) + n*4 (x) + n*4 (y) + n*4 (reorder) + chunks*4); the FP64 CSR drop-in,
FP64 SELL and BF16 SELL run on the same matrix and are reported beside it with their error
against an FP64 product (scipy, float64).

Prints ONE JSON line (rank 0, N = 1).  --gpus N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "SpMV achieved GB/s (% HBM peak), SpGEMM GFLOP/s; mixed vs FP64 at rel err"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def errors(y, y64, absrow):
    import numpy as np
    d = np.abs(y.astype(np.float64) - y64)
    cond = d / np.maximum(absrow, 1e-300)
    return {"max_rel_err": float(cond.max()), "rms_rel_err": float(np.sqrt(np.mean(cond ** 2))),
            "normwise_rel_err": float(np.linalg.norm(d) / np.linalg.norm(y64))}


def time_spmv(M, x, y, steps, warmup, stream):
    """Returns (ms per step from the wall clock around K steps, mean kernel ms from HIP events)."""
    import torch
    from sparseops_amd import _lib as L
    import ctypes
    for _ in range(warmup):
        M.spmv(x, y, 1.0, 0.0, stream)
    torch.cuda.synchronize()
    ev = [ctypes.c_void_p() for _ in range(steps + 1)]
    for e in ev:
        L.call("spops_event_create", ctypes.byref(e))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L.call("spops_event_record", ev[0], stream)
    for i in range(steps):
        M.spmv(x, y, 1.0, 0.0, stream)
        L.call("spops_event_record", ev[i + 1], stream)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3 / steps
    ms = ctypes.c_float()
    tot = 0.0
    for i in range(steps):
        L.call("spops_event_elapsed_ms", ctypes.byref(ms), ev[i], ev[i + 1])
        tot += ms.value
    for e in ev:
        L.call("spops_event_destroy", e)
    return wall_ms, tot / steps


def cpu_baseline(A, seconds, threads):
    """The reference's own OpenMP SELL-C-sigma SpMV (kernel_flag 2, FP32, C=32, sigma=1024),
    built from its sources into oracle/_ref/ref_harness, on this host's cores."""
    from sparseops_amd.io import write_csr_bin
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if not os.path.exists(exe):
        return None
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        path = os.path.join(d, "A.csr")
        write_csr_bin(path, A)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PLACES="cores", OMP_PROC_BIND="spread")
        out = subprocess.run([exe, "bench_spmv", path, "sell", "32", str(threads), str(seconds), "1024", "32"],
                             capture_output=True, text=True, env=env, timeout=600)
    if out.returncode != 0:
        log(f"cpu baseline failed: {out.stderr[-500:]}")
        return None
    return json.loads(out.stdout.strip().splitlines()[-1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1 << 21)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "16") or 16))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--variants", default="sell_f32,csr_f64,sell_f64,sell_bf16,csr_f32",
                    help="comma list of <fmt>_<value type>; the first one is the headline")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="only run a device copy of this many MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is described in DESIGN.md, not run.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import scipy.sparse as sp
    import torch
    from sparseops_amd import spmv as S
    from sparseops_amd.gen import powerlaw, uniform_x

    dev = torch.device("cuda:0")
    if args.copy_calibration_mb:
        a = torch.ones(args.copy_calibration_mb << 18, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4, "launches": 3}))
        return
    t = time.time()
    A = powerlaw(args.n, seed=0)
    x64 = uniform_x(A.n_cols, 0)
    log(f"matrix n={A.n_rows} nnz={A.nnz} generated in {time.time() - t:.1f}s")
    As = sp.csr_matrix((A.val, A.col, A.rowptr), shape=(A.n_rows, A.n_cols))
    y64 = As @ x64
    absrow = abs(As) @ np.abs(x64)

    stream = torch.cuda.current_stream().cuda_stream
    variants = [(v, v.split("_")[0], v.split("_")[1]) for v in args.variants.split(",")]
    head_name = variants[0][0]
    res = {}
    for name, fmt, vt in variants:
        t = time.time()
        M = S.csr(A, vt) if fmt == "csr" else S.csr_to_sell_c_sigma(A, 1024, 32, vt)
        dt = M.xy_dtype
        x = torch.tensor(x64, dtype=dt, device=dev)
        y = torch.zeros(A.n_rows, dtype=dt, device=dev)
        wall_ms, kern_ms = time_spmv(M, x, y, args.steps, args.warmup, stream)
        yh = y.double().cpu().numpy()
        r = {"wall_ms": wall_ms, "kernel_ms": kern_ms, "algo_bytes": M.algo_bytes,
             "gbs": M.algo_bytes / (wall_ms * 1e-3) / 1e9, "kernel_gbs": M.algo_bytes / (kern_ms * 1e-3) / 1e9,
             "device_bytes": M.device_bytes, "build_s": time.time() - t}
        r.update(errors(yh, y64, absrow))
        res[name] = r
        log(f"{name}: {r['kernel_ms'] * 1e3:.1f} us/kernel, {r['kernel_gbs']:.0f} GB/s, "
            f"max rel err {r['max_rel_err']:.2e}, normwise {r['normwise_rel_err']:.2e}")
        M.close()
        del x, y

    head = res[head_name]
    cpu = None
    if not args.no_cpu:
        t = time.time()
        c = cpu_baseline(A, args.cpu_seconds, args.cpu_threads)
        if c:
            cpu = {"value": head["algo_bytes"] / c["seconds_per_iter"] / 1e9, "unit": "GB/s",
                   "cores": c["threads"], "kind": "reference",
                   "sample": f"the full matrix, {c['iterations']} SpMV iterations of the reference's OpenMP "
                             f"SELL-C-sigma kernel (kernel_flag 2, FP32, C=32, sigma=1024), "
                             f"{c['seconds_per_iter'] * 1e3:.2f} ms each"}
        log(f"cpu baseline done in {time.time() - t:.1f}s: {cpu}")

    out = {
        "metric": METRIC,
        "value": round(head["gbs"], 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["wall_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "SELL-C-sigma SpMV, FP32 values + FP32 accumulation, power-law n=2,097,152 "
                               f"(nnz={A.nnz}, L^-2.1 on [3,8192]), C=32, sigma=1024",
                   "n": A.n_rows, "nnz": A.nnz, "C": 32, "sigma": 1024},
        "pct_hbm_peak": round(100.0 * head["gbs"] / HBM_PEAK_GBS, 2),
        "roofline": {"bound": "hbm", "achieved": round(head["kernel_gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(head["kernel_gbs"] / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": cpu,
        "rel_err_vs_fp64": {k: {e: v[e] for e in ("max_rel_err", "rms_rel_err", "normwise_rel_err")}
                            for k, v in res.items()},
        "variants": {k: {"gbs": round(v["gbs"], 1), "kernel_us": round(v["kernel_ms"] * 1e3, 2),
                         "pct_hbm_peak": round(100 * v["kernel_gbs"] / HBM_PEAK_GBS, 2)} for k, v in res.items()},
        "fp32_sell_over_fp64_csr": (round(res["csr_f64"]["kernel_ms"] / res["sell_f32"]["kernel_ms"], 3)
                                    if "csr_f64" in res and "sell_f32" in res else None),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
