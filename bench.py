#!/usr/bin/env python3
"""Headline benchmark: SpMV on the power-law matrix (n = 2,097,152, row lengths ~ L^-2.1 on
[3, 8192], columns uniform without replacement, values N(0,1)) on one MI355X.

A step is one y = A*x of that matrix with A, x, y resident in HBM.  The headline line is the
FP32-value SELL-C-sigma (C = 32, sigma = 1024) SpMV, in achieved GB/s of algorithmic bytes
((n+1)*4 + nnz*(4+4) + n*4 (x) + n*4 (y) + n*4 (reorder) + chunks*4); the FP64 CSR drop-in,
FP64 SELL, BF16 SELL, FP32 CSR and the FP64 CSR row kernels run on the same matrix and are
reported beside it with their error against an FP64 product (scipy, float64).  The SELL and
CSR handles of this matrix use the swept device layout for SpMV (row blocks summed in FP64 in
LDS, each block's entries sorted by column).  The roofline object covers one SpMV of the headline variant (its two launches:
the swept kernel and the pass that adds the two partial planes), timed with HIP events on
the launch stream; its HBM traffic comes from the committed rocprofv3 counter summary of
the same command.  A secondary LeSpGEMM measurement (C = A*A,
R-MAT, MinHash-LSH clustering, GFLOP/s) is reported with the reference's OpenMP pipeline
timed on the same host; so are the differentiable SpMM layer (R-MAT scale 20, BF16 X / G,
forward and backward), the batched SpMM (512 band matrices, one HIP-graph replay per step)
and the plain FP64 CSR SpMV on the 5-point Laplacian of a 256 x 256 grid (device-resident,
through the drop-in driver, and the reference's OpenMP CSR kernel on the same host).

Prints ONE JSON line (rank 0, N = 1).  --gpus N > 1 exits: the benchmark runs on one GPU.
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
HOST = {}            # the CPU baselines' host (host_cpu()), reported with each of them
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


def host_cpu():
    """The host the CPU baselines run on: logical CPUs (nproc), the CPUs this process may run
    on, the physical cores among them (distinct (socket, core) pairs of /proc/cpuinfo), the
    sockets, the CPU model, and the OpenMP thread count the baselines use: one per physical
    core this process may use, capped by OMP_NUM_THREADS when that is set."""
    allowed = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count() or 1))
    model, cores, sockets = "unknown", set(), set()
    try:
        cur = {}
        with open("/proc/cpuinfo") as f:
            for line in list(f) + ["\n"]:
                if not line.strip():
                    if "processor" in cur and int(cur["processor"]) in allowed:
                        cores.add((cur.get("physical id", "0"), cur.get("core id", cur["processor"])))
                        sockets.add(cur.get("physical id", "0"))
                        model = cur.get("model name", model)
                    cur = {}
                    continue
                k, _, v = line.partition(":")
                cur[k.strip()] = v.strip()
    except OSError:
        pass
    phys = len(cores) or len(allowed)
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(phys, int(env)) if env.isdigit() and int(env) > 0 else phys
    return {"nproc": os.cpu_count(), "allowed_cpus": len(allowed), "physical_cores": phys,
            "sockets": max(1, len(sockets)), "cpu_model": model, "threads": threads}


def ref_bytes_per_spmv(fmt, vt, A, M):
    """The reference's own GB/s model, bytes_per_spmv (include/spmv_benchmark.h:43-51 for CSR,
    :107-122 for SELL-C-sigma, which counts each chunk's values with sizeof(IndexType)), for
    int32 indices; None for BF16 values, which the reference does not have."""
    if vt == "bf16":
        return None
    sI, sV, n, nnz = 4, 8 if vt == "f64" else 4, A.n_rows, A.nnz
    if fmt in ("csr", "csrrows"):
        return 2 * sI * n + sI * nnz + 2 * sV * nnz + 2 * sV * n
    cl = M.layout(1).astype("int64")
    C = M.info[10]
    return sI * len(cl) + sI * n + int(cl.sum()) * C * (sI + sI) + sV * nnz + 2 * sV * n


def cpu_env(threads):
    return dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PLACES="cores", OMP_PROC_BIND="spread")


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
        out = subprocess.run([exe, "bench_spmv", path, "sell", "32", str(threads), str(seconds), "1024", "32"],
                             capture_output=True, text=True, env=cpu_env(threads), timeout=600)
    if out.returncode != 0:
        log(f"cpu baseline failed: {out.stderr[-500:]}")
        return None
    return json.loads(out.stdout.strip().splitlines()[-1])


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "spmv_sell_f32_powerlaw_n2097152_hbm_traffic.txt")


def committed_traffic():
    """HBM bytes per SpMV from the committed counter summary: per kernel FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch) scaled by the copy-kernel calibration factors."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    vals = {}
    with open(TRAFFIC_FILE) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_spmv"):
                vals["bytes"] = float(line.split("=")[1])
    return vals.get("bytes")


def spgemm_roofline(nnz_a, nnz_b, n_a, n_b, nnz_c, n_c, ms):
    """Roofline of one C = A * B: compulsory bytes = A, B and C once each as int64 + FP64 CSR
    ((rows + 1) * 8 + nnz * 16 each) over the time of the whole product (plan + panels)."""
    comp = (n_a + 1) * 8 + nnz_a * 16 + (n_b + 1) * 8 + nnz_b * 16 + (n_c + 1) * 8 + nnz_c * 16
    gbs = comp / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None, "compulsory_bytes": int(comp)}


def lespgemm_measure(scale, iterations, cpu_threads, cpu):
    """LeSpGEMM C = A*A on R-MAT(scale, edge factor 16), k=64, bands=16, seed=7,
    cluster_size=8, hc_v=0: GPU pipeline and, unless cpu is False, the reference's own
    OpenMP pipeline (kernel 3) on the same matrix."""
    import torch
    from sparseops_amd import lespgemm as LG
    from sparseops_amd.gen import rmat
    from sparseops_amd.io import write_csr_bin
    A = rmat(scale, 16, seed=0)
    r = LG.lespgemm(A, A, iterations=iterations)
    gpu_ms = sum(r["runs_ms"]) / len(r["runs_ms"])
    out = {"workload": f"R-MAT scale {scale} (n={A.n_rows}, nnz={A.nnz}), C = A*A, kernel 3, cluster_size 8, "
                       f"k=64, bands=16, seed=7, hc_v=0",
           "flop": int(r["flop"]), "gpu_ms": round(gpu_ms, 3), "gpu_gflops": round(r["flop"] / gpu_ms / 1e6, 2),
           "gpu_genpairs_ms": round(r["times"]["genpairs_ms"], 2), "gpu_hc_ms": round(r["times"]["hc_ms"], 2),
           "gpu_conversion_ms": round(r["times"]["conversion_ms"], 2), "nnz_C": int(len(r["C_col"])),
           "clusters": int(r["nclusters"]),
           "roofline": spgemm_roofline(A.nnz, A.nnz, A.n_rows, A.n_rows, len(r["C_col"]), A.n_rows, gpu_ms)}
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if cpu and os.path.exists(exe):
        with tempfile.TemporaryDirectory(dir="/tmp") as d:
            path = os.path.join(d, "A.csr")
            write_csr_bin(path, A)
            env = cpu_env(cpu_threads)
            p = subprocess.run([exe, "bench_lespgemm", path, path, str(cpu_threads), "3", "64", "16", "7", "8", "3"],
                               capture_output=True, text=True, env=env, timeout=900)
        if p.returncode == 0:
            c = json.loads(p.stdout.strip().splitlines()[-1])
            out["cpu_reference"] = {**HOST, "gflops": round(c["gflops"], 3), "lespgemm_ms": c["lespgemm_ms"],
                                    "genpairs_ms": c["genpairs_ms"], "hc_ms": c["hc_ms"],
                                    "conversion_ms": c["conversion_ms"], "threads": c["threads"],
                                    "kind": "reference", "iterations": 3}
            out["gpu_over_cpu"] = round(out["gpu_gflops"] / max(c["gflops"], 1e-12), 2)
            # cluster-C column counts of both pipelines (equal when clusterings and structure match)
            out["nnzc_gpu"], out["nnzc_cpu"] = int(r["stats"]["nnzc"]), int(c["nnzc"])
        else:
            log(f"cpu lespgemm failed: {p.stderr[-400:]}")
    torch.cuda.synchronize()
    return out


def rowwise_topk_measure(scale, iterations, cpu_threads, cpu):
    """The row-wise LeSpGEMM C = A*A (the reference's LeSpGEMM<sortOutput = true>, kernels 1-2;
    spops_spgemm_csr on the device: A's duplicates merged, one row per cluster, the cluster
    pipeline, CSR out) and the A*A^T top-k Jaccard candidate front end (HashSpGEMMTopK, top 7;
    spops_spgemm_topk_jaccard) on R-MAT(scale, edge factor 16), FP64, device arrays in and out,
    timed with HIP events around each call (output allocation included), beside the reference's
    OpenMP kernels on this host's cores."""
    import ctypes
    import torch
    from sparseops_amd import _lib as L
    from sparseops_amd import lespgemm as LG
    from sparseops_amd.gen import rmat
    from sparseops_amd.io import write_csr_bin
    A = rmat(scale, 16, seed=0)
    dev = torch.device("cuda:0")
    dA = LG.DeviceCsr(A, dev)
    flop = LG.spgemm_flop(dA, dA)
    stream = L.current_stream()
    ev = [ctypes.c_void_p() for _ in range(2)]
    for e in ev:
        L.call("spops_event_create", ctypes.byref(e))

    def timed(call):
        """Mean ms of `iterations` calls after one warm-up; each returns device outputs, freed after."""
        tot, nnz = 0.0, 0
        for it in range(iterations + 1):
            outs = [ctypes.c_void_p() for _ in range(3)]
            n = ctypes.c_int64()
            torch.cuda.synchronize()
            L.call("spops_event_record", ev[0], stream)
            call(n, outs)
            L.call("spops_event_record", ev[1], stream)
            torch.cuda.synchronize()
            ms = ctypes.c_float()
            L.call("spops_event_elapsed_ms", ctypes.byref(ms), ev[0], ev[1])
            if it > 0:
                tot += ms.value
            nnz = n.value
            for o in outs:
                L.lib.spops_free(o)
        return tot / iterations, nnz

    st = (ctypes.c_int64 * 16)()
    rw_ms, rw_nnz = timed(lambda n, o: L.call(
        "spops_spgemm_csr", dA.n_rows, dA.n_cols, dA.nnz, dA.rowptr.data_ptr(), dA.col.data_ptr(), dA.val.data_ptr(),
        dA.n_cols, dA.nnz, dA.rowptr.data_ptr(), dA.col.data_ptr(), dA.val.data_ptr(), L.VALUE_F64,
        ctypes.byref(n), ctypes.byref(o[0]), ctypes.byref(o[1]), ctypes.byref(o[2]), st, stream))
    tk_ms, tk_nnz = timed(lambda n, o: L.call(
        "spops_spgemm_topk_jaccard", dA.n_rows, dA.n_cols, dA.nnz, dA.rowptr.data_ptr(), dA.col.data_ptr(),
        L.VALUE_F64, 7, 0, ctypes.byref(n), ctypes.byref(o[0]), ctypes.byref(o[1]), ctypes.byref(o[2]), stream))
    for e in ev:
        L.call("spops_event_destroy", e)
    out = {"workload": f"R-MAT scale {scale} (n={A.n_rows}, nnz={A.nnz}), FP64: row-wise C = A*A and the A*A^T "
                       f"top-7 Jaccard candidates",
           "flop": int(flop), "rowwise_gpu_ms": round(rw_ms, 3), "rowwise_gpu_gflops": round(flop / rw_ms / 1e6, 2),
           "rowwise_nnz_C": int(rw_nnz), "topk_gpu_ms": round(tk_ms, 3), "topk_nnz": int(tk_nnz),
           "iterations": iterations}
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if cpu and os.path.exists(exe):
        with tempfile.TemporaryDirectory(dir="/tmp") as d:
            path = os.path.join(d, "A.csr")
            write_csr_bin(path, A)
            env = cpu_env(cpu_threads)
            for name, args in (("rowwise_kernel1", ["bench_rowwise", path, path, str(cpu_threads), "1", "1"]),
                               ("rowwise_kernel2", ["bench_rowwise", path, path, str(cpu_threads), "1", "2"]),
                               ("topk", ["bench_topk", path, str(cpu_threads), "1", "7"])):
                p = subprocess.run([exe] + args, capture_output=True, text=True, env=env, timeout=900)
                if p.returncode != 0:
                    log(f"cpu {name} failed: {p.stderr[-300:]}")
                    continue
                c = json.loads(p.stdout.strip().splitlines()[-1])
                ms = c["seconds"] * 1e3
                o = {**HOST, "ms": round(ms, 2), "threads": c["threads"], "kind": "reference", "nnz_C": c["nnz_C"]}
                if name.startswith("rowwise"):
                    o["gflops"] = round(flop / ms / 1e6, 3)
                    o["gpu_over_cpu"] = round(ms / rw_ms, 1)
                    o["nnz_C_equal"] = c["nnz_C"] == rw_nnz
                else:
                    o["gpu_over_cpu"] = round(ms / tk_ms, 1)
                    o["nnz_equal"] = c["nnz_C"] == tk_nnz
                out["cpu_" + name] = o
    del dA
    torch.cuda.synchronize()
    return out


def lespgemm_panels_measure(scale, budget_gb, cpu_threads, cpu, cpu_share=0.02):
    """LeSpGEMM C = A*A on R-MAT(scale, edge factor 16) (k=64, bands=16, seed=7,
    cluster_size=8, hc_v=0, kernel 3) when C is too large for one piece: the plan (B
    preparation + symbolic phase of every cluster) and the numeric phase in cluster panels
    whose output stays under budget_gb, output buffers reused from panel to panel.  Timed
    twice; the second (warm) pass is reported.  The CPU leg runs the reference's OpenMP
    pipeline and its LeSpGEMM_VLength on two cluster ranges of about cpu_share of the work
    each; the GPU computes the same ranges, and their checksums are compared.  C = A*A^T on
    the same clustering is timed the same way (GPU only) under "aat"."""
    import numpy as np
    import torch
    from sparseops_amd import lespgemm as LG
    from sparseops_amd.gen import rmat
    from sparseops_amd.io import write_csr_bin
    dev = torch.device("cuda:0")
    t = time.time()
    A = rmat(scale, 16, seed=0)
    gen_s = time.time() - t
    dA = LG.DeviceCsr(A, dev)
    torch.cuda.synchronize()
    t = time.perf_counter()
    pi, pj, ps = LG.lsh_candidate_pairs(dA, 64, 16, 7, constrain=True)
    genpairs_ms = (time.perf_counter() - t) * 1e3
    t = time.perf_counter()
    perm, offsets = LG.hierarchical_clustering(A, pi, pj, ps, 8, version=0)
    hc_ms = (time.perf_counter() - t) * 1e3
    t = time.perf_counter()
    Acl = LG.ClusterMatrix.build(dA, perm, offsets)
    torch.cuda.synchronize()
    conv_ms = (time.perf_counter() - t) * 1e3
    flop = LG.spgemm_flop(dA, dA)
    budget = int(budget_gb * (1 << 30))
    C = None
    passes = []
    for _ in range(2):
        torch.cuda.synchronize()
        t = time.perf_counter()
        plan = LG.SpgemmPlan(Acl, dA)
        ranges = plan.panel_ranges(budget)
        numeric_ms = 0.0
        for c0, c1 in ranges:
            C = plan.numeric(c0, c1, C)
            numeric_ms += C.stats["symbolic_ms"] + C.stats["malloc_ms"] + C.stats["numeric_ms"]
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t) * 1e3
        passes.append({"wall_ms": wall, "plan_ms": plan.setup_ms + plan.symbolic_ms, "panels_ms": numeric_ms})
        last = plan
        if len(passes) < 2:
            plan.close()
    plan = last
    w = passes[-1]
    # nonzeros of C (its CSR nnz, exact zeros dropped): an untimed pass over the panels
    nnz_c = 0
    for c0, c1 in ranges:
        C = plan.numeric(c0, c1, C)
        nnz_c += C.checksum()["nonzeros"]
    out = {"workload": f"R-MAT scale {scale} (n={A.n_rows}, nnz={A.nnz}), C = A*A, kernel 3, cluster_size 8, k=64, "
                       f"bands=16, seed=7, hc_v=0; output in cluster panels of <= {budget_gb} GiB",
           "flop": int(flop), "clusters": int(plan.clusters), "nnzc": int(plan.nnzc), "nnzv": int(plan.nnzv),
           "output_bytes": int(plan.nnzc * 4 + plan.nnzv * 8), "panels": len(ranges),
           "gpu_ms": round(w["wall_ms"], 2), "gpu_gflops": round(flop / w["wall_ms"] / 1e6, 2),
           "gpu_plan_ms": round(w["plan_ms"], 2), "gpu_panels_ms": round(w["panels_ms"], 2),
           "gpu_genpairs_ms": round(genpairs_ms, 1), "gpu_hc_ms": round(hc_ms, 1),
           "gpu_conversion_ms": round(conv_ms, 1), "generate_s": round(gen_s, 1), "pairs": int(len(pi)),
           "nnz_C": int(nnz_c),
           "roofline": spgemm_roofline(A.nnz, A.nnz, A.n_rows, A.n_rows, nnz_c, A.n_rows, w["wall_ms"])}
    # per-cluster flop in get_spgemm_flop's sense, and two ranges of about cpu_share of it
    cf = LG.cluster_flops(A, A.rowptr, perm, offsets)
    tot = int(cf.sum())
    picks = LG.sample_ranges(cf, cpu_share)
    gpu_pick = []
    for c0, c1 in picks:
        P = plan.numeric(c0, c1)
        ms = P.stats["symbolic_ms"] + P.stats["malloc_ms"] + P.stats["numeric_ms"]
        P2 = plan.numeric(c0, c1, P)
        ms = P2.stats["symbolic_ms"] + P2.stats["malloc_ms"] + P2.stats["numeric_ms"]
        gpu_pick.append({"c0": c0, "c1": c1, "flop": int(cf[c0:c1].sum()), "gpu_ms": round(ms, 3),
                         "nnzc": int(P2.nnzc), "nnzv": int(P2.nnzv), **P2.checksum()})
        P2.close()
    out["sampled_panels"] = gpu_pick
    fp_perm = int(np.sum(perm.astype(np.uint64) * (np.arange(len(perm), dtype=np.uint64) + np.uint64(1)),
                         dtype=np.uint64))
    fp_off = int(np.sum(offsets.astype(np.uint64) * (np.arange(len(offsets), dtype=np.uint64) + np.uint64(1)),
                        dtype=np.uint64))
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if cpu and os.path.exists(exe):
        with tempfile.TemporaryDirectory(dir="/tmp") as d:
            path = os.path.join(d, "A.csr")
            write_csr_bin(path, A)
            env = cpu_env(cpu_threads)
            rl = ",".join(f"{a}:{b}" for a, b in picks)
            proc = subprocess.Popen([exe, "bench_lespgemm_panels", path, path, str(cpu_threads), "1", "64", "16", "7",
                                     "8", "3", rl], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, env=env)
            lines = []
            for line in proc.stdout:
                lines.append(json.loads(line))
                log(f"cpu lespgemm panels: {line.strip()}")
            proc.wait(timeout=1200)
        if proc.returncode == 0 and len(lines) == 1 + len(picks):
            head, cp = lines[0], lines[1:]
            cpu_ms = sum(c["lespgemm_ms"] for c in cp)
            gpu_ms = sum(g["gpu_ms"] for g in gpu_pick)
            pflop = sum(g["flop"] for g in gpu_pick)
            match = all(c["nnzc"] == g["nnzc"] and c["nnzv"] == g["nnzv"] and int(c["colid_sum"]) == g["colid_sum"]
                        and int(c["nonzeros"]) == g["nonzeros"]
                        and abs(c["sum"] - g["sum"]) <= 1e-10 * max(c["sum_abs"], 1e-300) for c, g in zip(cp, gpu_pick))
            out["cpu_reference"] = {**HOST, "kind": "reference", "threads": head["threads"],
                                    "genpairs_ms": head["genpairs_ms"],
                                    "hc_ms": head["hc_ms"], "conversion_ms": head["conversion_ms"],
                                    "sampled_lespgemm_ms": round(cpu_ms, 2),
                                    "sampled_gflops": round(pflop / cpu_ms / 1e6, 3),
                                    "sample": f"LeSpGEMM_VLength kernel 3 on cluster ranges {rl} "
                                              f"({pflop:.3e} of {tot:.3e} flop)"}
            out["gpu_over_cpu_same_panels"] = round(cpu_ms / max(gpu_ms, 1e-9), 2)
            out["clustering_matches_reference"] = (head["clusters"] == plan.clusters
                                                   and head["perm_fingerprint"] == fp_perm
                                                   and head["offsets_fingerprint"] == fp_off)
            out["sampled_panels_match_reference"] = bool(match)
        else:
            log(f"cpu lespgemm panels failed: rc={proc.returncode} {proc.stderr.read()[-400:]}")
    if C is not None:
        C.close()
    plan.close()
    # C = A * A^T on the same clustering (the second product of the R-MAT configuration), GPU
    # only, timed like C = A * A (plan + panels, second of two passes)
    from sparseops_amd.gen import transpose
    At = transpose(A)
    dAt = LG.DeviceCsr(At, dev)
    flop_t = LG.spgemm_flop(dA, dAt)
    Ct, walls = None, []
    for _ in range(2):
        torch.cuda.synchronize()
        t = time.perf_counter()
        pt = LG.SpgemmPlan(Acl, dAt)
        ranges_t = pt.panel_ranges(budget)
        for c0, c1 in ranges_t:
            Ct = pt.numeric(c0, c1, Ct)
        torch.cuda.synchronize()
        walls.append((time.perf_counter() - t) * 1e3)
        nnzc_t = pt.nnzc
        if len(walls) == 2:        # untimed: C's nonzeros for the roofline
            nnz_ct = 0
            for c0, c1 in ranges_t:
                Ct = pt.numeric(c0, c1, Ct)
                nnz_ct += Ct.checksum()["nonzeros"]
        pt.close()
    Ct.close()
    out["aat"] = {"workload": "C = A * A^T, same matrix and clustering", "flop": int(flop_t),
                  "gpu_ms": round(walls[-1], 2), "gpu_gflops": round(flop_t / walls[-1] / 1e6, 2),
                  "nnzc": int(nnzc_t), "panels": len(ranges_t), "nnz_C": int(nnz_ct),
                  "roofline": spgemm_roofline(A.nnz, At.nnz, A.n_rows, At.n_rows, nnz_ct, A.n_rows, walls[-1])}
    del dAt
    # the approximate GPU clustering behind hc_v = 3 (no reference counterpart): MinHash + LSH +
    # rounds of mutual best pairs on the device, then the same product on its clusters (plan +
    # panels, the second of two passes)
    torch.cuda.synchronize()
    t = time.perf_counter()
    perm_g, off_g, rounds_g, npairs_g = LG.cluster_rows_gpu(dA, 64, 16, 7, True, 8)
    pre_g_ms = (time.perf_counter() - t) * 1e3
    Acl_g = LG.ClusterMatrix.build(dA, perm_g, off_g)
    Cg, walls_g = None, []
    for _ in range(2):
        torch.cuda.synchronize()
        t = time.perf_counter()
        pg = LG.SpgemmPlan(Acl_g, dA)
        for c0, c1 in pg.panel_ranges(budget):
            Cg = pg.numeric(c0, c1, Cg)
        torch.cuda.synchronize()
        walls_g.append((time.perf_counter() - t) * 1e3)
        nnzc_g = pg.nnzc
        pg.close()
    Cg.close()
    out["gpu_clustering"] = {
        "workload": "hc_v = 3: constrained LSH pairs clustered on the GPU by rounds of mutual best pairs "
                    "(cluster_size 8), then C = A*A on those clusters",
        "genpairs_and_clustering_ms": round(pre_g_ms, 1), "rounds": int(rounds_g), "pairs": int(npairs_g),
        "exact_genpairs_plus_hc_ms": round(genpairs_ms + hc_ms, 1),
        "clusters": len(off_g) - 1, "clusters_exact_hc": len(offsets) - 1,
        "nnzc_A": int(Acl_g.nnzc), "nnzc_A_exact_hc": int(Acl.nnzc), "nnzc_C": int(nnzc_g),
        "gpu_ms": round(walls_g[-1], 2), "gpu_gflops": round(flop / walls_g[-1] / 1e6, 2)}
    Acl_g.close()
    Acl.close()
    torch.cuda.synchronize()
    return out


def autograd_measure(scale, k, steps, cpu, cpu_threads):
    """The differentiable SpMM layer Y = A(vals) X: R-MAT(scale, edge factor 32) with symmetric
    normalised FP32 values, X of width k in BF16, FP32 accumulation, Y rounded once to BF16 (so
    its gradient G is BF16, as the workload states).  Forward and backward
    (dX = A^T G and dvals = SDDMM(G, X)) timed with HIP events; the CPU leg is the oracle's
    single-thread CSR SpMM on the first 1/64 of the rows, scaled to the whole matrix."""
    import ctypes
    import numpy as np
    import torch
    from sparseops_amd import _lib as L
    from sparseops_amd import autograd as AG
    from sparseops_amd.gen import rmat_symnorm
    dev = torch.device("cuda:0")
    A = rmat_symnorm(scale, 32, seed=0)
    pat = AG.SparsePattern(A, "f32")
    rng = np.random.default_rng(0)
    vals = torch.tensor(A.val, device=dev, requires_grad=True)
    X = torch.tensor(rng.standard_normal((A.n_cols, k), dtype=np.float32), device=dev).to(torch.bfloat16)
    X.requires_grad_()
    G = torch.tensor(rng.standard_normal((A.n_rows, k), dtype=np.float32), device=dev).to(torch.bfloat16)
    stream = torch.cuda.current_stream().cuda_stream
    ev = [ctypes.c_void_p() for _ in range(3)]
    for e in ev:
        L.call("spops_event_create", ctypes.byref(e))
    fw = bw = 0.0
    ms = ctypes.c_float()
    for it in range(steps + 2):
        vals.grad = None
        X.grad = None
        torch.cuda.synchronize()
        L.call("spops_event_record", ev[0], stream)
        Y = AG.spmm(pat, vals, X, out_dtype=torch.bfloat16)
        L.call("spops_event_record", ev[1], stream)
        Y.backward(G)
        L.call("spops_event_record", ev[2], stream)
        torch.cuda.synchronize()
        if it >= 2:
            L.call("spops_event_elapsed_ms", ctypes.byref(ms), ev[0], ev[1])
            fw += ms.value
            L.call("spops_event_elapsed_ms", ctypes.byref(ms), ev[1], ev[2])
            bw += ms.value
    for e in ev:
        L.call("spops_event_destroy", e)
    fw, bw = fw / steps, bw / steps
    n, nnz = A.n_rows, A.nnz
    # forward: entries (column + FP32 value), row pointers, X read, Y written; backward: dX = A^T G
    # (entries, row pointers, G read, dX written) and dvals (columns, row pointers, G, X, dvals)
    b_fwd = nnz * 8 + (n + 1) * 4 + n * k * 2 + n * k * 2
    b_bwd = (nnz * 8 + (n + 1) * 4 + n * k * 2 + n * k * 2) + (nnz * 4 + (n + 1) * 4 + n * k * 2 + n * k * 2 + nnz * 4)
    out = {"workload": f"Y = A X, R-MAT scale {scale} edge factor 32 (n={n}, nnz={nnz}), symmetric-normalised FP32 "
                       f"values, X/Y/G BF16 width {k}, FP32 accumulation; backward dX = A^T G, dvals = SDDMM(G, X)",
           "forward_ms": round(fw, 3), "backward_ms": round(bw, 3),
           "forward_gbs": round(b_fwd / fw / 1e6, 1), "backward_gbs": round(b_bwd / bw / 1e6, 1),
           "forward_gflops": round(2 * nnz * k / fw / 1e6, 1),
           "algo_bytes_forward": b_fwd, "algo_bytes_backward": b_bwd}
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if cpu and os.path.exists(exe):
        from sparseops_amd.io import write_csr_bin
        res = {}
        with tempfile.TemporaryDirectory(dir="/tmp") as d:
            path = os.path.join(d, "A.csr")
            write_csr_bin(path, A)
            for name, tr in (("forward", 0), ("backward_dx", 1)):
                r = subprocess.run([exe, "bench_spmm", path, str(k), str(cpu_threads), "2", "32", str(tr)],
                                   capture_output=True, text=True, env=cpu_env(cpu_threads), timeout=900)
                if r.returncode == 0:
                    res[name] = json.loads(r.stdout.strip().splitlines()[-1])
                else:
                    log(f"cpu autograd {name} failed: {r.stderr[-300:]}")
        if len(res) == 2:
            # dvals = SDDMM(G, X) has no reference counterpart: the oracle's one-thread loop on
            # the first 1/64 of the rows, scaled to the whole matrix
            sys.path.insert(0, os.path.join(ROOT, "tests"))
            import oracle_bind as orc
            from sparseops_amd.gen import Csr
            r1 = max(1, n // 64)
            sub = Csr(r1, A.n_cols, A.rowptr[:r1 + 1].astype(np.int64), A.col[:A.rowptr[r1]].astype(np.int64),
                      A.val[:A.rowptr[r1]].astype(np.float64))
            Xh = X.detach().double().cpu().numpy()
            Gh = G.double().cpu().numpy()
            t = time.perf_counter()
            orc.sddmm(sub, Gh[:r1], Xh)
            sddmm_ms = (time.perf_counter() - t) * n / r1 * 1e3
            fwd_ms = res["forward"]["seconds_per_spmm"] * 1e3
            dx_ms = res["backward_dx"]["seconds_per_spmm"] * 1e3
            out["cpu_baseline"] = {"value": round(fwd_ms, 2), "unit": "ms (forward)", "kind": "reference",
                                   "cores": cpu_threads, "forward_ms": round(fwd_ms, 2),
                                   "backward_dx_ms": round(dx_ms, 2),
                                   "backward_dvals_ms_port_1thread": round(sddmm_ms, 1),
                                   "sample": f"the whole matrix: forward and dX = A^T G as {k} calls of the "
                                             f"reference's OpenMP SpOps_csr (kernel_flag 2, FP32) per product, "
                                             f"on A and on csr_transpose(A); dvals (no reference counterpart) "
                                             f"by the oracle's one-thread SDDMM on rows 0..{r1 - 1}, scaled",
                                   **HOST}
            out["gpu_over_cpu_forward"] = round(fwd_ms / fw, 1)
            out["gpu_over_cpu_backward_dx"] = round(dx_ms / bw, 1)
    pat.close()
    return out


def bsr_measure(n, b, blocks_per_row, steps, warmup):
    """BSR SpMV on dense b x b blocks: the matrix-core kernel (v_mfma_f64_16x16x4f64) and the
    VALU kernel on the same matrix, FP64 and FP32 values (algorithmic bytes: block row
    pointers, block columns, block values, x, y)."""
    import numpy as np
    import torch
    from sparseops_amd import spmv as S
    from sparseops_amd.gen import block_sparse, uniform_x
    A = block_sparse(n, b, blocks_per_row, seed=1)
    x64 = uniform_x(A.n_cols, 0)
    stream = torch.cuda.current_stream().cuda_stream
    out = {"workload": f"BSR {b}x{b} dense blocks, {blocks_per_row} per block row, n={n} (nnz={A.nnz}); "
                       f"and {2 * b}x{2 * b} blocks, {blocks_per_row // 4} per block row (MFMA kernel only)"}

    def run(M, key):
        x = torch.tensor(x64, dtype=M.xy_dtype, device="cuda:0")
        y = torch.zeros(A.n_rows, dtype=M.xy_dtype, device="cuda:0")
        _, kern = time_spmv(M, x, y, steps, warmup, stream)
        out[key + "_us"] = round(kern * 1e3, 2)
        out[key + "_gbs"] = round(M.algo_bytes / (kern * 1e-3) / 1e9, 1)
        M.close()

    for vt in ("f64", "f32"):
        for mode in ("always", "never"):
            with S.bsr_mfma(mode):
                M = S.csr_to_bsr(A, b, b, vt)
            assert M.info[9] == (1 if mode == "always" else 0)
            run(M, f"{vt}_{'mfma' if mode == 'always' else 'valu'}")
    A = block_sparse(n, 2 * b, max(1, blocks_per_row // 4), seed=2)
    x64 = uniform_x(A.n_cols, 0)
    for vt in ("f64", "f32"):
        M = S.csr_to_bsr(A, 2 * b, 2 * b, vt)
        assert M.info[9] == 1
        run(M, f"{vt}_{2 * b}x{2 * b}_mfma")
    return out


def batch_measure(count, steps, warmup, cpu, cpu_threads):
    """Batched SpMM: `count` band matrices (n log-uniform in [1k, 64k], half-width U{1..64},
    0-8 off-band entries per row, k U{1..32}, FP32 N(0,1) values and X), formats chosen by the
    autotuner, one persistent launch per step replayed from a captured HIP graph and timed with
    HIP events on the replay stream.  The CPU leg is the oracle's single-thread CSR SpMM over
    whole matrices of the batch, in order, for about cpu_seconds, scaled by flop."""
    import ctypes
    import numpy as np
    import torch
    from sparseops_amd import _lib as L
    from sparseops_amd.batch import FORMATS as FORMAT_NAMES, BatchSpMM
    from sparseops_amd.gen import Csr, band_batch
    t = time.time()
    mats, ks = band_batch(count, seed=0)
    gen_s = time.time() - t
    t = time.time()
    B = BatchSpMM(mats, ks, tune=True)
    create_s = time.time() - t
    X, Y = B.new_x(), B.new_y()
    X.copy_(torch.randn(X.numel(), generator=torch.Generator().manual_seed(0)).to(X.device))
    s = torch.cuda.Stream()

    def replay_ms(Bx, Yx):
        """Mean ms of one graph replay of Bx's launch (HIP events on the replay stream)."""
        torch.cuda.synchronize()
        Bx.capture(X, Yx, s.cuda_stream)
        ev = [ctypes.c_void_p() for _ in range(2)]
        for e in ev:
            L.call("spops_event_create", ctypes.byref(e))
        for _ in range(warmup):
            Bx.replay(s.cuda_stream)
        s.synchronize()
        L.call("spops_event_record", ev[0], s.cuda_stream)
        for _ in range(steps):
            Bx.replay(s.cuda_stream)
        L.call("spops_event_record", ev[1], s.cuda_stream)
        s.synchronize()
        ms = ctypes.c_float()
        L.call("spops_event_elapsed_ms", ctypes.byref(ms), ev[0], ev[1])
        for e in ev:
            L.call("spops_event_destroy", e)
        return ms.value / steps

    step_ms = replay_ms(B, Y)
    # the autotuner's choice per class, and the same batch on the untuned default rule
    classes = {}
    for i in range(count):
        c = classes.setdefault(str(int(B.cls[i])), {"matrices": 0, "plans": {}})
        c["matrices"] += 1
        key = f"{FORMAT_NAMES.get(int(B.fmt[i]), B.fmt[i])}/lanes={int(B.lanes[i])}/sigma={int(B.sigma[i])}"
        c["plans"][key] = c["plans"].get(key, 0) + 1
    t = time.time()
    B0 = BatchSpMM(mats, ks, tune=False)
    Y0 = B0.new_y()
    untuned_ms = replay_ms(B0, Y0)
    untuned = {"ms_per_step": round(untuned_ms, 4), "gbs": round(B0.algo_bytes / untuned_ms / 1e6, 1),
               "create_s": round(time.time() - t, 1), "tuned_speedup": round(untuned_ms / step_ms, 3)}
    fmts0 = B0.formats()
    untuned["formats"] = {f: fmts0.count(f) for f in set(fmts0)}
    B0.close()
    del Y0
    fmts = B.formats()
    out = {"workload": f"{count} band matrices (n log-uniform 1k-64k, half-width 1-64, 0-8 off-band entries per row), "
                       f"k U{{1..32}}, FP32 values / X / accumulation, one persistent launch replayed from a HIP graph",
           "matrices": count, "rows": int(sum(A.n_rows for A in mats)), "nnz": int(sum(A.nnz for A in mats)),
           "flop": B.flop, "algo_bytes": B.algo_bytes, "ms_per_step": round(step_ms, 4),
           "gbs": round(B.algo_bytes / step_ms / 1e6, 1),
           "pct_hbm_peak": round(100 * B.algo_bytes / step_ms / 1e6 / HBM_PEAK_GBS, 2),
           "gflops": round(B.flop / step_ms / 1e6, 1), "tasks": B.ntasks,
           "formats": {f: fmts.count(f) for f in set(fmts)}, "autotuner_classes": classes,
           "untuned_default_rule": untuned, "generate_s": round(gen_s, 1),
           "create_and_tune_s": round(create_s, 1), "kernel": "k_batch_spmm"}
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as orc
    Xh, Yh = X.double().cpu(), Y.double().cpu()
    worst = 0.0
    for i in range(0, count, max(1, count // 16)):
        A = mats[i]
        Xi = B.x_view(Xh, i).numpy()
        A64 = Csr(A.n_rows, A.n_cols, A.rowptr, A.col, A.val.astype(np.float64))
        want = orc.spmm(A64, Xi)
        scale = orc.spmm(Csr(A.n_rows, A.n_cols, A.rowptr, A.col, np.abs(A64.val)), np.abs(Xi))
        got = B.y_view(Yh, i).numpy()
        worst = max(worst, float(np.max(np.abs(got - want) / (scale + 1e-30))))
    out["max_scaled_err_sampled"] = worst
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if cpu and os.path.exists(exe):
        # the reference's OpenMP SpOps_csr (kernel_flag 2, FP32) once per right-hand side, on the
        # first matrices of the batch up to share_flop of its flop, scaled by flop
        from sparseops_amd.io import write_csr_bin
        share_flop, done = 0.05, 0
        with tempfile.TemporaryDirectory(dir="/tmp") as d:
            lines = []
            for i in range(count):
                pth = os.path.join(d, f"m{i}.csr")
                write_csr_bin(pth, mats[i])
                lines.append(f"{pth} {ks[i]}\n")
                done += 2 * mats[i].nnz * ks[i]
                if done >= share_flop * B.flop:
                    break
            lst = os.path.join(d, "list.txt")
            with open(lst, "w") as f:
                f.writelines(lines)
            r = subprocess.run([exe, "bench_spmm_batch", lst, str(cpu_threads), "32"], capture_output=True, text=True,
                               env=cpu_env(cpu_threads), timeout=900)
        if r.returncode == 0:
            c = json.loads(r.stdout.strip().splitlines()[-1])
            cpu_ms = B.flop / c["flop"] * c["seconds"] * 1e3
            out["cpu_baseline"] = {"value": round(cpu_ms, 1), "unit": "ms per batch", "kind": "reference",
                                   "cores": c["threads"], "gflops": round(c["flop"] / c["seconds"] / 1e9, 3),
                                   "sample": f"matrices 0..{c['matrices'] - 1} of the batch ({c['flop']:.3e} of "
                                             f"{B.flop:.3e} flop): each k-column product as k calls of the "
                                             f"reference's OpenMP SpOps_csr (kernel_flag 2, FP32) after a warm-up, "
                                             f"scaled by flop to the whole batch", **HOST}
            out["gpu_over_cpu"] = round(cpu_ms / step_ms, 1)
        else:
            log(f"cpu batch baseline failed: {r.stderr[-300:]}")
    B.close()
    return out


def laplacian_measure(nx, steps, warmup, cpu_threads, cpu, cpu_seconds):
    """Plain FP64 CSR SpMV on the 5-point Laplacian of an nx-by-nx grid (x ~ U[-1,1), seed 0):
    the device-resident kernel (HIP events, many steps: one SpMV moves 5.2 MB, so it is
    launch-bound), the drop-in driver benchmark_spmv_csr on the same matrix as a .mtx (its
    host-array timing includes the PCIe copies of x and y), and the reference's OpenMP CSR
    kernel (kernel_flag 2) on this host's cores.  Error against an FP64 scipy product."""
    import numpy as np
    import scipy.sparse as sp
    import torch
    from sparseops_amd import spmv as S
    from sparseops_amd.gen import laplacian_2d, uniform_x
    from sparseops_amd.io import write_csr_bin, write_mtx

    A = laplacian_2d(nx)
    x64 = uniform_x(A.n_cols, 0)
    As = sp.csr_matrix((A.val, A.col, A.rowptr), shape=(A.n_rows, A.n_cols))
    y64 = As @ x64
    absrow = abs(As) @ np.abs(x64)
    M = S.csr(A, "f64")
    dev = torch.device("cuda:0")
    x = torch.tensor(x64, device=dev)
    y = torch.zeros(A.n_rows, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    wall_ms, kern_ms = time_spmv(M, x, y, steps, warmup, stream)
    err = errors(y.cpu().numpy(), y64, absrow)
    # back-to-back SpMVs without the per-call launch from Python: 100 calls captured in one HIP
    # graph, the graph replayed and timed with events on its stream
    per_graph = 100
    g = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream()
    cap.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(cap):
        M.spmv(x, y, 1.0, 0.0, cap.cuda_stream)
        torch.cuda.synchronize()
        with torch.cuda.graph(g, stream=cap):
            for _ in range(per_graph):
                M.spmv(x, y, 1.0, 0.0, cap.cuda_stream)
    torch.cuda.synchronize()
    y.zero_()
    g.replay()
    torch.cuda.synchronize()
    graph_err = errors(y.cpu().numpy(), y64, absrow)["max_rel_err"]
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    with torch.cuda.stream(cap):
        e0.record(cap)
        for _ in range(reps):
            g.replay()
        e1.record(cap)
    torch.cuda.synchronize()
    graph_us = e0.elapsed_time(e1) * 1e3 / (reps * per_graph)
    del g
    out = {"n": A.n_rows, "nnz": A.nnz, "algo_bytes": M.algo_bytes, "kernel_us": round(kern_ms * 1e3, 3),
           "wall_us_per_step": round(wall_ms * 1e3, 3),
           "graph_us_per_spmv": round(graph_us, 3), "graph_gbs": round(M.algo_bytes / (graph_us * 1e-6) / 1e9, 1),
           "graph_max_rel_err": graph_err,
           "kernel_gbs": round(M.algo_bytes / (kern_ms * 1e-3) / 1e9, 1),
           "pct_hbm_peak": round(100.0 * M.algo_bytes / (kern_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 2),
           "gflops": round(2.0 * A.nnz / (kern_ms * 1e-3) / 1e9, 2), "max_rel_err": err["max_rel_err"],
           "dropin": None, "cpu_reference": None}
    M.close()
    exe = os.path.join(ROOT, "sparseops_amd", "cli", "benchmark_spmv_csr")
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        if os.path.exists(exe):
            path = os.path.join(d, f"laplacian_{nx}.mtx")
            write_mtx(path, A)
            r = subprocess.run([exe, path, "--precision=64", f"--perf_file={os.path.join(d, 'perf.txt')}"],
                               capture_output=True, text=True, timeout=300)
            dropin = {"returncode": r.returncode, "passed": "PASSED" in r.stdout}
            for line in r.stdout.splitlines():
                if "benchmarking" in line and "[" in line:
                    tag = line.split("[")[1].split("]")[0]
                    ms = float(line.split("]:")[1].split("ms")[0])
                    dropin["host_arrays_us" if tag == "host arrays" else "device_us"] = round(ms * 1e3, 3)
            out["dropin"] = dropin
        if cpu and os.path.exists(harness):
            path = os.path.join(d, "A.csr")
            write_csr_bin(path, A)
            env = cpu_env(cpu_threads)
            r = subprocess.run([harness, "bench_spmv", path, "csr", "64", str(cpu_threads), str(cpu_seconds)],
                               capture_output=True, text=True, env=env, timeout=300)
            if r.returncode == 0:
                c = json.loads(r.stdout.strip().splitlines()[-1])
                out["cpu_reference"] = {**HOST, "us": round(c["seconds_per_iter"] * 1e6, 3), "cores": c["threads"],
                                        "iterations": c["iterations"],
                                        "gbs": round(out["algo_bytes"] / c["seconds_per_iter"] / 1e9, 1)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1 << 21)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="OpenMP threads of the CPU baselines (0: one per physical core this process may use, "
                         "capped by OMP_NUM_THREADS)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--variants", default="sell_f32,csr_f64,sell_f64,sell_bf16,csr_f32,csrrows_f64",
                    help="comma list of <fmt>_<value type> (fmt: sell, csr, or csrrows = CSR held to its row "
                         "kernels); the first one is the headline")
    ap.add_argument("--lespgemm-scale", type=int, default=16, help="R-MAT scale of the LeSpGEMM line (0: skip)")
    ap.add_argument("--lespgemm-panels-scale", type=int, default=20,
                    help="R-MAT scale of the panelled LeSpGEMM line (0: skip)")
    ap.add_argument("--lespgemm-budget-gb", type=float, default=32.0, help="output bytes per LeSpGEMM panel")
    ap.add_argument("--lespgemm-iters", type=int, default=5)
    ap.add_argument("--rowwise-scale", type=int, default=16,
                    help="R-MAT scale of the row-wise LeSpGEMM / A*A^T top-k line (0: skip)")
    ap.add_argument("--autograd-scale", type=int, default=20, help="R-MAT scale of the autograd SpMM line (0: skip)")
    ap.add_argument("--batch-count", type=int, default=512, help="matrices of the batched SpMM line (0: skip)")
    ap.add_argument("--bsr-n", type=int, default=1 << 18, help="rows of the dense-block BSR line (0: skip)")
    ap.add_argument("--laplacian-nx", type=int, default=256,
                    help="grid side of the FP64 CSR Laplacian line (0: skip)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="only run a device copy of this many MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    HOST.update(host_cpu())
    if args.cpu_threads <= 0:
        args.cpu_threads = HOST["threads"]
    HOST["threads"] = args.cpu_threads
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the sharded multi-GPU form is a design, not run here.", file=sys.stderr)
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
        if fmt == "csrrows":                   # CSR with its row kernels only (no swept layout)
            with S.spmv_layout("sliced"):
                M = S.csr(A, vt)
        else:
            M = S.csr(A, vt) if fmt == "csr" else S.csr_to_sell_c_sigma(A, 1024, 32, vt)
        dt = M.xy_dtype
        x = torch.tensor(x64, dtype=dt, device=dev)
        y = torch.zeros(A.n_rows, dtype=dt, device=dev)
        wall_ms, kern_ms = time_spmv(M, x, y, args.steps, args.warmup, stream)
        yh = y.double().cpu().numpy()
        r = {"wall_ms": wall_ms, "kernel_ms": kern_ms, "algo_bytes": M.algo_bytes,
             "gbs": M.algo_bytes / (wall_ms * 1e-3) / 1e9, "kernel_gbs": M.algo_bytes / (kern_ms * 1e-3) / 1e9,
             "device_bytes": M.device_bytes, "build_s": time.time() - t, "sweep": M.sweep,
             "kernels": ((["k_sweep_spmv", "k_sweep_reduce4" if vt in ("f32", "bf16") and A.n_rows % 4 == 0
                                           else "k_sweep_reduce"]
                          if M.sweep[1] > 1 else ["k_sweep_spmv"]) if M.sweep
                         else ["k_csr_spmv"] if M.format not in ("sell", "ell")
                         else (["k_slice_spmv", "k_slice_reduce"] if M.slices > 1 else ["k_slice_spmv"]))}
        r.update(errors(yh, y64, absrow))
        rb = ref_bytes_per_spmv(fmt, vt, A, M)
        r["ref_bytes_per_spmv"] = rb
        r["ref_model_gbs"] = round(rb / (kern_ms * 1e-3) / 1e9, 1) if rb else None
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
                             f"{c['seconds_per_iter'] * 1e3:.2f} ms each", **HOST}
        log(f"cpu baseline done in {time.time() - t:.1f}s: {cpu}")

    traffic = committed_traffic() if head_name == "sell_f32" and args.n == 1 << 21 else None
    lg = None
    if args.lespgemm_scale > 0:
        t = time.time()
        lg = lespgemm_measure(args.lespgemm_scale, args.lespgemm_iters, args.cpu_threads, not args.no_cpu)
        log(f"lespgemm done in {time.time() - t:.1f}s: {lg}")

    rt = None
    if args.rowwise_scale > 0:
        t = time.time()
        rt = rowwise_topk_measure(args.rowwise_scale, 3, args.cpu_threads, not args.no_cpu)
        log(f"rowwise / topk done in {time.time() - t:.1f}s: {rt}")

    lp = None
    if args.lespgemm_panels_scale > 0:
        t = time.time()
        lp = lespgemm_panels_measure(args.lespgemm_panels_scale, args.lespgemm_budget_gb, args.cpu_threads,
                                     not args.no_cpu)
        log(f"lespgemm panels done in {time.time() - t:.1f}s: {lp}")

    ag = None
    if args.autograd_scale > 0:
        t = time.time()
        ag = autograd_measure(args.autograd_scale, 64, 5, not args.no_cpu, args.cpu_threads)
        log(f"autograd spmm done in {time.time() - t:.1f}s: {ag}")

    bt = None
    if args.batch_count > 0:
        t = time.time()
        bt = batch_measure(args.batch_count, max(5, args.steps // 10), max(2, args.warmup // 5), not args.no_cpu,
                           args.cpu_threads)
        log(f"batched spmm done in {time.time() - t:.1f}s: {bt}")

    bs = None
    if args.bsr_n > 0:
        t = time.time()
        bs = bsr_measure(args.bsr_n, 16, 8, max(10, args.steps // 4), max(3, args.warmup // 4))
        log(f"bsr done in {time.time() - t:.1f}s: {bs}")

    lap = None
    if args.laplacian_nx > 0:
        t = time.time()
        lap = laplacian_measure(args.laplacian_nx, max(200, args.steps * 5), max(20, args.warmup), args.cpu_threads,
                                not args.no_cpu, min(2.0, args.cpu_seconds))
        log(f"laplacian done in {time.time() - t:.1f}s: {lap}")

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
        "config": {"workload": "SELL-C-sigma SpMV, FP32 values and x/y (products and row sums in FP64), "
                               f"power-law n=2,097,152 (nnz={A.nnz}, mean row length {A.nnz / A.n_rows:.2f}, "
                               "L^-2.1 on [3,8192]), C=32, sigma=1024",
                   "n": A.n_rows, "nnz": A.nnz, "C": 32, "sigma": 1024},
        "pct_hbm_peak": round(100.0 * head["gbs"] / HBM_PEAK_GBS, 2),
        "roofline": {"bound": "hbm", "achieved": round(head["kernel_gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(head["kernel_gbs"] / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algo_bytes": head["algo_bytes"], "kernel_us": round(head["kernel_ms"] * 1e3, 2),
                     "kernels": head.get("kernels")},
        "cpu_baseline": cpu,
        "rel_err_vs_fp64": {k: {e: v[e] for e in ("max_rel_err", "rms_rel_err", "normwise_rel_err")}
                            for k, v in res.items()},
        "variants": {k: {"gbs": round(v["gbs"], 1), "kernel_us": round(v["kernel_ms"] * 1e3, 2),
                         "pct_hbm_peak": round(100 * v["kernel_gbs"] / HBM_PEAK_GBS, 2), "kernels": v["kernels"],
                         "sweep": v["sweep"], "algo_bytes": v["algo_bytes"],
                         "ref_bytes_per_spmv": v["ref_bytes_per_spmv"], "ref_model_gbs": v["ref_model_gbs"]}
                     for k, v in res.items()},
        "reference_bytes_model": {"bytes_per_spmv": head["ref_bytes_per_spmv"], "gbs": head["ref_model_gbs"],
                                  "note": "the reference's bytes_per_spmv (x counted once per entry, padded SELL "
                                          "slots included); value and roofline use the algorithmic bytes"},
        "fp32_sell_over_fp64_csr": (round(res["csr_f64"]["kernel_ms"] / res["sell_f32"]["kernel_ms"], 3)
                                    if "csr_f64" in res and "sell_f32" in res else None),
        "fp32_sell_over_fp64_sell": (round(res["sell_f64"]["kernel_ms"] / res["sell_f32"]["kernel_ms"], 3)
                                     if "sell_f64" in res and "sell_f32" in res else None),
        "fp32_sell_over_fp64_csr_row_kernels": (round(res["csrrows_f64"]["kernel_ms"] / res["sell_f32"]["kernel_ms"], 3)
                                                if "csrrows_f64" in res and "sell_f32" in res else None),
        "lespgemm": lg,
        "lespgemm_panels": lp,
        "rowwise_topk": rt,
        "autograd_spmm": ag,
        "batch_spmm": bt,
        "bsr_dense_blocks": bs,
        "laplacian_csr_f64": lap,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
