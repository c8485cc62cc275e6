#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ from the reference itself.

The reference (lsl036/SparseOps) ships no golden vectors, so its own functions are run
here, built from its sources by oracle/Makefile into oracle/_ref/ref_harness, on seeded
synthetic inputs (sparseops_amd.gen).  Only inputs' fingerprints and outputs are stored
(.npz, no pickles); the inputs are regenerated from their seeds by the tests.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from sparseops_amd.features import NAMES as FEATURE_NAMES  # noqa: E402
from sparseops_amd.gen import laplacian_2d, powerlaw, random_rect, rmat, transpose, uniform_x  # noqa: E402
from sparseops_amd.io import write_csr_bin, write_mtx  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
OUT = os.path.dirname(os.path.abspath(__file__))


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def rd(d, name, dt):
    p = os.path.join(d, name)
    return np.fromfile(p, dtype=dt) if os.path.exists(p) else None


def run(args, d):
    subprocess.check_call([HARNESS, *args], stdout=subprocess.DEVNULL, env=dict(os.environ, OMP_NUM_THREADS="4"))


def spmv_case(name, A, sigma, C, bR, bC, prec, full_layout):
    with tempfile.TemporaryDirectory() as d:
        write_csr_bin(os.path.join(d, "A.csr"), A)
        x = uniform_x(A.n_cols, 1)
        x.tofile(os.path.join(d, "x.f64"))
        run(["spmv", os.path.join(d, "A.csr"), os.path.join(d, "x.f64"), d, str(sigma), str(C), str(bR), str(bC),
             str(prec)], d)
        vt = np.float64 if prec == 64 else np.float32
        out = {"input_digest": digest(A.rowptr, A.col, A.val), "sigma": sigma, "C": C, "bR": bR, "bC": bC,
               "x_seed": 1, "alpha": 0.8, "beta": 0.7}
        for y in ["y0", "y_csr_a1b0", "y_csr_ab", "y_csr_lb_ab", "y_sell_ab", "y_sellr_ab", "y_ell_ab", "y_bsr_ab"]:
            v = rd(d, y, vt)
            if v is not None:
                out[y] = v
        out["sell_reorder"] = rd(d, "sell_reorder", np.int32)
        out["sell_chunk_len"] = rd(d, "sell_chunk_len", np.int32)
        out["sellr_reorder"] = rd(d, "sellr_reorder", np.int32)
        out["sellr_chunk_len"] = rd(d, "sellr_chunk_len", np.int32)
        out["sell_meta"] = rd(d, "sell_meta", np.int64)
        out["ell_width"] = rd(d, "ell_width", np.int64)
        sc, sv = rd(d, "sell_col", np.int32), rd(d, "sell_val", vt)
        bp, bc, bd = rd(d, "bsr_row_ptr", np.int32), rd(d, "bsr_colidx", np.int32), rd(d, "bsr_data", vt)
        if full_layout:
            out.update(sell_col=sc, sell_val=sv, bsr_row_ptr=bp, bsr_colidx=bc, bsr_data=bd)
        out["sell_layout_digest"] = digest(sc, sv)
        if bp is not None:
            out["bsr_layout_digest"] = digest(bp, bc, bd)
    np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **out)
    print("wrote", name)


def lespgemm_case(name, A, B, k=64, bands=16, seed=7, hc_v=0, cs=8, prec=64, full_pairs=True, full_c=False):
    with tempfile.TemporaryDirectory() as d:
        write_csr_bin(os.path.join(d, "A.csr"), A)
        write_csr_bin(os.path.join(d, "B.csr"), B)
        run(["lespgemm", os.path.join(d, "A.csr"), os.path.join(d, "B.csr"), d, str(k), str(bands), str(seed),
             str(hc_v), str(cs), "3", str(prec)], d)
        sig = rd(d, "signatures", np.uint64).reshape(A.n_rows, k)
        pi, pj, ps = rd(d, "pairs_i", np.int64), rd(d, "pairs_j", np.int64), rd(d, "pairs_score", np.float64)
        perm, off = rd(d, "perm", np.int64), rd(d, "offsets", np.int64)
        crp, ccol, cval = rd(d, "C_rowptr", np.int64), rd(d, "C_col", np.int64), rd(d, "C_val", np.float64)
        # the reference emits C's columns in cluster column order; sort within rows
        rows = np.repeat(np.arange(A.n_rows), np.diff(crp))
        o = np.lexsort((ccol, rows))
        ccol, cval = ccol[o], cval[o]
        out = {"input_digest": digest(A.rowptr, A.col, A.val), "B_digest": digest(B.rowptr, B.col, B.val),
               "k": k, "bands": bands, "seed": seed, "hc_v": hc_v, "cluster_size": cs, "precision": prec,
               "sig_head": sig[:64].copy(), "sig_digest": digest(sig), "npairs": len(pi),
               "pairs_digest": digest(pi, pj, ps), "perm": perm, "offsets": off,
               "acl_meta": rd(d, "acl_meta", np.int64), "ccl_meta": rd(d, "ccl_meta", np.int64),
               "ccl_rowptr": rd(d, "ccl_rowptr", np.int64), "ccl_min": rd(d, "ccl_min", np.int64),
               "ccl_max": rd(d, "ccl_max", np.int64), "flops": rd(d, "flops", np.int64),
               "C_rowptr": crp, "C_col_digest": digest(ccol),
               "C_rowsum": np.bincount(rows[o], weights=cval, minlength=A.n_rows)}
        if full_pairs:
            out.update(pairs_i=pi, pairs_j=pj, pairs_score=ps)
        if full_c:
            out.update(C_col=ccol, C_val=cval)
    np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **out)
    print("wrote", name, "pairs", out["npairs"], "clusters", len(out["offsets"]) - 1, "nnz(C)", out["C_rowptr"][-1])


# labels MTX::FeaturesPrint / ExtraFeaturesPrint use (include/sparse_features.h:44-150) for
# the library's feature names, and the factor they are printed with (percentages x100)
PERCENT = {"nnz_ratio", "nnz_closed2diag_ratio", "NoneZero_Row_Ratio", "NoneZero_Col_Ratio",
           "diagonal_dominant_ratio"}
PRINTED = {"real_nnz": "real_nnz", "uniq_R": "uniq_R", "uniq_C": "uniq_C"}


def features_case(name, A):
    """The reference's feature extractor on A written as a general Matrix Market file; the
    printed values (6 significant digits) as a vector in the library's order (NaN: not
    printed)."""
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "A.mtx")
        write_mtx(path, A)
        text = subprocess.check_output([HARNESS, "features", path], text=True)
    got = {}
    for line in text.splitlines():
        for sep in (":", "="):
            if sep in line:
                k, v = line.split(sep, 1)
                k, v = k.strip(), v.strip()
                if k == "Tile inner size":
                    a, b = v.split("*")
                    got["tile_rows"], got["tile_cols"] = float(a), float(b)
                else:
                    try:
                        got[k] = float(v.rstrip("%"))
                    except ValueError:
                        pass
                break
    vec = np.full(len(FEATURE_NAMES), np.nan)
    for i, n in enumerate(FEATURE_NAMES):
        if n in got:
            vec[i] = got[n] / 100.0 if n in PERCENT else got[n]
    np.savez(os.path.join(OUT, f"{name}.npz"), features=vec, input_digest=digest(A.rowptr, A.col, A.val))


def classify_case(name, A, fractions, cs=8):
    """Kernel 3's cluster classification (src/spgemm_Vlength_mixed.cpp:268-319) as the
    reference computes it, for several L2 fractions of its budget (-1: its default 0.75)."""
    out = {"input_digest": digest(A.rowptr, A.col, A.val), "fractions": np.array(fractions, np.float64),
           "cluster_size": cs}
    with tempfile.TemporaryDirectory() as d:
        write_csr_bin(os.path.join(d, "A.csr"), A)
        for t, f in enumerate(fractions):
            run(["lespgemm", os.path.join(d, "A.csr"), os.path.join(d, "A.csr"), d, "64", "16", "7", "0", str(cs),
                 "3", "64", repr(f)], d)
            out[f"acc_flag_{t}"] = rd(d, "ccl_acc_flag", np.int8)
            out[f"dense_count_{t}"] = rd(d, "dense_cluster_count", np.int64)
            if t == 0:
                out["perm"], out["offsets"] = rd(d, "perm", np.int64), rd(d, "offsets", np.int64)
                out["ccl_min"], out["ccl_max"] = rd(d, "ccl_min", np.int64), rd(d, "ccl_max", np.int64)
                out["ccl_rowptr"] = rd(d, "ccl_rowptr", np.int64)
    np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **out)
    print("wrote", name, "dense counts", [int(out[f"dense_count_{t}"][0]) for t in range(len(fractions))])


def fingerprint_case(name, A, block=1):
    """C = A * A at a scale whose C is too large to keep: the reference's clustering and C
    as digests (perm, offsets, C row_ptr, C columns sorted within rows) plus C's row sums
    (summed over blocks of `block` rows)."""
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        write_csr_bin(os.path.join(d, "A.csr"), A)
        run(["lespgemm", os.path.join(d, "A.csr"), os.path.join(d, "A.csr"), d, "64", "16", "7", "0", "8", "3",
             "64"], d)
        perm, off = rd(d, "perm", np.int64), rd(d, "offsets", np.int64)
        crp = rd(d, "C_rowptr", np.int64)
        ccol, cval = rd(d, "C_col", np.int64), rd(d, "C_val", np.float64)
        rows = np.repeat(np.arange(A.n_rows), np.diff(crp))
        o = np.lexsort((ccol, rows))
        rs = np.bincount(rows, weights=cval, minlength=A.n_rows)
        out = {"input_digest": digest(A.rowptr, A.col, A.val), "perm_digest": digest(perm),
               "offsets_digest": digest(off), "nclusters": len(off) - 1, "C_rowptr_digest": digest(crp),
               "C_col_digest": digest(ccol[o]), "nnz": int(crp[-1]), "block": block,
               "C_rowsum": rs.reshape(-1, block).sum(axis=1), "flops": rd(d, "flops", np.int64),
               "ccl_meta": rd(d, "ccl_meta", np.int64)}
    np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **out)
    print("wrote", name, "nnz(C)", out["nnz"], "clusters", out["nclusters"])


def panels_case(name, scale, share=0.02):
    """C = A * A on R-MAT(scale, 16, seed 0) without forming C: the reference's clustering
    fingerprints and its LeSpGEMM_VLength checksums on two cluster ranges of about `share` of
    the flop each (the ranges the sampled-panel benchmark uses)."""
    from sparseops_amd.lespgemm import cluster_flops, sample_ranges
    A = rmat(scale, 16, seed=0)
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        path = os.path.join(d, "A.csr")
        write_csr_bin(path, A)
        env = dict(os.environ, OMP_NUM_THREADS=str(os.cpu_count() or 8))
        subprocess.check_call([HARNESS, "bench_lespgemm_panels", path, path, env["OMP_NUM_THREADS"], "0", "64", "16",
                               "7", "8", "3", "", d], stdout=subprocess.DEVNULL, env=env)
        perm, off = rd(d, "perm", np.int64), rd(d, "offsets", np.int64)
        picks = sample_ranges(cluster_flops(A, A.rowptr, perm, off), share)
        rl = ",".join(f"{a}:{b}" for a, b in picks)
        txt = subprocess.check_output([HARNESS, "bench_lespgemm_panels", path, path, env["OMP_NUM_THREADS"], "0",
                                       "64", "16", "7", "8", "3", rl], env=env, text=True)
    import json
    lines = [json.loads(t) for t in txt.splitlines() if t.strip()]
    head, ps = lines[0], lines[1:]
    out = {"input_digest": digest(A.rowptr, A.col, A.val), "scale": scale, "share": share,
           "perm_digest": digest(perm), "offsets_digest": digest(off), "nclusters": len(off) - 1,
           "perm_fingerprint": np.uint64(head["perm_fingerprint"]),
           "offsets_fingerprint": np.uint64(head["offsets_fingerprint"]),
           "ranges": np.array(picks, np.int64),
           "nnzc": np.array([p["nnzc"] for p in ps], np.int64), "nnzv": np.array([p["nnzv"] for p in ps], np.int64),
           "colid_sum": np.array([int(p["colid_sum"]) for p in ps], np.int64),
           "nonzeros": np.array([int(p["nonzeros"]) for p in ps], np.int64),
           "sum": np.array([p["sum"] for p in ps]), "sum_abs": np.array([p["sum_abs"] for p in ps])}
    np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **out)
    print("wrote", name, "ranges", picks, "nnzc", out["nnzc"])


def rowwise_case(name, A, B, prec=64, full=False):
    """The row-wise LeSpGEMM (kernel flags 1, 2, 3, sortOutput = true) on A * B.  A must not have
    more columns than B: the reference sizes its hash tables from the products of A's columns
    below B.num_cols only (src/spgemm_bin.cpp:100-104), and a wider A overflows a table, whose
    probe loop then never ends."""
    out = {"A_digest": digest(A.rowptr, A.col, A.val), "B_digest": digest(B.rowptr, B.col, B.val),
           "precision": np.int64(prec)}
    with tempfile.TemporaryDirectory() as d:
        write_csr_bin(os.path.join(d, "A.csr"), A)
        write_csr_bin(os.path.join(d, "B.csr"), B)
        for kern in (1, 2, 3):
            run(["rowwise", os.path.join(d, "A.csr"), os.path.join(d, "B.csr"), d, str(kern), str(prec)], d)
            rp, col, val = rd(d, "C_rowptr", np.int64), rd(d, "C_col", np.int64), rd(d, "C_val", np.float64)
            out[f"k{kern}_rowptr"] = rp
            out[f"k{kern}_col_digest"] = digest(col)
            rows = np.repeat(np.arange(A.n_rows), np.diff(rp))
            out[f"k{kern}_rowsum"] = np.bincount(rows, weights=val, minlength=A.n_rows)
            if full:
                out[f"k{kern}_col"] = col
                out[f"k{kern}_val"] = val
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **out)


def topk_case(name, A, top_k=7, cs=8, prec=64):
    """HashSpGEMMTopK and the test_spgemm_hc flow on its pairs (text round trip,
    hierarchical_clustering_v0, reordered_dict_to_permutation_and_offsets).  A needs at least as
    many rows as columns (the hash-table sizing of rowwise_case, with B = A^T)."""
    out = {"A_digest": digest(A.rowptr, A.col, A.val), "top_k": np.int64(top_k), "cluster_size": np.int64(cs),
           "precision": np.int64(prec)}
    with tempfile.TemporaryDirectory() as d:
        write_csr_bin(os.path.join(d, "A.csr"), A)
        run(["topk_hc", os.path.join(d, "A.csr"), d, str(top_k), str(cs), str(prec)], d)
        out["topk_rowptr"] = rd(d, "topk_rowptr", np.int64)
        out["topk_col"] = rd(d, "topk_col", np.int64)
        out["topk_val"] = rd(d, "topk_val", np.float64)
        out["pair_keys"] = rd(d, "pair_keys", np.int64).reshape(-1, 2)
        out["pair_scores"] = rd(d, "pair_scores", np.float64)
        out["perm"] = rd(d, "perm", np.int64)
        out["offsets"] = rd(d, "offsets", np.int64)
        out["roots"] = rd(d, "roots", np.int64)
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **out)


def reader_case(name="reader_mtx"):
    """The reference's read_csr_matrix on the Matrix Market files of tests/golden/mtx/ (file
    order kept, symmetric mirroring, pattern -> 1.0, duplicates kept or summed by compact)."""
    src = os.path.join(OUT, "mtx")
    out = {}
    for fn in sorted(os.listdir(src)):
        base = fn[:-4]
        for compact in (0, 1):
            with tempfile.TemporaryDirectory() as d:
                run(["read_mtx", os.path.join(src, fn), d, str(compact)], d)
                dims = rd(d, "dims", np.int64)
                out[f"{base}_c{compact}_dims"] = dims
                if dims[0] > 0:
                    out[f"{base}_c{compact}_rowptr"] = rd(d, "A_rowptr", np.int64)
                    out[f"{base}_c{compact}_col"] = rd(d, "A_col", np.int64)
                    out[f"{base}_c{compact}_val"] = rd(d, "A_val", np.float64)
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **out)


def features_text_case(name, A=None, mtx=None):
    """The reference's whole feature printout (test/Get_features.cpp: MtxLoad, CalculateFeatures,
    CalculateTilesExtraFeatures, FeaturesPrint, ExtraFeaturesPrint) as text, for a generated
    matrix or a file of tests/golden/mtx/."""
    with tempfile.TemporaryDirectory() as d:
        path = mtx if mtx else os.path.join(d, name + ".mtx")
        if A is not None:
            write_mtx(path, A)
        text = subprocess.check_output([HARNESS, "features_full", path, "3"], text=True)
    keep = [ln for ln in text.splitlines() if not ln.startswith("- ")]      # drop the file-path lines
    with open(os.path.join(OUT, "features_text", name + ".txt"), "w") as f:
        f.write("\n".join(keep) + "\n")


def main():
    if not os.path.exists(HARNESS):
        sys.exit(f"{HARNESS} missing: run make -C oracle ref (needs the reference sources)")
    only = set(sys.argv[1:])
    if only:   # regenerate the named groups only: classify, featurestext, fingerprints, panels, reader, rowwise, topk
        groups = {"classify": lambda: (classify_case("classify_rmat10", rmat(10, 16, seed=10), [-1.0, 2e-4, 5e-5, 0.0]),
                                       classify_case("classify_rmat12", rmat(12, 16, seed=12), [-1.0, 2e-3, 1e-3, 2e-4])),
                  "fingerprints": lambda: (fingerprint_case("lsh_rmat14_fp", rmat(14, 16, seed=14)),
                                           fingerprint_case("lsh_rmat16_fp", rmat(16, 16, seed=16), block=16)),
                  "panels": lambda: panels_case("lsh_rmat20_panels", 20),
                  "rowwise": lambda: (rowwise_case("rowwise_rmat10", rmat(10, 16, seed=10), rmat(10, 16, seed=10)),
                                      rowwise_case("rowwise_rect_f64", random_rect(300, 200, 12, seed=5),
                                                   random_rect(200, 500, 9, seed=6), full=True),
                                      rowwise_case("rowwise_rect_f32", random_rect(300, 200, 12, seed=5),
                                                   random_rect(200, 500, 9, seed=6), prec=32, full=True)),
                  "reader": reader_case,
                  "featurestext": lambda: (
                      os.makedirs(os.path.join(OUT, "features_text"), exist_ok=True),
                      features_text_case("ft_rmat10", rmat(10, 16, seed=10)),
                      features_text_case("ft_pl3000", powerlaw(3000, lmax=300, seed=2)),
                      features_text_case("ft_rect700x1900", random_rect(700, 1900, 20, seed=1)),
                      features_text_case("ft_lap40", laplacian_2d(40)),
                      [features_text_case("ft_" + f[:-4], mtx=os.path.join(OUT, "mtx", f))
                       for f in sorted(os.listdir(os.path.join(OUT, "mtx")))]),
                  "topk": lambda: (topk_case("topk_rmat10", rmat(10, 16, seed=10)),
                                   topk_case("topk_rmat10_f32", rmat(10, 16, seed=10), prec=32),
                                   topk_case("topk_rmat12_k16", rmat(12, 16, seed=12), top_k=16, cs=4),
                                   topk_case("topk_rect900x400", random_rect(900, 400, 10, seed=7), top_k=5))}
        for g in sorted(only):
            groups[g]()
        return
    spmv_case("spmv_lap16_f64", laplacian_2d(16), 64, 8, 4, 4, 64, True)
    spmv_case("spmv_lap16_f32", laplacian_2d(16), 64, 8, 4, 8, 32, True)
    spmv_case("spmv_pl3000_f64", powerlaw(3000, lmax=300, seed=2), 256, 32, 16, 4, 64, False)
    spmv_case("spmv_pl3000_f32", powerlaw(3000, lmax=300, seed=2), 256, 32, 16, 8, 32, False)
    R8, R10, R12 = rmat(8, 16, seed=8), rmat(10, 16, seed=10), rmat(12, 16, seed=12)
    lespgemm_case("lsh_rmat8", R8, R8, full_pairs=True, full_c=True)
    lespgemm_case("lsh_rmat10", R10, R10, full_pairs=True)
    lespgemm_case("lsh_rmat10_AAt", R10, transpose(R10), full_pairs=False)
    lespgemm_case("lsh_rmat10_hc1", R10, R10, hc_v=1, full_pairs=False)
    lespgemm_case("lsh_rmat10_hc2", R10, R10, hc_v=2, full_pairs=False)
    lespgemm_case("lsh_rmat10_f32", R10, R10, prec=32, full_pairs=False)
    lespgemm_case("lsh_rmat10_k128", R10, R10, k=128, bands=32, seed=42, full_pairs=False)
    lespgemm_case("lsh_rmat12", R12, R12, full_pairs=False)
    features_case("features_pl3000", powerlaw(3000, lmax=300, seed=2))
    features_case("features_lap40", laplacian_2d(40))
    features_case("features_rmat10", rmat(10, 16, seed=10))
    features_case("features_rect700x1900", random_rect(700, 1900, 20, seed=1))
    features_case("features_rect100x60", random_rect(100, 60, 12, seed=4))


if __name__ == "__main__":
    main()
