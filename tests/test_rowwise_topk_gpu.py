"""Row-wise SpGEMM (spops_spgemm_csr) and the A * A^T top-k Jaccard front end
(spops_spgemm_topk_jaccard) on the GPU, against the oracle and the reference's own outputs:

  * C = A * B row-wise equals the reference's row-wise LeSpGEMM (kernel flags 1 and 2; flag 3 on
    structure, see test_rowwise_topk_cpu) and the oracle on R-MAT, rectangular (A wider than B
    too), empty and FP32 cases: structure exact, values 1e-12 relative to the row's sum of
    |products| (FP32: 2^-20);
  * the top-k candidates equal the reference's HashSpGEMMTopK bit for bit (columns, order and
    scores, FP64 and FP32), also when the product runs in many small panels, for top_k 0, 1
    and beyond the row count;
  * the test_spgemm_hc flow: GPU candidates -> text round trip -> map-based clustering v0 gives
    the reference's permutation and offsets;
  * the reference's own generate_candidate_pairs and test_spgemm_hc callers, compiled against the
    drop-in headers, run end to end (candidate file identical to the oracle's pairs; C equal to
    scipy's product).
"""
import os
import re
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle_bind as orc
from sparseops_amd import lespgemm as LG
from sparseops_amd.gen import Csr, random_rect, rmat
from sparseops_amd.io import read_mtx, write_mtx

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
BIN = os.path.join(ROOT, "oracle", "_ref", "callers")


def gold(name):
    return np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)


def _check_product(A, B, C, tol):
    want = orc.spgemm_rows(A, B)
    np.testing.assert_array_equal(C.rowptr, want.rowptr)
    np.testing.assert_array_equal(C.col, want.col)
    absA = Csr(A.n_rows, A.n_cols, A.rowptr, A.col, np.abs(A.val))
    absB = Csr(B.n_rows, B.n_cols, B.rowptr, B.col, np.abs(B.val))
    scale = orc.spgemm_rows(absA, absB).val
    err = np.abs(C.val - want.val) / np.maximum(scale, 1e-300)
    assert err.size == 0 or err.max() <= tol, err.max()


def _with_duplicates(A, seed):
    """A copy of A in which every third non-empty row repeats its first entry (value halved)."""
    rng = np.random.default_rng(seed)
    rp, ci, va = [0], [], []
    for i in range(A.n_rows):
        c = A.col[A.rowptr[i]:A.rowptr[i + 1]].tolist()
        v = A.val[A.rowptr[i]:A.rowptr[i + 1]].tolist()
        if i % 3 == 0 and c:
            k = int(rng.integers(0, len(c) + 1))
            c.insert(k, c[0])
            v.insert(k, v[0] * 0.5)
        ci += c
        va += v
        rp.append(len(ci))
    return Csr(A.n_rows, A.n_cols, np.array(rp), np.array(ci), np.array(va))


@pytest.mark.parametrize("case", ["rmat10", "rmat12", "rect", "rect_wide_a", "rect_f32", "empty_rows", "duplicates"])
def test_spgemm_rows_matches_oracle(case, dev):
    print(f"spgemm_rows {case}", flush=True)
    f32 = case.endswith("f32")
    if case == "rmat10":
        A = B = rmat(10, 16, seed=10)
    elif case == "rmat12":
        A, B = rmat(12, 16, seed=12), rmat(12, 8, seed=3)
    elif case in ("rect", "rect_f32"):
        A, B = random_rect(300, 200, 12, seed=5), random_rect(200, 500, 9, seed=6)
    elif case == "rect_wide_a":
        A, B = random_rect(300, 500, 12, seed=5), random_rect(500, 200, 9, seed=6)
    elif case == "duplicates":
        A, B = _with_duplicates(rmat(10, 16, seed=2), 1), _with_duplicates(rmat(10, 8, seed=5), 2)
    else:
        A, B = random_rect(200, 60, 3, seed=9), random_rect(60, 70, 2, seed=10)
    vdt = torch.float32 if f32 else torch.float64
    if f32:
        A = Csr(A.n_rows, A.n_cols, A.rowptr, A.col, A.val.astype(np.float32).astype(np.float64))
        B = Csr(B.n_rows, B.n_cols, B.rowptr, B.col, B.val.astype(np.float32).astype(np.float64))
    C = LG.spgemm_rows(LG.DeviceCsr(A, dev, vdt), LG.DeviceCsr(B, dev, vdt))
    _check_product(A, B, C, 2.0 ** -20 if f32 else 1e-12)


@pytest.mark.parametrize("name,f32", [("rowwise_rmat10", False), ("rowwise_rect_f64", False), ("rowwise_rect_f32", True)])
def test_spgemm_rows_matches_reference(name, f32, dev):
    g = gold(name)
    if name == "rowwise_rmat10":
        A = B = rmat(10, 16, seed=10)
    else:
        A, B = random_rect(300, 200, 12, seed=5), random_rect(200, 500, 9, seed=6)
    vdt = torch.float32 if f32 else torch.float64
    C = LG.spgemm_rows(LG.DeviceCsr(A, dev, vdt), LG.DeviceCsr(B, dev, vdt))
    for kern in (1, 2, 3):
        np.testing.assert_array_equal(C.rowptr, g[f"k{kern}_rowptr"])
    rows = np.repeat(np.arange(A.n_rows), np.diff(C.rowptr))
    rs = np.bincount(rows, weights=C.val.astype(np.float64), minlength=A.n_rows)
    tol = 1e-5 if f32 else 1e-12
    for kern in (1, 2):
        np.testing.assert_allclose(rs, g[f"k{kern}_rowsum"], rtol=tol, atol=tol * np.abs(rs).max())
        if f"k{kern}_col" in g:
            np.testing.assert_array_equal(C.col, g[f"k{kern}_col"])


TOPK = {"topk_rmat10": (lambda: rmat(10, 16, seed=10), 7, 64),
        "topk_rmat10_f32": (lambda: rmat(10, 16, seed=10), 7, 32),
        "topk_rmat12_k16": (lambda: rmat(12, 16, seed=12), 16, 64),
        "topk_rect900x400": (lambda: random_rect(900, 400, 10, seed=7), 5, 64)}


@pytest.mark.parametrize("name", sorted(TOPK))
def test_topk_matches_reference(name, dev):
    make, k, prec = TOPK[name]
    g = gold(name)
    A = make()
    print(f"topk {name}", flush=True)
    T = LG.topk_jaccard(LG.DeviceCsr(A, dev), k, precision=prec)
    np.testing.assert_array_equal(T.rowptr, g["topk_rowptr"])
    np.testing.assert_array_equal(T.col, g["topk_col"])
    np.testing.assert_array_equal(T.val.astype(np.float64), g["topk_val"])


def test_topk_panels_and_edge_k(dev):
    A = rmat(11, 16, seed=4)
    dA = LG.DeviceCsr(A, dev)
    for k, budget in [(7, 0), (7, 1 << 16), (1, 1 << 12), (0, 0), (5000, 0)]:
        print(f"topk k={k} budget={budget}", flush=True)
        T = LG.topk_jaccard(dA, k, budget_bytes=budget)
        W = orc.topk_jaccard(A, k)
        np.testing.assert_array_equal(T.rowptr, W.rowptr)
        np.testing.assert_array_equal(T.col, W.col)
        np.testing.assert_array_equal(T.val, W.val)


def test_topk_with_duplicates_and_empty_rows(dev):
    rng = np.random.default_rng(3)
    rp, ci = [0], []
    for i in range(500):
        m = 0 if i % 11 == 3 else int(rng.integers(1, 12))
        cols = rng.integers(0, 300, m)
        if i % 5 == 0 and m:
            cols = np.append(cols, cols[:2])
        ci.extend(cols.tolist())
        rp.append(len(ci))
    A = Csr(500, 300, np.array(rp), np.array(ci), np.ones(len(ci)))
    for prec in (64, 32):
        T = LG.topk_jaccard(LG.DeviceCsr(A, dev), 9, precision=prec)
        W = orc.topk_jaccard(A, 9, score_f32=(prec == 32))
        np.testing.assert_array_equal(T.rowptr, W.rowptr)
        np.testing.assert_array_equal(T.col, W.col)
        np.testing.assert_array_equal(T.val.astype(np.float64), W.val)


@pytest.mark.parametrize("name", ["topk_rmat10", "topk_rmat12_k16", "topk_rmat10_f32"])
def test_topk_hc_v0_flow_matches_reference(name, dev):
    """GPU candidates, written and read back as test/generate_candidate_pairs.cpp and
    test/test_spgemm_hc.cpp do ("%20.16g" text, map keyed by (i, j)), clustered by the map-based
    v0: the reference's permutation, offsets and roots."""
    make, k, prec = TOPK[name]
    g = gold(name)
    A = make()
    T = LG.topk_jaccard(LG.DeviceCsr(A, dev), k, precision=prec)
    rows = np.repeat(np.arange(A.n_rows), np.diff(T.rowptr))
    order = np.lexsort((T.col, rows))                       # map order: by (i, j)
    pi, pj = rows[order], T.col[order]
    ps = np.array([float("%20.16g" % v) for v in T.val[order].astype(np.float64)])
    if prec == 32:
        ps = ps.astype(np.float32)
    np.testing.assert_array_equal(np.stack([pi, pj], 1), g["pair_keys"])
    np.testing.assert_array_equal(ps.astype(np.float64), g["pair_scores"])
    perm, offsets, roots = LG.hierarchical_clustering_v0(A, pi, pj, ps, int(g["cluster_size"]))
    np.testing.assert_array_equal(perm, g["perm"])
    np.testing.assert_array_equal(offsets, g["offsets"])
    np.testing.assert_array_equal(roots, g["roots"])


def _run(cwd, name, *args, env=None):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} is not built: the reference's callers are compiled where its sources are present")
    print(f"run {name} {' '.join(args)}", flush=True)
    return subprocess.run([exe, *args], capture_output=True, text=True, timeout=300, cwd=cwd,
                          env=dict(os.environ, **(env or {})))


@pytest.mark.parametrize("kernel", ["1", "3"])
def test_candidate_pairs_and_spgemm_hc_callers(kernel, tmp_path):
    A = rmat(10, 16, seed=10)
    mtx = str(tmp_path / "rmat10.mtx")
    write_mtx(mtx, A)
    cand = str(tmp_path / "cand.mtx")
    p = _run(str(tmp_path), "generate_candidate_pairs", mtx, f"--output={cand}", "--topk=7",
             env={"CLOSE_PAIR_DATA_PATH": str(tmp_path)})
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    assert "Basic validation: PASSED" in p.stdout
    W = orc.topk_jaccard(A, 7)
    lines = [ln.split() for ln in open(cand) if not ln.startswith("%")][1:]
    got = np.array([[int(a), int(b)] for a, b, _ in lines])
    rows = np.repeat(np.arange(A.n_rows), np.diff(W.rowptr))
    np.testing.assert_array_equal(got, np.stack([rows, W.col], 1))          # 0-based, in score order
    np.testing.assert_array_equal(np.array([float(v) for _, _, v in lines]),
                                  np.array([float("%20.16g" % v) for v in W.val]))
    p = _run(str(tmp_path), "test_spgemm_hc", mtx, mtx, cand, "--test_type=correctness", f"--kernel={kernel}")
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    assert int(re.search(r"Clusters: (\d+)", p.stdout).group(1)) == len(gold("topk_rmat10")["offsets"]) - 1
    suffix = "hashvlengthcluster_hc" if kernel == "1" else "mixedvlengthcluster_hc"
    C = read_mtx(str(tmp_path / f"rmat10_SpOps_{suffix}.mtx"))
    S = sp.csr_matrix((A.val, A.col, A.rowptr), shape=(A.n_rows, A.n_cols))
    want = (S @ S).tocsr()
    want.sort_indices()
    got = sp.csr_matrix((C.val, C.col, C.rowptr), shape=(C.n_rows, C.n_cols))
    got.sum_duplicates()
    got.sort_indices()
    np.testing.assert_array_equal(got.indptr, want.indptr)
    np.testing.assert_array_equal(got.indices, want.indices)
    np.testing.assert_allclose(got.data, want.data, rtol=1e-12, atol=0)
