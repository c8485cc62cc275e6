"""LeSpGEMM parity on the GPU (through the C ABI): MinHash signatures, LSH pairs, clusters and
permutations bit-exact; C structure (row_ptr, sorted col_idx) bit-exact; FP64 values within
1e-12 relative (FP32: 2e-6) of the oracle and of the reference fixtures."""
import hashlib
import os

import numpy as np
import pytest
import torch

import oracle_bind as orc
from sparseops_amd import lespgemm as LG
from sparseops_amd.gen import Csr, rmat, transpose

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gold(name):
    return np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


@pytest.mark.parametrize("scale,k,seed", [(10, 64, 7), (10, 128, 42), (9, 32, 12345), (12, 64, 7), (8, 200, 3)])
def test_minhash_bit_exact(scale, k, seed, dev):
    A = rmat(scale, 16, seed=scale)
    got = LG.minhash_signatures(LG.DeviceCsr(A, dev), k, seed).cpu().numpy().view(np.uint64)
    np.testing.assert_array_equal(got, orc.minhash(A, k, seed))


@pytest.mark.parametrize("name", ["lsh_rmat8", "lsh_rmat10", "lsh_rmat10_k128", "lsh_rmat12", "lsh_rmat10_hc1"])
def test_lsh_pairs_match_reference(name, dev):
    g = gold(name)
    scale = {"lsh_rmat8": 8, "lsh_rmat12": 12}.get(name, 10)
    A = rmat(scale, 16, seed=scale)
    dA = LG.DeviceCsr(A, dev)
    sig = LG.minhash_signatures(dA, int(g["k"]), int(g["seed"])).cpu().numpy().view(np.uint64)
    assert digest(sig) == str(g["sig_digest"])
    pi, pj, ps = LG.lsh_candidate_pairs(dA, int(g["k"]), int(g["bands"]), int(g["seed"]),
                                        constrain=int(g["hc_v"]) == 0)
    assert len(pi) == int(g["npairs"])
    assert digest(pi, pj, ps) == str(g["pairs_digest"])


def test_lsh_large_bucket_window(dev):
    """300 identical rows + 700 empty rows: buckets above 256 members use the 16-window."""
    n = 1000
    lens = np.r_[np.ones(300, np.int64), np.zeros(700, np.int64)]
    rp = np.r_[0, np.cumsum(lens)]
    A = Csr(n, 50, rp, np.full(300, 7, np.int64), np.ones(300))
    dA = LG.DeviceCsr(A, dev)
    for constrain in (True, False):
        got = LG.lsh_candidate_pairs(dA, 8, 2, 5, constrain=constrain)
        want = orc.lsh_pairs(orc.minhash(A, 8, 5), 8, 2, constrain)
        for a, b in zip(got, want):
            np.testing.assert_array_equal(a, b)


def test_lsh_score_f32(dev):
    A = rmat(10, 16, seed=10)
    pi, pj, ps = LG.lsh_candidate_pairs(LG.DeviceCsr(A, dev), 64, 16, 7, score_f32=True)
    wi, wj, ws = orc.lsh_pairs(orc.minhash(A, 64, 7), 64, 16)
    np.testing.assert_array_equal(pi, wi)
    np.testing.assert_array_equal(ps, ws.astype(np.float32))


@pytest.mark.parametrize("scale,cs", [(10, 8), (12, 8), (11, 4)])
def test_cluster_build_matches_oracle(scale, cs, dev):
    A = rmat(scale, 16, seed=scale)
    sig = orc.minhash(A, 64, 7)
    perm, off = orc.hc(A, *orc.lsh_pairs(sig, 64, 16), cluster_size=cs)
    Acl = LG.ClusterMatrix.build(LG.DeviceCsr(A, dev), perm, off)
    csz, rp, rpv, col, val = Acl.download()
    wr, wrv, wcsz, wcol, wval = orc.cluster_build(A, perm, off)
    np.testing.assert_array_equal(csz, wcsz)
    np.testing.assert_array_equal(rp, wr)
    np.testing.assert_array_equal(rpv, wrv)
    np.testing.assert_array_equal(col, wcol)
    np.testing.assert_array_equal(val, wval)


CASES = [("lsh_rmat8", 8, None, 64), ("lsh_rmat10", 10, None, 64), ("lsh_rmat10_AAt", 10, "T", 64),
         ("lsh_rmat10_hc1", 10, None, 64), ("lsh_rmat10_hc2", 10, None, 64), ("lsh_rmat10_f32", 10, None, 32),
         ("lsh_rmat12", 12, None, 64)]


@pytest.mark.parametrize("name,scale,B_of,prec", CASES)
def test_lespgemm_pipeline_matches_reference(name, scale, B_of, prec, dev):
    g = gold(name)
    A = rmat(scale, 16, seed=scale)
    B = transpose(A) if B_of == "T" else A
    out = LG.lespgemm(A, B, k=int(g["k"]), bands=int(g["bands"]), seed=int(g["seed"]),
                      cluster_size=int(g["cluster_size"]), hc_v=int(g["hc_v"]), device=dev, precision=prec)
    np.testing.assert_array_equal(out["perm"], g["perm"])
    np.testing.assert_array_equal(out["offsets"], g["offsets"])
    assert out["flop"] == int(g["flops"][0])
    assert out["nclusters"] == int(g["acl_meta"][0]) and out["nnzc_A"] == int(g["acl_meta"][1])
    np.testing.assert_array_equal(out["cmin"], g["ccl_min"])
    np.testing.assert_array_equal(out["cmax"], g["ccl_max"])
    np.testing.assert_array_equal(np.r_[0, np.cumsum(out["nz"])], g["ccl_rowptr"])
    np.testing.assert_array_equal(out["C_rowptr"], g["C_rowptr"])
    assert digest(out["C_col"]) == str(g["C_col_digest"])
    rows = np.repeat(np.arange(A.n_rows), np.diff(out["C_rowptr"]))
    rs = np.bincount(rows, weights=out["C_val"].astype(np.float64), minlength=A.n_rows)
    np.testing.assert_allclose(rs, g["C_rowsum"], rtol=1e-12 if prec == 64 else 2e-6, atol=0)
    if prec == 64:
        want = orc.lespgemm(A, B, k=int(g["k"]), bands=int(g["bands"]), seed=int(g["seed"]),
                            cluster_size=int(g["cluster_size"]), hc_version=1 if int(g["hc_v"]) == 1 else 0,
                            constrain=int(g["hc_v"]) == 0, use_long_double=True)
        np.testing.assert_array_equal(out["C_col"], want["C_col"])
        rel = np.abs(out["C_val"] - want["C_val"]) / np.abs(want["C_val"])
        assert rel.max() <= 1e-12, rel.max()


def test_lespgemm_exercises_both_accumulators(dev):
    A = rmat(12, 16, seed=12)
    out = LG.lespgemm(A, A, device=dev)
    assert out["stats"]["small_clusters"] > 0 and out["stats"]["large_clusters"] > 0


def test_lespgemm_error_paths(dev):
    A = rmat(6, 4, seed=1)
    with pytest.raises(ValueError):
        LG.lespgemm(A, rmat(7, 4, seed=1), device=dev)
    with pytest.raises(ValueError):
        LG.lespgemm(A, A, k=10, bands=3, device=dev)


def _random_b(rows, cols, per_row, seed, shuffle=False):
    """B with `per_row` distinct uniform columns per row, values U(0,1]; rows sorted unless
    shuffle (then each row's entries are in random order)."""
    rng = np.random.default_rng(seed)
    col = np.concatenate([np.sort(rng.choice(cols, size=per_row, replace=False)) for _ in range(rows)])
    val = 1.0 - rng.random(col.size)
    if shuffle:                                   # the same entries, each row in another order
        p = np.concatenate([r * per_row + np.random.default_rng(seed + 1 + r).permutation(per_row)
                            for r in range(rows)])
        col, val = col[p], val[p]
    rp = (np.arange(rows + 1) * per_row).astype(np.int64)
    return Csr(rows, cols, rp, col.astype(np.int64), val)


def _check_against_oracle(A, B, out):
    want = orc.lespgemm(A, B, use_long_double=True)
    np.testing.assert_array_equal(out["perm"], want["perm"])
    np.testing.assert_array_equal(out["C_rowptr"], want["C_rowptr"])
    np.testing.assert_array_equal(out["C_col"], want["C_col"])
    np.testing.assert_array_equal(out["cmin"], want["cmin"])
    np.testing.assert_array_equal(out["cmax"], want["cmax"])
    rel = np.abs(out["C_val"] - want["C_val"]) / np.abs(want["C_val"])
    assert rel.max() <= 1e-12, rel.max()


def test_lespgemm_wide_b_several_bitmap_tiles(dev):
    """B with 600,000 columns: the large-cluster symbolic and column emission run several
    LDS bitmap tiles of 2^18 columns."""
    A = rmat(10, 16, seed=3)
    B = _random_b(A.n_cols, 600_000, 48, seed=4)
    out = LG.lespgemm(A, B, device=dev)
    assert out["stats"]["large_clusters"] > 0 and int(out["cmax"].max()) >= 2 * (1 << 18)
    _check_against_oracle(A, B, out)


@pytest.mark.parametrize("case", ["rmat", "wide_b", "narrow_b", "wide_rows"])
def test_heavy_clusters_split_bitmaps(case, dev):
    """Every large cluster heavy (threshold 1 product): its symbolic bitmap is built in parts
    of the A columns by several workgroups, merged, kept by the plan, and C's columns are
    emitted from it -- whole and in panels.  B with 600,000 columns (several 2^18-column
    tiles with B-row cursors), with 700 columns (one tile, whole rows), and A rows of up to
    3000 columns (parts of at most 1024 A columns): same C as the oracle."""
    if case == "rmat":
        A = rmat(11, 16, seed=5)
        B = _random_b(A.n_cols, A.n_cols, 40, seed=6)
    elif case == "wide_b":
        A = rmat(10, 16, seed=3)
        B = _random_b(A.n_cols, 600_000, 48, seed=4)
    elif case == "narrow_b":
        A = rmat(10, 16, seed=8)
        B = _random_b(A.n_cols, 700, 30, seed=9)
    else:
        rng = np.random.default_rng(31)
        n = 5000
        lens = rng.integers(1, 10, n)
        lens[[3, 900, 4999]] = [3000, 2100, 1500]
        cols = [np.sort(rng.choice(n, size=L, replace=False)) for L in lens]
        rp = np.zeros(n + 1, np.int64)
        np.cumsum(lens, out=rp[1:])
        col = np.concatenate(cols).astype(np.int64)
        A = Csr(n, n, rp, col, rng.random(col.size) + 0.5)
        B = A
    with LG.heavy_products(1):
        out = LG.lespgemm(A, B, device=dev)
    assert out["stats"]["large_clusters"] > 0
    _check_against_oracle(A, B, out)
    # panels of a plan made with the heavy path, against the whole product
    dA, dB = LG.DeviceCsr(A, dev), LG.DeviceCsr(B, dev)
    perm, off = out["perm"], out["offsets"]
    Acl = LG.ClusterMatrix.build(dA, perm, off)
    whole = Acl.multiply(dB)
    csz, rp, rpv, col, val = whole.download()
    with LG.heavy_products(1):
        plan = LG.SpgemmPlan(Acl, dB)
    P = None
    for c0, c1 in plan.panel_ranges(1 << 16):
        P = plan.numeric(c0, c1, P)
        pcsz, prp, prpv, pcol, pval = P.download()
        np.testing.assert_array_equal(prp, rp[c0:c1 + 1] - rp[c0])
        np.testing.assert_array_equal(pcol, col[rp[c0]:rp[c1]])
        np.testing.assert_allclose(pval, val[rpv[c0]:rpv[c1]], rtol=1e-12, atol=0)
    plan.close()


def test_wide_clusters_over_wide_b_keep_cursors_in_memory(dev):
    """Rows of 1100-2500 A columns times a B of 600,000 columns with 1-3 entries per row: the
    clusters are not heavy (few products) but have more columns than the LDS cursors hold, so
    the symbolic and emit bitmap passes keep their B-row cursors in device memory across the
    2^18-column tiles; in every wide-cursor mode the product equals the oracle's."""
    rng = np.random.default_rng(41)
    n, nb_cols = 3000, 600_000
    lens = rng.integers(1, 8, n)
    lens[[7, 1200, 2999]] = [2500, 1100, 1800]
    cols = [np.sort(rng.choice(n, size=L, replace=False)) for L in lens]
    rp = np.zeros(n + 1, np.int64)
    np.cumsum(lens, out=rp[1:])
    col = np.concatenate(cols).astype(np.int64)
    A = Csr(n, n, rp, col, rng.random(col.size) + 0.5)
    blens = rng.integers(1, 4, n)
    bcols = [np.sort(rng.choice(nb_cols, size=L, replace=False)) for L in blens]
    brp = np.zeros(n + 1, np.int64)
    np.cumsum(blens, out=brp[1:])
    bcol = np.concatenate(bcols).astype(np.int64)
    B = Csr(n, nb_cols, brp, bcol, rng.random(bcol.size) + 0.5)
    for mode in (0, 1):
        with LG.wide_cursors(mode):
            out = LG.lespgemm(A, B, device=dev)
        assert int(out["cmax"].max()) - int(out["cmin"].min()) >= (1 << 18)
        _check_against_oracle(A, B, out)


def test_lespgemm_chunked_accumulator(dev):
    """Clusters whose C block has more than 1024 columns: the numeric pass runs several LDS
    accumulator chunks per cluster."""
    A = rmat(11, 16, seed=5)
    B = _random_b(A.n_cols, A.n_cols, 96, seed=6)
    out = LG.lespgemm(A, B, device=dev)
    assert int(out["nz"].max()) > 1024
    _check_against_oracle(A, B, out)


def test_lespgemm_unsorted_b_rows(dev):
    """B rows in random order give the same C (B is sorted into a copy first)."""
    A = rmat(10, 16, seed=7)
    Bs = _random_b(A.n_cols, A.n_cols, 24, seed=8)
    Bu = _random_b(A.n_cols, A.n_cols, 24, seed=8, shuffle=True)
    o1 = LG.lespgemm(A, Bs, device=dev)
    o2 = LG.lespgemm(A, Bu, device=dev)
    assert o1["stats"]["b_rows_sorted_copy"] == 0 and o2["stats"]["b_rows_sorted_copy"] == 1
    np.testing.assert_array_equal(o1["C_rowptr"], o2["C_rowptr"])
    np.testing.assert_array_equal(o1["C_col"], o2["C_col"])
    np.testing.assert_allclose(o2["C_val"], o1["C_val"], rtol=1e-12, atol=0)
    _check_against_oracle(A, Bs, o1)


@pytest.mark.parametrize("scale,budget", [(12, 1 << 16), (12, 1 << 20), (10, 1)])
def test_panels_match_the_whole_product(scale, budget, dev):
    """The product computed in cluster panels (output buffers reused from panel to panel)
    equals the whole product panel by panel: structure exactly, values to 1e-12 (the LDS
    accumulators add in a run-dependent order); budget 1 gives one cluster per panel.
    Checksums of the panels add up to the whole product's."""
    A = rmat(scale, 16, seed=scale)
    dA = LG.DeviceCsr(A, dev)
    sig = orc.minhash(A, 64, 7)
    perm, off = orc.hc(A, *orc.lsh_pairs(sig, 64, 16), cluster_size=8)
    Acl = LG.ClusterMatrix.build(dA, perm, off)
    whole = Acl.multiply(dA)
    csz, rp, rpv, col, val = whole.download()
    plan = LG.SpgemmPlan(Acl, dA)
    assert (plan.clusters, plan.nnzc, plan.nnzv) == (Acl.rows, whole.nnzc, whole.nnzv)
    ranges = plan.panel_ranges(budget)
    assert ranges[0][0] == 0 and ranges[-1][1] == Acl.rows
    assert all(a < b for a, b in ranges) and all(ranges[i][1] == ranges[i + 1][0] for i in range(len(ranges) - 1))
    if budget == 1:
        assert len(ranges) == Acl.rows
    elif budget == 1 << 16:
        assert len(ranges) > 4
    sums = np.zeros(4)
    P = None
    for c0, c1 in ranges[:: max(1, len(ranges) // 40)] if budget == 1 else ranges:
        P = plan.numeric(c0, c1, P)
        pcsz, prp, prpv, pcol, pval = P.download()
        np.testing.assert_array_equal(pcsz, csz[c0:c1])
        np.testing.assert_array_equal(prp, rp[c0:c1 + 1] - rp[c0])
        np.testing.assert_array_equal(prpv, rpv[c0:c1 + 1] - rpv[c0])
        np.testing.assert_array_equal(pcol, col[rp[c0]:rp[c1]])
        np.testing.assert_allclose(pval, val[rpv[c0]:rpv[c1]], rtol=1e-12, atol=0)
        assert P.csr_rows == int(csz[c0:c1].sum())
        ck = P.checksum()
        assert ck["colid_sum"] == int(pcol.sum()) and ck["nonzeros"] == int(np.count_nonzero(pval))
        np.testing.assert_allclose(ck["sum"], pval.sum(), rtol=1e-12)
        sums += [ck["sum"], ck["sum_abs"], ck["colid_sum"], ck["nonzeros"]]
    if budget != 1:
        wk = whole.checksum()
        np.testing.assert_allclose(sums[:2], [wk["sum"], wk["sum_abs"]], rtol=1e-12)
        assert sums[2] == wk["colid_sum"] and sums[3] == wk["nonzeros"]
    with pytest.raises(LG.L.SpopsError):
        plan.numeric(0, Acl.rows + 1)
    plan.close()


def test_lsh_more_raw_pairs_than_one_wave_each_fits_in_a_launch(dev):
    """2^20 identical rows: one bucket per band, each row paired with its next 16 (constrain),
    16 bands -> 2.7e8 raw pairs before de-duplication, more waves than one launch can hold
    (the scoring kernel strides).  Every one of the 16.8e6 unique pairs (i, i + d), d <= 16,
    has score 1."""
    n = 1 << 20
    A = Csr(n, 1, np.arange(n + 1, dtype=np.int64), np.zeros(n, np.int64), np.ones(n))
    pi, pj, sc = LG.lsh_candidate_pairs(LG.DeviceCsr(A, dev), 64, 16, 7, constrain=True)
    d = np.arange(1, 17)
    wi = np.repeat(np.arange(n, dtype=np.int64), 16)
    wj = wi + np.tile(d, n)
    keep = wj < n
    np.testing.assert_array_equal(pi, wi[keep])
    np.testing.assert_array_equal(pj, wj[keep])
    assert np.all(sc == 1.0)


@pytest.mark.parametrize("threshold,budget", [(500, 1 << 40), (2000, 1 << 18)])
def test_dense_accumulator_clusters(threshold, budget, dev):
    """Clusters with at least `threshold` products take the dense path (device-memory row
    block, many workgroups per cluster, atomic adds, gather at the sorted columns): the
    product equals the LDS-chunk path's, structure exactly and values to 1e-12, whole or in
    panels."""
    A = rmat(12, 16, seed=3)
    dA = LG.DeviceCsr(A, dev)
    sig = orc.minhash(A, 64, 7)
    perm, off = orc.hc(A, *orc.lsh_pairs(sig, 64, 16), cluster_size=8)
    Acl = LG.ClusterMatrix.build(dA, perm, off)
    ref = Acl.multiply(dA)
    csz, rp, rpv, col, val = ref.download()
    assert ref.stats["dense_clusters"] == 0
    with LG.dense_products(threshold):
        plan = LG.SpgemmPlan(Acl, dA)
    P = None
    dense = 0
    for c0, c1 in plan.panel_ranges(budget):
        P = plan.numeric(c0, c1, P)
        dense += P.stats["dense_clusters"]
        pcsz, prp, prpv, pcol, pval = P.download()
        np.testing.assert_array_equal(prp, rp[c0:c1 + 1] - rp[c0])
        np.testing.assert_array_equal(pcol, col[rp[c0]:rp[c1]])
        np.testing.assert_allclose(pval, val[rpv[c0]:rpv[c1]], rtol=1e-12, atol=0)
    assert dense > 10
    plan.close()


@pytest.mark.parametrize("products,budget", [(64, 1 << 40), (1000, 1 << 18)])
def test_split_cluster_tasks(products, budget, dev):
    """Large clusters split over workgroups by ranges of their output-column chunks (each task
    starts its B row cursors at its first chunk's column): the product equals the unsplit one,
    structure exactly and values to 1e-12, whole or in panels."""
    A = rmat(12, 16, seed=5)
    dA = LG.DeviceCsr(A, dev)
    sig = orc.minhash(A, 64, 7)
    perm, off = orc.hc(A, *orc.lsh_pairs(sig, 64, 16), cluster_size=8)
    Acl = LG.ClusterMatrix.build(dA, perm, off)
    with LG.task_products(1 << 40):
        ref = Acl.multiply(dA)
    csz, rp, rpv, col, val = ref.download()
    plan = LG.SpgemmPlan(Acl, dA)
    P = None
    with LG.task_products(products):
        for c0, c1 in plan.panel_ranges(budget):
            P = plan.numeric(c0, c1, P)
            pcsz, prp, prpv, pcol, pval = P.download()
            np.testing.assert_array_equal(prp, rp[c0:c1 + 1] - rp[c0])
            np.testing.assert_array_equal(pcol, col[rp[c0]:rp[c1]])
            np.testing.assert_allclose(pval, val[rpv[c0]:rpv[c1]], rtol=1e-12, atol=0)
    plan.close()


def test_wide_clusters_keep_cursors(dev):
    """Rows of 1500-3000 columns make clusters with more than 1024 A columns, whose tasks run
    in the numeric variant with 4096 LDS cursors when it is on (mode 1; also split over
    workgroups) and in the regular variant with per-chunk searches otherwise (also with the
    per-call task size, products 0): same product as the oracle, structure exactly and values
    to 1e-12."""
    rng = np.random.default_rng(21)
    n = 6000
    lens = rng.integers(1, 12, n)
    lens[[5, 700, 3001, 5999]] = [3000, 1500, 2200, 2600]
    cols = [np.sort(rng.choice(n, size=L, replace=False)) for L in lens]
    rp = np.zeros(n + 1, np.int64)
    np.cumsum(lens, out=rp[1:])
    col = np.concatenate(cols).astype(np.int64)
    A = Csr(n, n, rp, col, rng.random(col.size) + 0.5)
    want = orc.lespgemm(A, A, k=64, bands=16, seed=7, cluster_size=8, use_long_double=True)
    for products, mode in ((1 << 40, 1), (4096, 1), (0, 1), (1 << 40, 0), (0, 0), (0, 2)):
        with LG.task_products(products), LG.wide_cursors(mode):
            got = LG.lespgemm(A, A, k=64, bands=16, seed=7, cluster_size=8, device=dev)
        np.testing.assert_array_equal(got["perm"], want["perm"])
        np.testing.assert_array_equal(got["C_rowptr"], want["C_rowptr"])
        np.testing.assert_array_equal(got["C_col"], want["C_col"])
        rel = np.abs(got["C_val"] - want["C_val"]) / np.abs(want["C_val"])
        assert rel.max() <= 1e-12, rel.max()
