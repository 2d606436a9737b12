"""Pins the CPU oracle (oracle/sd_oracle.c) against independent brute force
and the reference tests' known answers.  CPU only."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import brute
from conftest import bits, self_rows_from_cross

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "fixtures.json")))
NA_INT = np.iinfo(np.int32).min


def rand_sparse(nrow, ncol, density, seed, signed=False, levels=None):
    rng = np.random.default_rng(seed)
    M = sp.random(nrow, ncol, density=density, random_state=rng, format="csc")
    if levels:
        M.data = rng.integers(1, levels + 1, size=M.nnz).astype(float)
    elif signed:
        M.data = rng.standard_normal(M.nnz)
    else:
        M.data = np.abs(rng.standard_normal(M.nnz)) + 0.01
    M.sort_indices()
    return M


def from_cols(cols):
    return sp.csc_matrix(np.array(cols, dtype=float).T)


@pytest.mark.parametrize("method", ["binary", "cosine", "euclidean", "manhattan", "pearson",
                                    "covariance"])
@pytest.mark.parametrize("dist", [True, False])
def test_slab_matches_closed_forms(oracle_lib, method, dist):
    if method in ("euclidean", "manhattan", "covariance") and not dist:
        pytest.skip("no similarity form")
    A = rand_sparse(40, 17, 0.3, 1, signed=method in ("cosine", "pearson", "covariance"))
    B = rand_sparse(40, 11, 0.3, 2, signed=method in ("cosine", "pearson", "covariance"))
    A = A.tolil()
    A[:, 3] = 0
    A = A.tocsc()
    A.eliminate_zeros()
    got = oracle_lib.slab(method, A, B, dist=dist)
    want = brute.slab(method, A, B, dist)
    if method == "binary":
        # exact integers and one IEEE division on both sides
        assert np.array_equal(bits(got), bits(want))
    else:
        np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-12)


def test_jaccard_counts_structure_not_values(oracle_lib):
    A = rand_sparse(30, 6, 0.4, 3)
    B = A.copy()
    B.data = B.data * 7.0  # same structure, different values
    np.testing.assert_array_equal(oracle_lib.slab("binary", A, A), oracle_lib.slab("binary", A, B))


def test_topk_block_brute(oracle_lib):
    rng = np.random.default_rng(5)
    D = rng.integers(0, 4, size=(37, 9)).astype(float)  # heavy ties
    D[rng.random(D.shape) < 0.1] = np.nan
    D[0, 1] = np.inf
    for dec in (False, True):
        self_row = np.zeros(9, dtype=np.int32)
        idx, val = oracle_lib.topk_block(D, 6, dec, self_row)
        ridx, rval = brute.ref_topk(D, 6, dec, False)
        assert np.array_equal(idx, ridx)
        ok = ridx != NA_INT
        assert np.array_equal(val[ok], rval[ok])


def test_topk_self_row_excluded(oracle_lib):
    rng = np.random.default_rng(6)
    D = rng.random((12, 12))
    idx, _ = oracle_lib.topk_block(D, 4, False, np.arange(1, 13, dtype=np.int32))
    ridx, _ = brute.ref_topk(D, 4, False, True)
    assert np.array_equal(idx, ridx)


def test_topk_rows_accum_streaming(oracle_lib):
    rng = np.random.default_rng(7)
    D = rng.integers(0, 5, size=(23, 31)).astype(float)
    k = 4
    for dec in (False, True):
        bi = np.full((23, k), NA_INT, dtype=np.int32, order="F")
        bv = np.full((23, k), np.nan, order="F")
        for c0 in range(0, 31, 7):
            blk = np.asfortranarray(D[:, c0:c0 + 7])
            oracle_lib.topk_rows_accum(blk, k, dec, bi, bv, c0, np.zeros(23, np.int32))
        ridx, rval = brute.ref_topk(D.T, k, dec, False)  # rows of D are the queries
        assert np.array_equal(bi, ridx)


@pytest.mark.parametrize("method", ["binary", "cosine", "euclidean", "manhattan", "pearson",
                                    "covariance"])
def test_knn_matches_full_matrix(oracle_lib, method):
    X = rand_sparse(60, 23, 0.3, 11)
    for dist in ([True, False] if method in ("binary", "cosine", "pearson") else [True]):
        dec = (not dist) if method in ("binary", "cosine", "pearson") else method == "covariance"
        idx, dst, _ = oracle_lib.knn(X, k=5, method=method, dist=dist, block_size=7)
        full = brute.slab(method, X, X, dist)
        ridx, rval = brute.ref_topk(full, 5, dec, True)
        if method == "binary":
            assert np.array_equal(idx, ridx)
        else:
            # closed forms agree to rounding; ranks agree unless values tie
            assert (idx == ridx).mean() > 0.97
        # the driver reports exactly the slab kernel's values ...
        own = oracle_lib.slab(method, X, X, dist=dist)
        assert np.array_equal(bits(dst), bits(np.take_along_axis(own.T, idx - 1, 1)))
        # ... and its ranks are exactly those of a plain sort of that slab
        oidx, _ = brute.ref_topk(own, 5, dec, True)
        assert np.array_equal(idx, oidx)
        # ... which match the closed form to rounding
        np.testing.assert_allclose(dst, np.take_along_axis(full.T, idx - 1, 1), rtol=1e-9,
                                   atol=1e-12)


def test_knn_block_size_invariance(oracle_lib):
    X = rand_sparse(50, 29, 0.2, 12, levels=3)
    base = oracle_lib.knn(X, k=5, method="binary", block_size=29)
    for bs in (1, 2, 7, 100):
        got = oracle_lib.knn(X, k=5, method="binary", block_size=bs)
        assert np.array_equal(got[0], base[0]) and np.array_equal(bits(got[1]), bits(base[1]))


def test_knn_degenerate_and_clamp(oracle_lib):
    X = rand_sparse(40, 10, 0.3, 13).toarray()
    X[:, 1] = 0
    X[:, 4] = 0
    Xs = sp.csc_matrix(X)
    idx, dst, warn = oracle_lib.knn(Xs, k=3, method="binary")
    assert (idx[[1, 4]] == NA_INT).all()
    assert not np.isin(idx[[0, 2, 3, 5, 6, 7, 8, 9]], [2, 5]).any()
    assert bits(dst[1, 0]) == 0x7FF00000000007A2
    idx, dst, warn = oracle_lib.knn(Xs, k=9, method="binary")
    assert warn and idx.shape[1] == 7  # 8 non-degenerate, self excluded


def test_knn_include_self_restores_exact_values(oracle_lib):
    X = rand_sparse(40, 12, 0.3, 14)
    idx, dst, _ = oracle_lib.knn(X, k=3, method="cosine", include_self=True)
    assert np.array_equal(idx[:, 0], np.arange(1, 13))
    assert np.array_equal(bits(dst[:, 0]), bits(np.zeros(12)))


@pytest.mark.parametrize("method", ["binary", "cosine"])
@pytest.mark.parametrize("include_self", [False, True])
def test_self_rows_from_cross_search(oracle_lib, method, include_self):
    """The GPU tests check sampled rows of large self-searches against a cross
    search with k + 1 neighbours, minus the query's own column (plus its
    restored value with include_self).  That construction must reproduce the
    oracle's self mode, including tied duplicate columns."""
    X = rand_sparse(60, 150, 0.15, 31).tolil()
    for c in range(40, 70):
        X[:, c] = X[:, 3]  # 31 identical columns: ties wider than k + 1
    X[:, 100] = 0  # an empty column: an NA row and never a candidate
    X = X.tocsc()
    X.eliminate_zeros()
    k = 8
    o = oracle_lib.knn(X, k=k, method=method, include_self=include_self)
    qcols = [0, 3, 5, 40, 41, 69, 70, 100, 149]
    c = oracle_lib.knn(X[:, qcols], X, k=k + 1, method=method)
    gi, gd = self_rows_from_cross(c[0], c[1], qcols, k, include_self)
    assert np.array_equal(gi, o[0][qcols])
    assert np.array_equal(bits(gd), bits(o[1][qcols]))


def test_cross_equals_two_knn_calls(oracle_lib):
    X = rand_sparse(80, 21, 0.1, 15, levels=2)
    Y = rand_sparse(80, 13, 0.1, 16, levels=2)
    for method in ("binary", "cosine", "euclidean"):
        f, r, _ = oracle_lib.knn_cross(X, Y, k=4, method=method, block_size=5)
        fi, fd, _ = oracle_lib.knn(X, Y, k=4, method=method, block_size=5)
        ri, rd, _ = oracle_lib.knn(Y, X, k=4, method=method, block_size=5)
        assert np.array_equal(f[0], fi) and np.array_equal(bits(f[1]), bits(fd))
        assert np.array_equal(r[0], ri) and np.array_equal(bits(r[1]), bits(rd))


def test_snn_weights_vs_sets(oracle_lib):
    X = rand_sparse(60, 30, 0.3, 17)
    idx, _, _ = oracle_lib.knn(X, k=5, method="binary")
    idx2 = idx.copy()
    idx2[3, 2] = NA_INT
    idx2[4, 1] = idx2[4, 0]  # duplicate
    idx2[5, 0] = 6          # self listed
    for inc in (True, False):
        w = oracle_lib.snn_jaccard(idx2, include_self=inc)
        want = brute.ref_snn(idx2, inc)
        assert np.array_equal(np.isnan(w), np.isnan(want))
        assert np.array_equal(bits(w[~np.isnan(w)]), bits(want[~np.isnan(want)]))


@pytest.mark.parametrize("symmetrise", [True, False])
@pytest.mark.parametrize("prune", [0.0, 1 / 15, 0.5])
def test_snn_graph_knn_edges(oracle_lib, symmetrise, prune):
    X = rand_sparse(50, 40, 0.3, 18)
    idx, _, _ = oracle_lib.knn(X, k=6, method="cosine")
    p, ri, x = oracle_lib.snn_graph(idx, include_self=True, prune=prune, symmetrise=symmetrise)
    g = sp.csc_matrix((x, ri, p), shape=(40, 40)).toarray()
    want = brute.ref_knn_graph(idx, True, prune, symmetrise)
    assert np.array_equal(bits(g), bits(want))


@pytest.mark.parametrize("prune", [0.0, 0.2, 0.5])
def test_snn_graph_shared_edges(oracle_lib, prune):
    X = rand_sparse(50, 35, 0.3, 19)
    idx, _, _ = oracle_lib.knn(X, k=4, method="binary")
    for inc in (True, False):
        p, ri, x = oracle_lib.snn_graph(idx, include_self=inc, prune=prune, edges="shared")
        g = sp.csc_matrix((x, ri, p), shape=(35, 35)).toarray()
        assert np.array_equal(bits(g), bits(brute.ref_shared(idx, inc, prune)))


# ---- the reference tests' literal known answers -----------------------------------
def _idx_array(rows):
    return np.array([[NA_INT if v is None else v for v in r] for r in rows], dtype=np.int32)


def test_fixture_topk_na(oracle_lib):
    f = FIX["topk_na_padding"]
    D = np.array([[1.0], [np.nan], [np.inf]])
    idx, val = oracle_lib.topk_block(D, 3, False, np.array([0], np.int32))
    assert idx[0].tolist() == [1, NA_INT, NA_INT]
    assert val[0, 0] == 1.0 and bits(val[0, 1]) == 0x7FF00000000007A2
    assert f["expect_idx"][0][0] == 1


def test_fixture_ties(oracle_lib):
    f = FIX["ties_lower_index"]
    X = from_cols(f["X"])
    idx, _, _ = oracle_lib.knn(X, k=3, method="euclidean")
    assert idx[f["row"] - 1].tolist() == f["expect_idx_row"]


def test_fixture_massive_ties(oracle_lib):
    f = FIX["massive_jaccard_ties"]
    X = np.zeros((f["nrow"], f["ncol_x"]))
    X[: f["ones_rows"]] = 1
    Y = np.zeros((f["nrow"], f["ncol_y"]))
    Y[: f["ones_rows"]] = 1
    for bs in f["block_sizes"]:
        (fi, fd), (ri, rd), _ = oracle_lib.knn_cross(sp.csc_matrix(X), sp.csc_matrix(Y), k=6,
                                                      method="binary", block_size=bs)
        assert (fi == np.array(f["expect_forward_row"])).all()
        assert (ri == np.array(f["expect_reverse_row"])).all()
        assert (fd == 0).all() and (rd == 0).all()


def test_fixture_signed_cancel(oracle_lib):
    f = FIX["signed_cancel_not_empty"]
    X = from_cols(f["X"])
    for m in f["methods"]:
        idx, _, _ = oracle_lib.knn(X, k=2, method=m)
        assert (idx[0] != NA_INT).all()
        assert (idx[1] == NA_INT).all()
        assert not (idx == 2).any()
        assert (idx[1:] == 1).any()


def test_fixture_cross_degenerate(oracle_lib):
    f = FIX["cross_degenerate"]
    clean, holed = from_cols(f["clean"]), from_cols(f["holed"])
    a, _, _ = oracle_lib.knn(holed, clean, k=2, method="binary")
    assert (a[1] == NA_INT).all() and (a[[0, 2]] != NA_INT).all()
    b, _, _ = oracle_lib.knn(clean, holed, k=2, method="binary")
    assert not (b == NA_INT).any() and not (b == 2).any()
    c, _, _ = oracle_lib.knn(holed, holed, k=2, method="binary")
    assert (c[1] == NA_INT).all() and not (c == 2).any()
    d, _, warn = oracle_lib.knn(clean, holed, k=3, method="binary")
    assert warn and d.shape[1] == 2


def test_fixture_pearson_constant(oracle_lib):
    f = FIX["pearson_constant_self"]
    X = from_cols(f["X"])
    idx, dst, _ = oracle_lib.knn(X, k=2, method="pearson", dist=False, include_self=True)
    at = list(idx[0]).index(1)
    assert dst[0, at] == 1.0
    idx, dst, _ = oracle_lib.knn(X, k=2, method="pearson", dist=True, include_self=True)
    assert idx[:, 0].tolist() == f["expect_dist_idx_col1"]
    assert dst[:, 0].tolist() == f["expect_dist_dist_col1"]


TESTTHAT_TOL = 1.5e-8  # testthat's expect_equal tolerance, which these asserts use


def check_cosine_conventions(slab):
    X = from_cols(FIX["cosine_extremes"]["X"])
    S = slab("cosine", X, X, False)
    assert not np.isnan(S).any()
    for i, j, v in FIX["cosine_extremes"]["expect_sim"]:
        assert abs(S[i - 1, j - 1] - v) < TESTTHAT_TOL
    X = from_cols(FIX["cosine_antiparallel"]["X"])
    assert abs(slab("cosine", X, X, False)[0, 1] + 1.0) < TESTTHAT_TOL
    f = FIX["empty_conventions"]
    X = from_cols(f["X"])
    for m in ("cosine", "binary"):
        S = slab(m, X, X, False)
        for i, j, v in f["expect_sim"]:
            assert abs(S[i - 1, j - 1] - v) < TESTTHAT_TOL, (m, i, j)
    Dd = slab("cosine", X, X, True)
    assert Dd[0, 0] == 0.0 and Dd[1, 1] == 0.0 and abs(Dd[2, 2]) < f["cosine_cross_dist_33_below"]


def test_fixture_cosine_conventions(oracle_lib):
    check_cosine_conventions(lambda m, A, B, d: oracle_lib.slab(m, A, B, dist=d))


def test_fixture_cosine_rejects_nonfinite(oracle_lib):
    X = sp.csc_matrix(np.array([[np.inf, 1.0], [0.0, 0.0]]))
    with pytest.raises(oracle_lib.OracleError, match="finite"):
        oracle_lib.slab("cosine", X, X)


def test_fixture_snn(oracle_lib):
    f = FIX["snn_self_listed"]
    w = oracle_lib.snn_jaccard(_idx_array(f["idx"]), include_self=True)
    assert w[:, 1].tolist() == f["expect_w_col2"]
    f = FIX["snn_disjoint"]
    w = oracle_lib.snn_jaccard(_idx_array(f["idx"]), include_self=False)
    assert w.tolist() == f["expect_w"]
    f = FIX["snn_na_padding"]
    w = oracle_lib.snn_jaccard(_idx_array(f["idx"]), include_self=True)
    exp = np.array([[np.nan if v is None else v for v in r] for r in f["expect_w"]])
    assert np.array_equal(np.isnan(w), np.isnan(exp))
    assert np.array_equal(bits(w[~np.isnan(w)]), bits(exp[~np.isnan(exp)]))
    for case in FIX["snn_out_of_range"]["cases"]:
        with pytest.raises(oracle_lib.OracleError, match="1-based row indices"):
            oracle_lib.snn_jaccard(_idx_array(case))
    f = FIX["snn_duplicates"]
    p, ri, x = oracle_lib.snn_graph(_idx_array(f["idx"]), include_self=True, symmetrise=False)
    g = sp.csc_matrix((x, ri, p), shape=(2, 2)).toarray()
    for i, j, v in f["expect_entries"]:
        assert g[i - 1, j - 1] == v
    f = FIX["snn_include_self_override"]
    ws = oracle_lib.snn_jaccard(_idx_array(f["idx"]), include_self=True)
    wn = oracle_lib.snn_jaccard(_idx_array(f["idx"]), include_self=False)
    assert ws[0, 1] == f["expect_with_self_1_2"] and wn[0, 1] == f["expect_without_1_2"]
    w0 = oracle_lib.snn_jaccard(_idx_array(f["idx_noself"]), include_self=True)
    assert np.array_equal(ws[:, 1], w0[:, 0])


def test_fixture_shared_awkward(oracle_lib):
    f = FIX["snn_shared_awkward"]
    idx = _idx_array(f["idx"])
    p, ri, x = oracle_lib.snn_graph(idx, include_self=True, prune=0.0, edges="shared")
    m = sp.csc_matrix((x, ri, p), shape=(6, 6)).toarray()
    assert (m != 0).sum() / 2 == f["expect_shared_pairs"]
    assert np.array_equal(m, m.T) and (np.diag(m) == 0).all()
    for i, j, v in f["expect_shared"]:
        assert m[i - 1, j - 1] == v
    p, ri, x = oracle_lib.snn_graph(idx, include_self=True, prune=0.0, edges="knn")
    knn = sp.csc_matrix((x, ri, p), shape=(6, 6)).toarray()
    off = knn.copy()
    np.fill_diagonal(off, 0)
    assert (off != 0).sum() / 2 == f["expect_knn_offdiag_pairs"]
    i, j, v = f["expect_knn_selfloop"]
    assert knn[i - 1, j - 1] == v
    for pr in f["prune_values"]:
        p, ri, x = oracle_lib.snn_graph(idx, include_self=True, prune=pr, edges="shared")
        got = sp.csc_matrix((x, ri, p), shape=(6, 6)).toarray()
        want = np.where(m > pr, m, 0.0)
        assert np.array_equal(got, want)
    p, ri, x = oracle_lib.snn_graph(idx, include_self=True, prune=1.0, edges="shared")
    assert p[-1] == 0


def _dense_js(P, Q):
    """sqrt(0.5 * sum(xlogx(p) + xlogx(q) - xlogx(p + q) + (p + q) log 2)) per
    column pair, numpy on dense arrays (0 log 0 = 0)."""
    def xlogx(x):
        with np.errstate(divide="ignore", invalid="ignore"):
            return np.where(x > 0, x * np.log(np.where(x > 0, x, 1.0)), 0.0)
    out = np.empty((P.shape[1], Q.shape[1]))
    for i in range(P.shape[1]):
        p = P[:, i:i + 1]
        s = xlogx(p) + xlogx(Q) - xlogx(p + Q) + (p + Q) * np.log(2.0)
        v = 0.5 * s.sum(axis=0)
        out[i] = np.where(v > 0, np.sqrt(np.maximum(v, 0)), 0.0)
    return out


def test_js_matches_scipy_on_normalised_columns(oracle_lib):
    """On columns that sum to 1 the reference's sum is the Jensen-Shannon
    divergence, which scipy computes another way (KL divergences to the
    midpoint distribution)."""
    from scipy.spatial.distance import jensenshannon
    A = rand_sparse(50, 13, 0.3, 21).toarray()
    B = rand_sparse(50, 9, 0.3, 22).toarray()
    A[:, 4] = 0.0
    A[:, 5] = B[:, 2] * 3.0  # same distribution, other scale
    for M in (A, B):
        s = M.sum(axis=0)
        M[:, s > 0] /= s[s > 0]
    got = oracle_lib.slab("js", sp.csc_matrix(A), sp.csc_matrix(B))
    for i in range(A.shape[1]):
        for j in range(B.shape[1]):
            if A[:, i].sum() == 0 or B[:, j].sum() == 0:
                continue
            assert abs(got[i, j] - jensenshannon(A[:, i], B[:, j])) < 1e-7, (i, j)
    # A[:, 5] is B[:, 2]: distance 0 up to rounding of the divergence sum
    assert got[5, 2] < 1e-7
    assert np.all(got <= np.sqrt(np.log(2.0)) + 1e-12) and np.all(got >= 0)


def test_js_matches_dense_closed_form(oracle_lib):
    """Unnormalised non-negative columns against the dense numpy form; a
    dropped term or a wrong sign would differ by far more than rounding."""
    A = rand_sparse(40, 11, 0.35, 23)
    B = rand_sparse(40, 7, 0.35, 24)
    got = oracle_lib.slab("js", A, B)
    np.testing.assert_allclose(got, _dense_js(A.toarray(), B.toarray()), rtol=1e-12, atol=1e-14)


def test_fixture_js(oracle_lib):
    f = FIX["js_standard_metric"]
    P = from_cols(f["P"])
    d = oracle_lib.slab("js", P, P)
    assert d[0, 1] == np.sqrt(np.log(2.0)) and d[1, 0] == np.sqrt(np.log(2.0))
    for case in FIX["js_invalid_input"]["cases"]:
        M = np.array([[np.nan if v is None else v for v in row] for row in case["M"]])
        with pytest.raises(Exception, match=case["error"]):
            oracle_lib.slab("js", sp.csc_matrix(M), sp.csc_matrix(M))
    f = FIX["js_self_exact_zero"]
    rng = np.random.default_rng(25)
    P = np.abs(rng.standard_normal((12, 6)))
    P /= P.sum(axis=0)
    idx, dst, _ = oracle_lib.knn(sp.csc_matrix(P), k=f["k"], method="js",
                                 include_self=f["include_self"])
    assert (bits(dst[:, 0]) == bits(np.zeros(6))).all()
    assert idx[:, 0].tolist() == list(range(1, 7))
    for m in FIX["distance_only_methods"]["methods"]:
        with pytest.raises(Exception, match=FIX["distance_only_methods"]["error"]):
            oracle_lib.knn(sp.csc_matrix(P), k=2, method=m, dist=False)


def test_js_knn_matches_full_matrix(oracle_lib):
    """The kNN driver's JS lists are the brute-force ranking of the dense
    closed form (values differ from it only by rounding)."""
    A = rand_sparse(30, 25, 0.4, 26)
    idx, dst, _ = oracle_lib.knn(A, k=4, method="js")
    full = _dense_js(A.toarray(), A.toarray())
    ridx, _ = brute.ref_topk(full, 4, False, True)
    assert np.array_equal(idx, ridx)
    np.testing.assert_allclose(dst, np.take_along_axis(full.T, idx - 1, 1), rtol=1e-12)
