"""GPU parity: libsd_hip.so against the CPU oracle on the same seeded inputs.

Integer and index results must be identical; FP64 values identical bit for
bit (the exact path and the FP64 re-rank use the reference's operation
order).  Every call goes through the C ABI."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import sparsedist_amd as sd
from conftest import bits, self_rows_from_cross
from sparsedist_amd import synth

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "fixtures.json")))
NA_INT = np.iinfo(np.int32).min
NT = min(16, os.cpu_count() or 1)
METHODS = ["binary", "cosine", "euclidean", "manhattan", "pearson", "covariance"]


def rand_sparse(nrow, ncol, density, seed, signed=False, levels=None, empty_cols=()):
    rng = np.random.default_rng(seed)
    M = sp.random(nrow, ncol, density=density, random_state=rng, format="csc")
    if levels:
        M.data = rng.integers(1, levels + 1, size=M.nnz).astype(float)
    elif signed:
        M.data = rng.standard_normal(M.nnz)
    else:
        M.data = np.abs(rng.standard_normal(M.nnz)) + 0.01
    if empty_cols:
        M = M.tolil()
        for c in empty_cols:
            M[:, c] = 0
        M = M.tocsc()
        M.eliminate_zeros()
    M.sort_indices()
    return M


def same(gpu, ora, label=""):
    gi, gd = gpu
    oi, od = ora[0], ora[1]
    assert gi.shape == oi.shape, label
    bad = np.argwhere(gi != oi)
    assert bad.size == 0, f"{label}: idx differs at {bad[:5].tolist()} gpu={gi[tuple(bad[0])]} " \
                          f"oracle={oi[tuple(bad[0])]}"
    db = np.argwhere(bits(gd) != bits(od))
    assert db.size == 0, f"{label}: dist differs at {db[:5].tolist()} " \
                         f"gpu={gd[tuple(db[0])]!r} oracle={od[tuple(db[0])]!r}"


@pytest.mark.parametrize("method", METHODS)
def test_slab_bitwise(oracle_lib, gpu_ctx, method):
    signed = method in ("cosine", "pearson", "covariance")
    A = rand_sparse(50, 23, 0.3, 1, signed=signed, empty_cols=(2,))
    B = rand_sparse(50, 17, 0.3, 2, signed=signed, empty_cols=(5,))
    for dist in ([True, False] if method in ("binary", "cosine", "pearson") else [True]):
        g = sd.sparseDist2(A, B, method=method, dist=dist)
        o = oracle_lib.slab(method, A, B, dist=dist)
        assert np.array_equal(bits(g), bits(o)), (method, dist)


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("precision", ["auto", "bf16", "exact", "fp32"])
def test_knn_self_bitwise(oracle_lib, gpu_ctx, method, precision):
    X = rand_sparse(70, 45, 0.25, 3, signed=method in ("pearson", "covariance"),
                    empty_cols=(4, 9))
    for dist in ([True, False] if method in ("binary", "cosine", "pearson") else [True]):
        for include_self in (False, True):
            for drop in (True, False):
                lab = f"{method} {precision} dist={dist} self={include_self} drop={drop}"
                g = sd.sparseKNN(X, k=6, method=method, dist=dist, include_self=include_self,
                                 drop_degenerate=drop, precision=precision, verbose=False)
                o = oracle_lib.knn(X, k=6, method=method, dist=dist, include_self=include_self,
                                   drop_degenerate=drop, nthreads=NT)
                same((g.idx, g.dist), o, lab)
                assert g.self_search is True


@pytest.mark.parametrize("method", ["binary", "cosine", "euclidean"])
def test_knn_against_reference_matrix(oracle_lib, gpu_ctx, method):
    X = rand_sparse(90, 140, 0.2, 5, empty_cols=(7,))
    Y = rand_sparse(90, 300, 0.2, 6, empty_cols=(11, 12))
    for k in (1, 5, 20, 40):
        g = sd.sparseKNN(X, Y, k=k, method=method, verbose=False)
        o = oracle_lib.knn(X, Y, k=k, method=method, nthreads=NT)
        same((g.idx, g.dist), o, f"{method} k={k}")
        assert g.self_search is False


def test_large_k_uses_exact_path(oracle_lib, gpu_ctx):
    X = rand_sparse(40, 300, 0.2, 7)
    for method in ("binary", "cosine"):
        g = sd.sparseKNN(X, k=120, method=method, verbose=False)
        o = oracle_lib.knn(X, k=120, method=method, nthreads=NT)
        same((g.idx, g.dist), o, method)


@pytest.mark.parametrize("method", ["binary", "cosine"])
def test_knn_c1_shape(oracle_lib, gpu_ctx, method):
    """C1 geometry (2000 rows, d = 0.05, |N(0,1)| values), 1500 columns."""
    X = synth.abs_normal(2000, 1500, 0.05, 11)
    g = sd.sparseKNN(X, k=15, method=method, verbose=False)
    o = oracle_lib.knn(X, k=15, method=method, nthreads=NT)
    same((g.idx, g.dist), o, method)


@pytest.mark.parametrize("method", ["binary", "cosine"])
@pytest.mark.parametrize("dist", [True, False])
def test_knn_count_matrix_shape(oracle_lib, gpu_ctx, method, dist):
    """The count-matrix workload's rows and density (20000 x d = 0.06,
    NB/log1p), 1200 columns so the oracle finishes in seconds."""
    X = synth.nb_log1p(20000, 1200, 0.06, 12)
    for prec in ("auto", "fp32"):
        g = sd.sparseKNN(X, k=20, method=method, dist=dist, precision=prec, verbose=False)
        o = oracle_lib.knn(X, k=20, method=method, dist=dist, nthreads=NT)
        same((g.idx, g.dist), o, f"{method} {prec}")


def test_certificate_counts(gpu_ctx):
    X = synth.nb_log1p(20000, 1200, 0.06, 13)
    gpu_ctx.stats(reset=True)
    sd.sparseKNN(X, k=20, method="cosine", precision="bf16", verbose=False)
    st = gpu_ctx.stats(reset=True)
    assert st["certified_queries"] + st["fallback_queries"] == 1200


def test_cross_equivalence(oracle_lib, gpu_ctx):
    X = rand_sparse(120, 61, 0.06, 8, levels=2, empty_cols=(0, 4, 60))
    Y = rand_sparse(120, 37, 0.06, 9, levels=2, empty_cols=(1, 29))
    for method in ("binary", "cosine", "euclidean", "manhattan"):
        for dist in ([True, False] if method in ("binary", "cosine") else [True]):
            cr = sd.sparseKNNCross(X, Y, k=7, method=method, dist=dist, verbose=False)
            fw = sd.sparseKNN(X, Y, k=7, method=method, dist=dist, verbose=False)
            rv = sd.sparseKNN(Y, X, k=7, method=method, dist=dist, verbose=False)
            same((cr["forward"].idx, cr["forward"].dist), (fw.idx, fw.dist), "fwd")
            same((cr["reverse"].idx, cr["reverse"].dist), (rv.idx, rv.dist), "rev")
            (fi, fd), (ri, rd), _ = oracle_lib.knn_cross(X, Y, k=7, method=method, dist=dist,
                                                          nthreads=NT)
            same((cr["forward"].idx, cr["forward"].dist), (fi, fd), f"{method} oracle fwd")
            same((cr["reverse"].idx, cr["reverse"].dist), (ri, rd), f"{method} oracle rev")


def test_dense_embedding_cross(oracle_lib, gpu_ctx):
    """The dense-embedding cross shape: unit-normalised N(0,1) FP32 columns
    (signed values, so the re-rank certificate uses its general bound), cosine
    similarity, both directions; the forward direction is large enough for the
    threshold pre-pass + filtered pass."""
    E = synth.unit_normal_dense(64, 600, 90).astype(np.float64)
    Y = synth.unit_normal_dense(64, 20000, 91).astype(np.float64)
    gpu_ctx.stats(reset=True)
    gpu_ctx.set_timing(True)  # launches are counted while timing is on
    try:
        cr = sd.sparseKNNCross(E, Y, k=10, method="cosine", dist=False, precision="bf16",
                               verbose=False)
    finally:
        gpu_ctx.set_timing(False)
    st = gpu_ctx.stats(reset=True)
    (fi, fd), (ri, rd), _ = oracle_lib.knn_cross(sp.csc_matrix(E), sp.csc_matrix(Y), k=10,
                                                  method="cosine", dist=False, nthreads=NT)
    same((cr["forward"].idx, cr["forward"].dist), (fi, fd), "dense fwd")
    same((cr["reverse"].idx, cr["reverse"].dist), (ri, rd), "dense rev")
    assert st["filter_launches"] >= 1


def test_filtered_pass_two_dimensional_grid(oracle_lib, gpu_ctx):
    """More (query tile, reference tile) pairs than one grid dimension holds
    (1100 x 1100 tiles of 256 > 2^20): the filtered pass runs on a 2-D grid
    and must still visit every tile.  Per-query results do not depend on the
    other queries, so a seeded sample of queries is checked against the oracle."""
    n = 256 * 1100
    Q = synth.unit_normal_dense(32, n, 92).astype(np.float64)
    R = synth.unit_normal_dense(32, n, 93).astype(np.float64)
    g = sd.sparseKNN(Q, R, k=5, method="cosine", verbose=False)
    rng = np.random.default_rng(94)
    sample = np.sort(rng.choice(n, size=160, replace=False))
    oi, od, _ = oracle_lib.knn(sp.csc_matrix(Q[:, sample]), sp.csc_matrix(R), k=5,
                               method="cosine", nthreads=NT)
    same((g.idx[sample], g.dist[sample]), (oi, od), "2-D grid sample")


def test_autotune_path_choice(oracle_lib, gpu_ctx):
    """sd_autotune measures plausible rates; with precision="auto" a single
    query against a very sparse reference takes the FP64 merge walk and a
    large batch the densified path, and both give the oracle's answer."""
    rates = gpu_ctx.autotune()
    assert rates["calibrated"] == 1
    assert 1e13 < rates["bf16_ops"] < 5e15 and 1e13 < rates["fp4_ops"] < 2e16
    assert 1e9 < rates["merge_steps"] < 1e13
    Y = synth.nb_log1p(30000, 3000, 0.0003, 95)     # about 9 stored rows per column
    X1 = synth.nb_log1p(30000, 1, 0.0003, 96)
    Xb = synth.nb_log1p(30000, 3000, 0.02, 97)
    Yb = synth.nb_log1p(30000, 3000, 0.02, 98)
    for method in ("binary", "cosine"):
        gpu_ctx.stats(reset=True)
        g = sd.sparseKNN(X1, Y, k=5, method=method, verbose=False)
        st = gpu_ctx.stats(reset=True)
        assert st["exact_calls"] == 1 and st["fused_calls"] == 0, method
        oi, od, _ = oracle_lib.knn(X1, Y, k=5, method=method, nthreads=NT)
        same((g.idx, g.dist), (oi, od), f"{method} tiny batch")
        g = sd.sparseKNN(Xb, Yb, k=5, method=method, verbose=False)
        st = gpu_ctx.stats(reset=True)
        assert st["fused_calls"] == 1 and st["exact_calls"] == 0, method
        e = sd.sparseKNN(Xb, Yb, k=5, method=method, precision="exact", verbose=False)
        same((g.idx, g.dist), (e.idx, e.dist), f"{method} batch vs exact path")


def test_topk_block_parity(oracle_lib, gpu_ctx):
    rng = np.random.default_rng(21)
    D = rng.integers(0, 6, size=(300, 40)).astype(float)
    D[rng.random(D.shape) < 0.05] = np.nan
    D[3, 2] = np.inf
    D[5, 2] = -0.0
    self_row = rng.integers(0, 301, size=40).astype(np.int32)
    for k in (1, 7, 64, 300):
        for dec in (False, True):
            g = sd.topKBlock(D, k, dec, self_row)
            o = oracle_lib.topk_block(D, k, dec, self_row)
            same(g, o, f"k={k} dec={dec}")


def test_topk_rows_accum_parity(oracle_lib, gpu_ctx):
    rng = np.random.default_rng(22)
    D = rng.integers(0, 4, size=(50, 90)).astype(float)
    k = 5
    self_col = rng.integers(0, 91, size=50).astype(np.int32)
    for dec in (False, True):
        gi = np.full((50, k), NA_INT, np.int32, order="F")
        gv = np.full((50, k), np.nan, order="F")
        oi, ov = gi.copy(order="F"), gv.copy(order="F")
        for c0 in range(0, 90, 13):
            blk = np.asfortranarray(D[:, c0:c0 + 13])
            sd.topKRowsAccum(blk, k, dec, gi, gv, c0, self_col)
            oracle_lib.topk_rows_accum(blk, k, dec, oi, ov, c0, self_col)
        same((gi, gv), (oi, ov), f"dec={dec}")


def test_fixture_topk_na(gpu_ctx):
    D = np.array([[1.0], [np.nan], [np.inf]])
    idx, val = sd.topKBlock(D, 3, False, [0])
    assert idx[0].tolist() == [1, NA_INT, NA_INT]
    assert val[0, 0] == 1.0
    assert sd.is_na_real(val[0, 1:]).all()
    for case in FIX["topk_validation"]["cases"]:
        with pytest.raises((ValueError, sd.SparseDistError), match=case["error"]):
            sd.topKBlock(np.random.rand(3, 2), case["k"], False, case["self_row"])


def test_fixture_ties_and_degenerate(gpu_ctx):
    f = FIX["ties_lower_index"]
    X = sp.csc_matrix(np.array(f["X"], dtype=float).T)
    g = sd.sparseKNN(X, k=3, method="euclidean", verbose=False)
    assert g.idx[f["row"] - 1].tolist() == f["expect_idx_row"]
    f = FIX["massive_jaccard_ties"]
    Xd = np.zeros((f["nrow"], f["ncol_x"]))
    Xd[: f["ones_rows"]] = 1
    Yd = np.zeros((f["nrow"], f["ncol_y"]))
    Yd[: f["ones_rows"]] = 1
    for bs in f["block_sizes"]:
        cr = sd.sparseKNNCross(sp.csc_matrix(Xd), sp.csc_matrix(Yd), k=6, block_size=bs,
                               verbose=False)
        assert (cr["forward"].idx == np.arange(1, 7)).all()
        assert (cr["reverse"].idx == np.arange(1, 7)).all()
        assert (cr["forward"].dist == 0).all()
    f = FIX["signed_cancel_not_empty"]
    X = sp.csc_matrix(np.array(f["X"], dtype=float).T)
    for m in f["methods"]:
        g = sd.sparseKNN(X, k=2, method=m, verbose=False)
        assert (g.idx[0] != NA_INT).all() and (g.idx[1] == NA_INT).all()
        assert not (g.idx == 2).any() and (g.idx[1:] == 1).any()
    f = FIX["cross_degenerate"]
    clean = sp.csc_matrix(np.array(f["clean"], dtype=float).T)
    holed = sp.csc_matrix(np.array(f["holed"], dtype=float).T)
    a = sd.sparseKNN(holed, clean, k=2, verbose=False).idx
    assert (a[1] == NA_INT).all() and (a[[0, 2]] != NA_INT).all()
    b = sd.sparseKNN(clean, holed, k=2, verbose=False).idx
    assert not (b == NA_INT).any() and not (b == 2).any()
    with pytest.warns(sd.SparseDistWarning, match="exceeds"):
        d = sd.sparseKNN(clean, holed, k=3, verbose=False)
    assert d.idx.shape[1] == 2


def test_fixture_pearson_and_cosine_conventions(gpu_ctx):
    from test_oracle import check_cosine_conventions
    check_cosine_conventions(lambda m, A, B, d: sd.sparseDist2(A, B, method=m, dist=d))
    f = FIX["pearson_constant_self"]
    X = sp.csc_matrix(np.array(f["X"], dtype=float).T)
    s = sd.sparseKNN(X, k=2, method="pearson", dist=False, include_self=True, verbose=False)
    at = list(s.idx[0]).index(1)
    assert s.dist[0, at] == 1.0
    d = sd.sparseKNN(X, k=2, method="pearson", dist=True, include_self=True, verbose=False)
    assert d.idx[:, 0].tolist() == [1, 2] and d.dist[:, 0].tolist() == [0.0, 0.0]
    X = sp.csc_matrix(np.array([[np.inf, 1.0], [0.0, 0.0]]))
    with pytest.raises(sd.SparseDistError, match="finite"):
        sd.sparseKNN(X, k=1, method="cosine", verbose=False)


def test_block_size_and_batch_invariance(gpu_ctx):
    X = synth.nb_log1p(3000, 700, 0.05, 14)
    base = sd.sparseKNN(X, k=10, method="cosine", block_size=700, verbose=False)
    for bs in (1, 61, 256, 10000):
        g = sd.sparseKNN(X, k=10, method="cosine", block_size=bs, verbose=False)
        same((g.idx, g.dist), (base.idx, base.dist), f"bs={bs}")
    # a query's result does not depend on which other queries share the launch
    Y = sd.DeviceMatrix(X)
    sub = X[:, 100:110]
    g = sd.sparseKNN(sub, Y, k=10, method="cosine", verbose=False)
    full = sd.sparseKNN(X, Y, k=10, method="cosine", verbose=False)
    same((g.idx, g.dist), (full.idx[100:110], full.dist[100:110]), "subset")


def test_certificate_escalation_and_fallback(oracle_lib, gpu_ctx):
    """A narrow first list (margin 1) forces the wider second pass, and
    near-duplicate columns (cosines within BF16 rounding of each other) force
    the exact FP64 fallback; every path must still match the oracle."""
    X = synth.nb_log1p(3000, 600, 0.05, 31)
    gpu_ctx.stats(reset=True)
    g = sd.sparseKNN(X, k=6, method="cosine", margin=1, precision="bf16", verbose=False)
    st = gpu_ctx.stats(reset=True)
    o = oracle_lib.knn(X, k=6, method="cosine", nthreads=NT)
    same((g.idx, g.dist), o, "margin=1")
    assert st["escalated_queries"] > 0
    rng = np.random.default_rng(32)
    base = synth.abs_normal(500, 1, 0.2, 33).toarray()
    cols = [base * (1 + 1e-7 * rng.standard_normal(base.shape)) for _ in range(150)]
    cols.append(synth.abs_normal(500, 100, 0.2, 34).toarray())
    Xd = sp.csc_matrix(np.hstack(cols))
    gpu_ctx.stats(reset=True)
    g = sd.sparseKNN(Xd, k=10, method="cosine", precision="bf16", verbose=False)
    st = gpu_ctx.stats(reset=True)
    o = oracle_lib.knn(Xd, k=10, method="cosine", nthreads=NT)
    same((g.idx, g.dist), o, "near duplicates")
    assert st["fallback_queries"] > 0


@pytest.mark.parametrize("method", ["binary", "cosine", "euclidean", "pearson"])
def test_two_pass_filter_path(oracle_lib, gpu_ctx, method):
    """Enough reference columns (>= 128 tiles of 128) for the threshold
    pre-pass + filtered full pass; queries are a subset so the oracle stays
    fast, and a heavily tied block exercises the buffer-overflow path."""
    Y = synth.nb_log1p(600, 20000, 0.05, 41)
    Yd = Y.tolil()
    # 4700 copies of column 0: more than any candidate buffer holds (4096)
    for c in range(3000, 7700):
        Yd[:, c] = Y[:, 0]
    Y = Yd.tocsc()
    X = Y[:, :300]
    gpu_ctx.stats(reset=True)
    g = sd.sparseKNN(X, Y, k=20, method=method, precision="bf16", verbose=False)
    st = gpu_ctx.stats(reset=True)
    o = oracle_lib.knn(X, Y, k=20, method=method, nthreads=NT)
    same((g.idx, g.dist), o, method)
    assert st["overflow_queries"] >= 1


DEGENERATE_COLS = (17, 12001, 19990)  # first tile, mid-range, the partial last tile


@pytest.mark.parametrize("method,data", [("binary", "counts"), ("cosine", "counts"),
                                         ("euclidean", "counts"), ("euclidean", "signed"),
                                         ("pearson", "counts"), ("pearson", "signed")])
@pytest.mark.parametrize("include_self", [False, True])
def test_symmetric_self_filter(oracle_lib, gpu_ctx, method, data, include_self):
    """A self-search over 20000 columns at once scores each unordered pair once
    (tiles with reference tile >= query tile, each value screened for both
    columns as the query, euclidean and pearson with the reference column's
    own key weights).  Signed euclidean / pearson data takes the split layout,
    whose mirrored values may differ from the threshold pre-pass's in the last
    bit: when that leaves a query fewer than L candidates (two pearson
    queries here, before every short buffer was sent to the full pass, came
    out with 17 and 18 neighbours; the whole-result check caught it) the query
    is searched again in full.  A heavily tied block overflows candidate
    buffers from both directions; the last tile is partial.  Sampled rows,
    including those, must match the oracle, and every row must match a cross
    search of the same columns."""
    n, k = 20000, 20
    if data == "counts":
        Y = synth.nb_log1p(600, n, 0.05, 43)
    else:
        Y = rand_sparse(600, n, 0.05, 44, signed=True)
    Yd = Y.tolil()
    for c in range(3000, 7700):
        Yd[:, c] = Y[:, 0]
    for c in DEGENERATE_COLS:  # empty: NA rows, never candidates
        Yd[:, c] = 0
    Y = Yd.tocsc()
    Y.eliminate_zeros()
    gpu_ctx.stats(reset=True)
    gpu_ctx.set_timing(True)  # launches are counted while timing is on
    try:
        g = sd.sparseKNN(Y, k=k, method=method, include_self=include_self, precision="bf16",
                         verbose=False)
    finally:
        gpu_ctx.set_timing(False)
    st = gpu_ctx.stats(reset=True)
    assert st["filter_launches"] == 1 and st["filter_pairs"] == n * (n + 1) // 2, st
    assert st["overflow_queries"] >= 1
    rng = np.random.default_rng(7)
    qcols = sorted(set(rng.choice(n, 600, replace=False).tolist()) |
                   {0, 1, 255, 256, 2999, 3000, 5000, 7699, 7700, 19967, 19968, 19999} |
                   set(DEGENERATE_COLS))
    o = oracle_lib.knn(Y[:, qcols], Y, k=k + 1, method=method, nthreads=NT)
    want = self_rows_from_cross(o[0], o[1], qcols, k, include_self)
    same((g.idx[qcols], g.dist[qcols]), want, f"{method} include_self={include_self}")
    # every row: a cross search against the same columns runs the full grid
    # of tiles; with the self rules applied its rows equal the symmetric ones
    c = sd.sparseKNN(Y, Y.copy(), k=k + 1, method=method, precision="bf16", verbose=False)
    allq = list(range(n))
    same((g.idx, g.dist), self_rows_from_cross(c.idx, c.dist, allq, k, include_self),
         f"{method} include_self={include_self}: symmetric vs full grid")


@pytest.mark.parametrize("prec", ["bf16", "fp32"])
def test_euclidean_count_matrix_certified(oracle_lib, gpu_ctx, prec):
    """Euclidean on the count-matrix rows and density: split-BF16 (or FP32)
    slab with per-pair lower-bound keys, FP64 re-rank.  Nearest distances on
    such data sit within a fraction of a percent of each other, so this also
    checks that the bound is tight enough to certify nearly every query."""
    X = synth.nb_log1p(20000, 1200, 0.06, 12)
    gpu_ctx.stats(reset=True)
    g = sd.sparseKNN(X, k=20, method="euclidean", precision=prec, verbose=False)
    st = gpu_ctx.stats(reset=True)
    o = oracle_lib.knn(X, k=20, method="euclidean", nthreads=NT)
    same((g.idx, g.dist), o, f"euclidean {prec}")
    assert st["fused_calls"] == 1
    # at most 5% need the wider second list, at most 1% the exact fallback
    assert st["escalated_queries"] <= 60 and st["fallback_queries"] <= 12, st


def test_euclidean_signed_and_near_duplicates(oracle_lib, gpu_ctx):
    """Signed values take the general bound (sum |q_i r_i| <= |q||r|); columns
    equal up to 1e-7 relative noise cannot be ordered by any slab and go to the
    exact fallback.  Both must match the oracle."""
    X = rand_sparse(300, 500, 0.1, 51, signed=True, empty_cols=(3,))
    Y = rand_sparse(300, 900, 0.1, 52, signed=True, empty_cols=(8,))
    for include_self in (False, True):
        g = sd.sparseKNN(X, k=7, method="euclidean", include_self=include_self,
                         precision="bf16", verbose=False)
        o = oracle_lib.knn(X, k=7, method="euclidean", include_self=include_self, nthreads=NT)
        same((g.idx, g.dist), o, f"signed self include_self={include_self}")
    g = sd.sparseKNN(X, Y, k=7, method="euclidean", precision="bf16", verbose=False)
    o = oracle_lib.knn(X, Y, k=7, method="euclidean", nthreads=NT)
    same((g.idx, g.dist), o, "signed cross")
    rng = np.random.default_rng(53)
    base = synth.abs_normal(400, 1, 0.2, 54).toarray()
    cols = [base * (1 + 1e-7 * rng.standard_normal(base.shape)) for _ in range(120)]
    cols.append(synth.abs_normal(400, 100, 0.2, 55).toarray())
    Xd = sp.csc_matrix(np.hstack(cols))
    gpu_ctx.stats(reset=True)
    dX = sd.DeviceMatrix(Xd)
    g = sd.sparseKNN(dX, k=10, method="euclidean", precision="bf16", verbose=False)
    st = gpu_ctx.stats(reset=True)
    o = oracle_lib.knn(Xd, k=10, method="euclidean", nthreads=NT)
    same((g.idx, g.dist), o, "near duplicates")
    assert st["fallback_queries"] > 0
    # more than 1% fell back on plain BF16 values, so this reference now uses
    # the split layout; the result is the same
    g = sd.sparseKNN(dX, k=10, method="euclidean", precision="bf16", verbose=False)
    same((g.idx, g.dist), o, "near duplicates, split layout")


def test_euclidean_out_of_range_values_take_exact_path(oracle_lib, gpu_ctx):
    """Values outside [1e-15, 1e15] (FP32 squares would overflow or BF16 lo
    parts become subnormal) are searched on the FP64 merge walk."""
    X = rand_sparse(60, 80, 0.3, 56)
    X.data[::7] *= 1e17
    gpu_ctx.stats(reset=True)
    g = sd.sparseKNN(X, k=5, method="euclidean", precision="bf16", verbose=False)
    st = gpu_ctx.stats(reset=True)
    o = oracle_lib.knn(X, k=5, method="euclidean", nthreads=NT)
    same((g.idx, g.dist), o, "large values")
    assert st["fused_calls"] == 0 and st["exact_calls"] == 1


@pytest.mark.parametrize("prec", ["bf16", "fp32"])
@pytest.mark.parametrize("dist", [True, False])
def test_pearson_count_matrix_certified(oracle_lib, gpu_ctx, prec, dist):
    """Pearson on the count-matrix rows and density: raw-value slab (split
    BF16 or FP32), per-pair upper-bound keys of the correlation, FP64
    re-rank in the library's centred definition."""
    X = synth.nb_log1p(20000, 1200, 0.06, 12)
    gpu_ctx.stats(reset=True)
    g = sd.sparseKNN(X, k=20, method="pearson", dist=dist, precision=prec, verbose=False)
    st = gpu_ctx.stats(reset=True)
    o = oracle_lib.knn(X, k=20, method="pearson", dist=dist, nthreads=NT)
    same((g.idx, g.dist), o, f"pearson {prec} dist={dist}")
    assert st["fused_calls"] == 1
    assert st["escalated_queries"] <= 60 and st["fallback_queries"] <= 12, st


def test_pearson_signed_constant_and_self(oracle_lib, gpu_ctx):
    """Signed values (general bound), constant columns (correlation 0 by
    convention), self restore and cross search on the fused pearson path."""
    X = rand_sparse(200, 400, 0.15, 61, signed=True, empty_cols=(5,))
    Xd = X.toarray()
    Xd[:, 7] = 2.5  # a constant column: every stored row holds the same value
    Xd[:, 9] = 0.0
    Xd[3, 9] = 1.0  # a single stored value
    X = sp.csc_matrix(Xd)
    Y = rand_sparse(200, 700, 0.15, 62, signed=True)
    for dist in (True, False):
        for include_self in (False, True):
            for drop in (True, False):
                lab = f"dist={dist} self={include_self} drop={drop}"
                g = sd.sparseKNN(X, k=8, method="pearson", dist=dist, include_self=include_self,
                                 drop_degenerate=drop, precision="bf16", verbose=False)
                o = oracle_lib.knn(X, k=8, method="pearson", dist=dist,
                                   include_self=include_self, drop_degenerate=drop, nthreads=NT)
                same((g.idx, g.dist), o, lab)
        g = sd.sparseKNN(X, Y, k=8, method="pearson", dist=dist, precision="bf16", verbose=False)
        o = oracle_lib.knn(X, Y, k=8, method="pearson", dist=dist, nthreads=NT)
        same((g.idx, g.dist), o, f"cross dist={dist}")


@pytest.mark.parametrize("method,dist,drop", [("binary", True, True), ("cosine", True, True),
                                              ("euclidean", True, True), ("pearson", True, True),
                                              ("cosine", False, False)])
def test_cross_search_one_filtered_pass(oracle_lib, gpu_ctx, method, dist, drop):
    """sparseKNNCross with both sides large enough for a filtered pass: one
    pass over the X x Y tiles feeds both directions (the reverse search's
    buffers through the mirror, with its own thresholds).  A tied block in Y
    overflows forward buffers and empty columns on both sides test the
    degenerate rules; euclidean and pearson keys use the reverse query's own
    weights.  Both directions must equal two independent sparseKNN calls on
    every row, and sampled rows the oracle."""
    nx, ny, k = 17000, 20000, 20
    X = synth.nb_log1p(600, nx, 0.05, 48).tolil()
    Y = synth.nb_log1p(600, ny, 0.05, 49).tolil()
    for c in range(2000, 6500):
        Y[:, c] = X[:, 5]
    for c in (3, 16999):
        X[:, c] = 0
    for c in (10, 19999):
        Y[:, c] = 0
    X, Y = X.tocsc(), Y.tocsc()
    X.eliminate_zeros()
    Y.eliminate_zeros()
    gpu_ctx.stats(reset=True)
    gpu_ctx.set_timing(True)  # launches are counted while timing is on
    flags = dict(method=method, dist=dist, drop_degenerate=drop)
    try:
        cr = sd.sparseKNNCross(X, Y, k=k, precision="bf16", verbose=False, **flags)
    finally:
        gpu_ctx.set_timing(False)
    st = gpu_ctx.stats(reset=True)
    assert st["filter_launches"] == 1 and st["filter_pairs"] == nx * ny, st
    fw = sd.sparseKNN(X, Y, k=k, precision="bf16", verbose=False, **flags)
    rv = sd.sparseKNN(Y, X, k=k, precision="bf16", verbose=False, **flags)
    same((cr["forward"].idx, cr["forward"].dist), (fw.idx, fw.dist), "forward")
    same((cr["reverse"].idx, cr["reverse"].dist), (rv.idx, rv.dist), "reverse")
    rng = np.random.default_rng(12)
    sf = sorted(set(rng.choice(nx, 200, replace=False).tolist()) | {3, 5, 16999})
    sr = sorted(set(rng.choice(ny, 200, replace=False).tolist()) | {10, 2000, 6499, 19999})
    of = oracle_lib.knn(X[:, sf], Y, k=k, nthreads=NT, **flags)
    orv = oracle_lib.knn(Y[:, sr], X, k=k, nthreads=NT, **flags)
    same((cr["forward"].idx[sf], cr["forward"].dist[sf]), of, "forward vs oracle")
    same((cr["reverse"].idx[sr], cr["reverse"].dist[sr]), orv, "reverse vs oracle")


@pytest.mark.parametrize("method", ["cosine", "euclidean"])
def test_symmetric_self_filter_keeps_degenerate(oracle_lib, gpu_ctx, method):
    """drop_degenerate = FALSE: empty columns are ordinary candidates and
    queries of the symmetric pass (an empty column is at distance |q| in
    euclidean; cosine takes the reference's 0 / 1 convention)."""
    n, k = 20000, 20
    Y = synth.nb_log1p(600, n, 0.05, 47).tolil()
    for c in DEGENERATE_COLS:
        Y[:, c] = 0
    Y = Y.tocsc()
    Y.eliminate_zeros()
    gpu_ctx.stats(reset=True)
    g = sd.sparseKNN(Y, k=k, method=method, drop_degenerate=False, precision="bf16",
                     verbose=False)
    st = gpu_ctx.stats(reset=True)
    assert st["filter_pairs"] == n * (n + 1) // 2, st
    rng = np.random.default_rng(11)
    qcols = sorted(set(rng.choice(n, 300, replace=False).tolist()) | set(DEGENERATE_COLS))
    o = oracle_lib.knn(Y[:, qcols], Y, k=k + 1, method=method, drop_degenerate=False,
                       nthreads=NT)
    same((g.idx[qcols], g.dist[qcols]), self_rows_from_cross(o[0], o[1], qcols, k, False),
         method)
    c = sd.sparseKNN(Y, Y.copy(), k=k + 1, method=method, drop_degenerate=False,
                     precision="bf16", verbose=False)
    same((g.idx, g.dist), self_rows_from_cross(c.idx, c.dist, list(range(n)), k, False),
         f"{method}: symmetric vs full grid")


@pytest.mark.parametrize("method", ["cosine", "euclidean"])
def test_escalation_lists_from_candidate_buffers(oracle_lib, gpu_ctx, method):
    """A narrow re-rank margin (L = 24 at k = 20) makes the first certificate
    fail often; the wider lists (L = 64) of the failed queries then come from
    the filtered pass's candidate buffers (about 770 entries per query here,
    every reference at or above the query's threshold) instead of a second
    slab pass.  Sampled rows must match the oracle."""
    n, k = 20000, 20
    Y = synth.nb_log1p(2000, n, 0.05, 45)
    gpu_ctx.stats(reset=True)
    g = sd.sparseKNN(Y, k=k, method=method, precision="bf16", margin=1, verbose=False)
    st = gpu_ctx.stats(reset=True)
    if method == "cosine":
        assert st["escalated_queries"] > 0, st
    if st["escalated_queries"] > 0:
        assert st["buffer_escalations"] == 1, st
    rng = np.random.default_rng(9)
    qcols = sorted(rng.choice(n, 400, replace=False).tolist())
    o = oracle_lib.knn(Y[:, qcols], Y, k=k + 1, method=method, nthreads=NT)
    want = self_rows_from_cross(o[0], o[1], qcols, k, False)
    same((g.idx[qcols], g.dist[qcols]), want, method)


def test_margin_adapts_per_reference(gpu_ctx):
    """With precision = "auto" the re-rank margin adapts per reference: after a
    filtered pass (candidate buffers serve the wider lists) more than 15% of
    queries escalating doubles it (at most 48), fewer than 1% lowers it by 4
    (at least 8).  Three searches on one resident reference follow that rule
    from the default 16, and give identical results."""
    Y = synth.nb_log1p(2000, 20000, 0.05, 46)
    dY = sd.DeviceMatrix(Y, gpu_ctx)
    margin, prev = 16, None
    for it in range(3):
        gpu_ctx.stats(reset=True)
        g = sd.sparseKNN(dY, k=20, method="cosine", verbose=False, context=gpu_ctx)
        st = gpu_ctx.stats(reset=True)
        assert st["fused_calls"] == 1 and st["last_margin"] == margin, (it, st)
        if prev is not None:
            assert np.array_equal(g.idx, prev.idx)
            assert np.array_equal(bits(g.dist), bits(prev.dist))
        prev = g
        # first-pass failures: escalated while L < 64, else sent to the exact path
        esc = st["escalated_queries"] if 20 + margin <= 48 else st["fallback_queries"]
        if esc * 100 > 15 * 20000:
            margin = min(2 * margin, 48)
        elif esc * 100 < 20000:
            margin = max(margin - 4, 8)


@pytest.mark.parametrize("method,copies", [("cosine", 29), ("binary", 16)])
@pytest.mark.parametrize("nref", [40000, 70000])
def test_sparse_sample_threshold_redo(oracle_lib, gpu_ctx, method, copies, nref):
    """Large references take thresholds from a sparse sample of tiles at a
    reduced list rank: every 64th tile at about half the rank from 256 tiles
    of 128 (L = 40 -> 24 for cosine, 24 -> 16 for binary at k = 20), every
    128th at about a quarter from 512 tiles (40 -> 16, 24 -> 8).  Such a
    threshold can exceed the true L-th key: here copies of the first query
    fill the first sampled tile beyond that rank, so the sample's key at that
    rank is the top score and fewer than L candidates pass; the query is
    searched again in full and every result still matches the oracle.  The
    rows are long enough (5000: 2.5 KB of FP4 patterns per column) for these
    sampling tiers; rows of 2 KB or less take a denser sample instead."""
    Y = synth.nb_log1p(5000, nref, 0.05, 43)
    Yd = Y.tolil()
    base = Y[:, 100]
    for c in range(copies):
        Yd[:, c] = base
    Y = Yd.tocsc()
    X = sp.hstack([base, Y[:, 5000:5060]]).tocsc()
    gpu_ctx.stats(reset=True)
    g = sd.sparseKNN(X, Y, k=20, method=method, precision="bf16", verbose=False)
    st = gpu_ctx.stats(reset=True)
    o = oracle_lib.knn(X, Y, k=20, method=method, nthreads=NT)
    same((g.idx, g.dist), o, method)
    assert st["overflow_queries"] >= 1


def test_escalation_with_incomplete_buffers(oracle_lib, gpu_ctx):
    """Short rows take a dense threshold sample at a low rank (every 64th
    tile of 128 references, rank 16 for L = 40).  Sixteen near-copies of
    query column 5000 in the first sampled tile put its threshold at the
    copies' level, and thirty more copies in an unsampled tile leave its
    buffer with 47 entries: at least L but fewer than 64.  Copies equal to
    within BF16 rounding make its certificate fail, so its wider list comes
    from a list-mode pass of its own, written back at its position in the
    list of failed queries.  That query is the 151st query column while few
    others fail, so its position and its column differ.  Every row must
    match the oracle."""
    rng = np.random.default_rng(46)
    Y = synth.nb_log1p(300, 20000, 0.05, 47).tolil()
    base = Y[:, 5000].toarray().ravel()
    for c in list(range(16)) + list(range(128, 158)):
        Y[:, c] = (base * (1 + 1e-7 * rng.standard_normal(base.size)))[:, None]
    Y = Y.tocsc()
    qcols = list(range(200, 350)) + [5000]
    X = Y[:, qcols]
    gpu_ctx.stats(reset=True)
    g = sd.sparseKNN(X, Y, k=20, method="cosine", precision="bf16", verbose=False)
    st = gpu_ctx.stats(reset=True)
    o = oracle_lib.knn(X, Y, k=20, method="cosine", nthreads=NT)
    same((g.idx, g.dist), o, "incomplete escalation buffers")
    assert st["buffer_escalations"] == 1 and st["escalation_list_queries"] >= 1, st
    assert st["escalated_queries"] < 100, st


@pytest.mark.parametrize("method", ["binary", "cosine", "euclidean", "pearson"])
def test_query_blocks_self_search(oracle_lib, gpu_ctx, method):
    """Under a small buffer budget a 40,000-column self-search runs its query
    columns in three blocks (18,432 + 18,432 + 3,136), each a whole search
    against every reference column that finds the query's own column at the
    block's offset.  Every row must equal the one-launch (symmetric) search,
    and sampled rows -- around the block edges, in the partial last block and
    a degenerate column in the second block -- the oracle, with and without
    the query's own column."""
    n = 40000
    Y = synth.nb_log1p(200, n, 0.1, 120).tolil()
    Y[:, 20001] = 0
    Y = Y.tocsc()
    Y.eliminate_zeros()
    k = 10
    rng = np.random.default_rng(121)
    qcols = sorted(set(rng.choice(n, 200, replace=False).tolist()) |
                   {0, 18431, 18432, 20001, 36863, 36864, 39999})
    for include_self in (False, True):
        lab = f"{method} include_self={include_self}"
        gpu_ctx.set_buffer_budget(0)
        gpu_ctx.stats(reset=True)
        one = sd.sparseKNN(Y, k=k, method=method, include_self=include_self, precision="bf16",
                           verbose=False)
        assert gpu_ctx.stats(reset=True)["query_blocks"] == 1, lab
        gpu_ctx.set_buffer_budget(1 << 30)
        try:
            blk = sd.sparseKNN(Y, k=k, method=method, include_self=include_self,
                               precision="bf16", verbose=False)
        finally:
            gpu_ctx.set_buffer_budget(0)
        st = gpu_ctx.stats(reset=True)
        assert st["query_blocks"] == 3, (lab, st)
        same((blk.idx, blk.dist), (one.idx, one.dist), f"{lab}: blocks vs one launch")
        o = oracle_lib.knn(Y[:, qcols], Y, k=k + 1, method=method, nthreads=NT)
        same((blk.idx[qcols], blk.dist[qcols]),
             self_rows_from_cross(o[0], o[1], qcols, k, include_self), f"{lab}: oracle")


def test_query_blocks_cross_search(oracle_lib, gpu_ctx):
    """A cross pair whose searches need blocks gives up the shared filtered
    pass (the reverse search's probe stands down) and runs each direction in
    blocks; both directions must equal the unblocked cross search."""
    X = synth.nb_log1p(200, 20000, 0.1, 122)
    Y = synth.nb_log1p(200, 18000, 0.1, 123)
    gpu_ctx.stats(reset=True)
    one = sd.sparseKNNCross(X, Y, k=8, method="cosine", precision="bf16", verbose=False)
    assert gpu_ctx.stats(reset=True)["query_blocks"] == 2
    gpu_ctx.set_buffer_budget(1 << 29)
    try:
        blk = sd.sparseKNNCross(X, Y, k=8, method="cosine", precision="bf16", verbose=False)
    finally:
        gpu_ctx.set_buffer_budget(0)
    st = gpu_ctx.stats(reset=True)
    assert st["query_blocks"] >= 4, st
    for d in ("forward", "reverse"):
        same((blk[d].idx, blk[d].dist), (one[d].idx, one[d].dist), f"blocked cross {d}")
    rng = np.random.default_rng(124)
    qs = np.sort(rng.choice(20000, 120, replace=False))
    oi, od, _ = oracle_lib.knn(X[:, qs], Y, k=8, method="cosine", nthreads=NT)
    same((blk["forward"].idx[qs], blk["forward"].dist[qs]), (oi, od), "blocked cross oracle")


def test_sample_tuner_explores_then_keeps_fastest(gpu_ctx):
    """A class of large searches times each threshold-sample candidate once
    (65,536-column self-search: every candidate has at least four sampled
    tiles; a search that allocated device memory is not timed), then runs
    the fastest -- the fixed rule's choice unless another is 3% faster; the
    results never change.  With tuning off the fixed rule applies (rows of
    400 bytes: every 64th tile at rank ~L/4)."""
    Y = synth.nb_log1p(200, 65536, 0.1, 130)
    sd.tune_samples(reset=True)
    strides = []

    def search():
        gpu_ctx.stats(reset=True)
        g = sd.sparseKNN(Y, k=10, method="cosine", precision="bf16", margin=16, verbose=False)
        st = gpu_ctx.stats(reset=True)
        strides.append((st["sample_stride"], st["sample_rank"]))
        return g

    try:
        first = search()
        for run in range(1, 8):
            g = search()
            same((g.idx, g.dist), (first.idx, first.dist), f"exploring run {run}")
            tab = sd.tune_samples()
            if tab and min(tab[0]["runs"]) >= 1:
                break
        assert len(tab) == 1, tab
        e = tab[0]
        assert e["symmetric"] and e["runs"] == [1, 1, 1, 1], e
        assert set(strides) == set(e["candidates"]), (strides, e)
        rule = 3  # rows of 400 bytes
        t = e["ns_per_mpair"]
        best = min(range(4), key=lambda c: t[c])
        if best != rule and t[best] > 0.97 * t[rule]:
            best = rule
        g = search()
        same((g.idx, g.dist), (first.idx, first.dist), "tuned run")
        assert strides[-1] == e["candidates"][best], (strides, e)
        assert sd.tune_samples()[0]["runs"][best] == 2
        gpu_ctx.set_tuning(False)
        search()
        assert strides[-1] == e["candidates"][rule], strides
    finally:
        gpu_ctx.set_tuning(True)
        sd.tune_samples(reset=True)
