"""GPU checks of the library's edges: cancellation (reference
Progress::check_abort, src/sparseDist.cpp:177-179), lists longer than one
top-k pass holds (src/topK.cpp:67-149 has no upper bound on k), launches
over more items than one grid dimension counts, the Jensen-Shannon method
(src/sparseDist.cpp:1074-1115) and every specialisation of the SNN weight
kernel (src/snn.cpp:198-250).  Every call goes through the C ABI."""
import ctypes as C
import os
import threading
import time

import numpy as np
import pytest
import scipy.sparse as sp

import sparsedist_amd as sd
from sparsedist_amd import _lib, synth
from sparsedist_amd.api import _params
from conftest import bits

pytestmark = pytest.mark.gpu
NT = min(16, os.cpu_count() or 1)
NA_INT = np.iinfo(np.int32).min


def rand_sparse(nrow, ncol, density, seed, signed=False):
    rng = np.random.default_rng(seed)
    M = sp.random(nrow, ncol, density=density, random_state=rng, format="csc")
    M.data = rng.standard_normal(M.nnz) if signed else np.abs(rng.standard_normal(M.nnz)) + 0.01
    M.sort_indices()
    return M


def same(gpu, ora, label=""):
    gi, gd = gpu
    oi, od = ora[0], ora[1]
    assert gi.shape == oi.shape, label
    bad = np.argwhere(gi != oi)
    assert bad.size == 0, f"{label}: idx differs at {bad[:5].tolist()}"
    db = np.argwhere(bits(gd) != bits(od))
    assert db.size == 0, f"{label}: dist differs at {db[:5].tolist()}"


def raw_knn(ctx, dX, dY, k, method, precision, idx, dst):
    """sd_knn straight through the C ABI into the caller's buffers."""
    p = _params(k, method, True, False, True, 256, precision, 0)
    keff, warn = C.c_int32(0), C.c_int32(0)
    return _lib.load().sd_knn(ctx.handle, dX.handle, dY.handle if dY is not None else None,
                              C.byref(p), idx.ctypes.data, dst.ctypes.data, C.byref(keff),
                              C.byref(warn))


# ---- cancellation -------------------------------------------------------------
def test_cancel_before_the_call_applies_to_it(gpu_ctx):
    """A cancel request made before the call starts (a caller's interrupt poll
    can run before its worker thread enters the call) stops that call with
    no output written; the call clears the request, so the next call runs."""
    ctx = sd.Context(0)
    X = synth.nb_log1p(500, 400, 0.05, 601)
    dX = sd.DeviceMatrix(X, ctx)
    idx = np.full(400 * 5, 12345, np.int32)
    dst = np.full(400 * 5, 7.0)
    ctx.cancel()
    rc = raw_knn(ctx, dX, None, 5, "cosine", "auto", idx, dst)
    assert rc == _lib.SD_ECANCELLED
    assert _lib.load().sd_last_error().decode() == "Computation interrupted by the user."
    assert (idx == 12345).all() and (dst == 7.0).all()
    assert raw_knn(ctx, dX, None, 5, "cosine", "auto", idx, dst) == _lib.SD_OK
    want = sd.sparseKNN(dX, k=5, method="cosine", verbose=False, context=ctx)
    assert np.array_equal(idx.reshape((400, 5), order="F"), want.idx)
    dX.free()
    ctx.close()


def test_cancel_from_another_thread(gpu_ctx):
    """sd_cancel from a host thread during a long search (the exact FP64 path,
    polled between slab blocks): the call returns SD_ECANCELLED well before
    the search could finish and leaves the caller's buffers untouched; every
    kernel it launched ran to its normal end."""
    ctx = sd.Context(0)
    n = 12000
    X = synth.nb_log1p(20000, n, 0.06, 602)
    dX = sd.DeviceMatrix(X, ctx)
    idx = np.full(n * 10, 12345, np.int32)
    dst = np.full(n * 10, 7.0)
    timer = threading.Timer(0.3, ctx.cancel)
    t0 = time.perf_counter()
    timer.start()
    rc = raw_knn(ctx, dX, None, 10, "cosine", "exact", idx, dst)
    took = time.perf_counter() - t0
    msg = _lib.load().sd_last_error().decode()
    timer.join()
    assert rc == _lib.SD_ECANCELLED, (rc, msg)
    assert took < 2.0, took  # the full search takes about 4 s
    assert (idx == 12345).all() and (dst == 7.0).all()
    # the context is usable afterwards
    small = synth.nb_log1p(500, 300, 0.05, 603)
    g = sd.sparseKNN(small, k=5, method="cosine", verbose=False, context=ctx)
    want = sd.sparseKNN(small, k=5, method="cosine", verbose=False)
    same((g.idx, g.dist), (want.idx, want.dist), "after a cancel")
    dX.free()
    ctx.close()


def test_cancelled_cross_search_writes_nothing(gpu_ctx):
    """sparseKNNCross keeps the forward results back until the reverse search
    has succeeded, so a cancelled call writes neither direction."""
    ctx = sd.Context(0)
    X = synth.nb_log1p(300, 200, 0.05, 604)
    Y = synth.nb_log1p(300, 250, 0.05, 605)
    dX, dY = sd.DeviceMatrix(X, ctx), sd.DeviceMatrix(Y, ctx)
    p = _params(4, "cosine", True, False, True, 256, "auto", 0)
    fi, ri = np.full(200 * 4, 99, np.int32), np.full(250 * 4, 99, np.int32)
    fd, rd = np.full(200 * 4, 3.0), np.full(250 * 4, 3.0)
    kf, kr, warn = C.c_int32(0), C.c_int32(0), C.c_int32(0)
    ctx.cancel()
    rc = _lib.load().sd_knn_cross(ctx.handle, dX.handle, dY.handle, C.byref(p), fi.ctypes.data,
                                  fd.ctypes.data, C.byref(kf), ri.ctypes.data, rd.ctypes.data,
                                  C.byref(kr), C.byref(warn))
    assert rc == _lib.SD_ECANCELLED
    assert (fi == 99).all() and (ri == 99).all() and (fd == 3.0).all() and (rd == 3.0).all()
    dX.free()
    dY.free()
    ctx.close()


# ---- lists longer than one pass ---------------------------------------------------
@pytest.mark.parametrize("method", ["binary", "cosine"])
def test_knn_k_beyond_one_pass(oracle_lib, gpu_ctx, method):
    """k = 2000 (above the 1792 entries one workgroup list holds): the exact
    path's top-k runs in passes, each after the previous pass's last entry."""
    X = rand_sparse(60, 2600, 0.2, 611)
    g = sd.sparseKNN(X, k=2000, method=method, verbose=False)
    o = oracle_lib.knn(X, k=2000, method=method, nthreads=NT)
    same((g.idx, g.dist), o, method)


def test_topk_block_and_rows_beyond_one_pass(oracle_lib, gpu_ctx):
    rng = np.random.default_rng(612)
    D = rng.integers(0, 50, size=(4000, 6)).astype(float)  # many ties
    D[rng.random(D.shape) < 0.05] = np.nan
    self_row = np.array([0, 1, 4000, 17, 0, 3], np.int32)
    for k in (1800, 3000, 4500):
        for dec in (False, True):
            g = sd.topKBlock(D, k, dec, self_row)
            o = oracle_lib.topk_block(D, k, dec, self_row)
            same(g, o, f"topKBlock k={k} dec={dec}")
    D = rng.integers(0, 30, size=(7, 2500)).astype(float)
    k = 2000
    self_col = np.array([0, 3, 2500, 0, 1, 0, 7], np.int32)
    for dec in (False, True):
        gi = np.full((7, k), NA_INT, np.int32, order="F")
        gv = np.full((7, k), np.nan, order="F")
        oi, ov = gi.copy(order="F"), gv.copy(order="F")
        for c0 in range(0, 2500, 900):
            blk = np.asfortranarray(D[:, c0:c0 + 900])
            sd.topKRowsAccum(blk, k, dec, gi, gv, c0, self_col)
            oracle_lib.topk_rows_accum(blk, k, dec, oi, ov, c0, self_col)
        same((gi, gv), (oi, ov), f"topKRowsAccum k={k} dec={dec}")


def test_topk_block_more_columns_than_one_grid(gpu_ctx):
    """2^24 + 7 columns: more workgroups than one dispatch dimension counts
    (its work-item count is 32 bits); the launch is cut into grids."""
    n = (1 << 24) + 7
    rng = np.random.default_rng(613)
    D = rng.standard_normal((1, n))
    D[0, rng.choice(n, 1000, replace=False)] = np.nan
    idx, val = sd.topKBlock(D, 1, False, np.zeros(n, np.int32))
    fin = np.isfinite(D[0])
    assert (idx[fin, 0] == 1).all() and (idx[~fin, 0] == NA_INT).all()
    assert np.array_equal(bits(val[fin, 0]), bits(D[0, fin]))
    assert sd.is_na_real(val[~fin, 0]).all()


# ---- Jensen-Shannon ----------------------------------------------------------------
def normalised(nrow, ncol, density, seed):
    M = rand_sparse(nrow, ncol, density, seed).toarray()
    s = M.sum(axis=0)
    M[:, s > 0] /= s[s > 0]
    return M


# The device FP64 log and the host C library's may differ in the last bit,
# so JS values are compared at 1e-12 relative (a dropped or mis-signed term
# would be off by orders of magnitude more); idx must match exactly.
JS_RTOL = 1e-12


def test_js_slab(oracle_lib, gpu_ctx):
    A = normalised(80, 30, 0.3, 621)
    B = normalised(80, 25, 0.3, 622)
    A[:, 3] = 0.0
    g = sd.sparseDist2(A, B, method="js")
    o = oracle_lib.slab("js", sp.csc_matrix(A), sp.csc_matrix(B))
    np.testing.assert_allclose(g, o, rtol=JS_RTOL, atol=1e-15)
    assert (g >= 0).all() and (g <= np.sqrt(np.log(2.0)) + 1e-12).all()


@pytest.mark.parametrize("include_self", [False, True])
def test_js_knn(oracle_lib, gpu_ctx, include_self):
    P = normalised(120, 400, 0.2, 623)
    P[:, 7] = 0.0  # degenerate
    g = sd.sparseKNN(P, k=8, method="js", include_self=include_self, verbose=False)
    o = oracle_lib.knn(sp.csc_matrix(P), k=8, method="js", include_self=include_self,
                       nthreads=NT)
    assert np.array_equal(g.idx, o[0])
    ok = g.idx != NA_INT
    np.testing.assert_allclose(g.dist[ok], o[1][ok], rtol=JS_RTOL, atol=1e-15)
    assert sd.is_na_real(g.dist[~ok]).all() and (g.idx[7] == NA_INT).all()
    if include_self:
        live = np.flatnonzero((g.idx != NA_INT).any(axis=1))
        assert (bits(g.dist[live, 0]) == 0).all()  # exact +0.0
        assert (g.idx[live, 0] == live + 1).all()
    Q = normalised(120, 150, 0.2, 624)
    cr = sd.sparseKNNCross(Q, P, k=6, method="js", verbose=False)
    (fi, fd), (ri, rd), _ = oracle_lib.knn_cross(sp.csc_matrix(Q), sp.csc_matrix(P), k=6,
                                                  method="js", nthreads=NT)
    assert np.array_equal(cr["forward"].idx, fi) and np.array_equal(cr["reverse"].idx, ri)


def test_js_fixtures_and_errors(gpu_ctx):
    import json
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fixtures.json")))
    P = np.array(fx["js_standard_metric"]["P"], dtype=float).T
    d = sd.sparseDist2(P, P, method="js")
    assert d[0, 1] == np.sqrt(np.log(2.0)) and d[1, 0] == np.sqrt(np.log(2.0))
    for case in fx["js_invalid_input"]["cases"]:
        M = np.array([[np.nan if v is None else v for v in row] for row in case["M"]])
        with pytest.raises(sd.SparseDistError, match=case["error"]):
            sd.sparseKNN(M, k=1, method="js", verbose=False)
    for m in fx["distance_only_methods"]["methods"]:
        with pytest.raises((ValueError, sd.SparseDistError),
                           match=fx["distance_only_methods"]["error"]):
            sd.sparseKNN(P, k=1, method=m, dist=False, verbose=False)


# ---- SNN weights for every set width --------------------------------------------------
@pytest.mark.parametrize("k", [3, 7, 19, 27, 31, 40])
def test_snn_every_set_width(oracle_lib, gpu_ctx, k):
    """Set rows of 4, 8, 20, 28 and 32 entries take the register kernels;
    k = 40 (44-entry rows) takes the merge walk.  Index matrices with NA,
    repeated and self entries, against the oracle bit for bit."""
    n = 3000
    rng = np.random.default_rng(630 + k)
    idx = rng.integers(1, n + 1, size=(n, k)).astype(np.int32)
    idx[rng.random(idx.shape) < 0.05] = NA_INT
    idx[::17, 0] = np.arange(1, n + 1, 17)  # self listed
    idx[::13, -1] = idx[::13, 0]            # repeats
    idx = np.asfortranarray(idx)
    for include_self in (True, False):
        w = sd.snnJaccard(idx, include_self=include_self)
        ow = oracle_lib.snn_jaccard(idx, include_self=include_self, nthreads=NT)
        assert np.array_equal(bits(w), bits(ow)), (k, include_self)
    for edges in ("knn", "shared"):
        G = sd.sparseSNN(idx, prune=1 / 15, edges=edges)
        p, ri, x = oracle_lib.snn_graph(idx, include_self=True, prune=1 / 15, symmetrise=True,
                                        edges=edges, nthreads=NT)
        assert np.array_equal(G.indptr, p) and np.array_equal(G.indices, ri), (k, edges)
        assert np.array_equal(bits(G.data), bits(x)), (k, edges)
