"""GPU parity of the SNN builder against the CPU oracle and the reference
tests' known answers."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import sparsedist_amd as sd
from conftest import bits
from sparsedist_amd import synth

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "fixtures.json")))
NA_INT = np.iinfo(np.int32).min
NT = min(16, os.cpu_count() or 1)


def idx_array(rows):
    return np.array([[NA_INT if v is None else v for v in r] for r in rows], dtype=np.int32)


def same_graph(g, p, ri, x):
    g = g.tocsc()
    assert np.array_equal(g.indptr.astype(np.int64), p)
    assert np.array_equal(g.indices, ri)
    assert np.array_equal(bits(g.data), bits(x))


def knn_idx(oracle_lib, n, k, seed, method="binary"):
    X = synth.abs_normal(400, n, 0.05, seed)
    idx, _, _ = oracle_lib.knn(X, k=k, method=method, nthreads=NT)
    return idx


def test_snn_weights(oracle_lib, gpu_ctx):
    idx = knn_idx(oracle_lib, 500, 10, 1)
    idx[3, 2] = NA_INT
    idx[4, 1] = idx[4, 0]
    idx[5, 0] = 6
    for inc in (True, False):
        g = sd.snnJaccard(idx, include_self=inc)
        o = oracle_lib.snn_jaccard(idx, include_self=inc)
        assert np.array_equal(bits(g), bits(o))


@pytest.mark.parametrize("edges", ["knn", "shared"])
@pytest.mark.parametrize("prune", [0.0, 1 / 15, 0.5, 1.0])
def test_snn_graph(oracle_lib, gpu_ctx, edges, prune):
    idx = knn_idx(oracle_lib, 700, 12, 2)
    for inc in (True, False):
        for sym in ((True, False) if edges == "knn" else (True,)):
            g = sd.sparseSNN(idx, include_self=inc, prune=prune, symmetrise=sym, edges=edges)
            p, ri, x = oracle_lib.snn_graph(idx, include_self=inc, prune=prune,
                                            symmetrise=sym, edges=edges, nthreads=NT)
            same_graph(g, p, ri, x)


def test_snn_from_knn_result_and_c1_pipeline(oracle_lib, gpu_ctx):
    X = synth.abs_normal(2000, 5000, 0.05, 3)
    nn = sd.sparseKNN(X, k=15, method="cosine", verbose=False)
    g = sd.sparseSNN(nn, prune=1 / 15)
    oi, _, _ = oracle_lib.knn(X, k=15, method="cosine", nthreads=NT)
    assert np.array_equal(nn.idx, oi)
    p, ri, x = oracle_lib.snn_graph(oi, include_self=True, prune=1 / 15, symmetrise=True,
                                    nthreads=NT)
    same_graph(g, p, ri, x)
    assert (g.data > 1 / 15).all()
    assert (abs(g - g.T) > 0).nnz == 0


def test_fixtures(gpu_ctx):
    f = FIX["snn_self_listed"]
    w = sd.snnJaccard(idx_array(f["idx"]), include_self=True)
    assert w[:, 1].tolist() == f["expect_w_col2"]
    f = FIX["snn_disjoint"]
    idx = idx_array(f["idx"])
    assert sd.snnJaccard(idx, include_self=False).tolist() == f["expect_w"]
    g = sd.sparseSNN(idx, include_self=False, prune=0, symmetrise=False)
    assert (g.data > 0).all()
    f = FIX["snn_na_padding"]
    w = sd.snnJaccard(idx_array(f["idx"]), include_self=True)
    assert sd.is_na_real(w[1, 1])
    g = sd.sparseSNN(idx_array(f["idx"]), symmetrise=False)
    assert not np.isnan(g.data).any()
    for case in FIX["snn_out_of_range"]["cases"]:
        with pytest.raises(sd.SparseDistError, match="1-based row indices"):
            sd.snnJaccard(idx_array(case))
    f = FIX["snn_duplicates"]
    g = sd.sparseSNN(idx_array(f["idx"]), include_self=True, symmetrise=False).toarray()
    for i, j, v in f["expect_entries"]:
        assert g[i - 1, j - 1] == v
    f = FIX["snn_include_self_override"]
    ws = sd.snnJaccard(idx_array(f["idx"]), include_self=True)
    wn = sd.snnJaccard(idx_array(f["idx"]), include_self=False)
    assert ws[0, 1] == f["expect_with_self_1_2"] and wn[0, 1] == f["expect_without_1_2"]


def test_fixture_shared_awkward(gpu_ctx):
    f = FIX["snn_shared_awkward"]
    idx = idx_array(f["idx"])
    m = sd.sparseSNN(idx, edges="shared", prune=0).toarray()
    assert (m != 0).sum() / 2 == f["expect_shared_pairs"]
    assert np.array_equal(m, m.T) and (np.diag(m) == 0).all()
    for i, j, v in f["expect_shared"]:
        assert m[i - 1, j - 1] == v
    knn = sd.sparseSNN(idx, edges="knn", prune=0).toarray()
    off = knn.copy()
    np.fill_diagonal(off, 0)
    assert (off != 0).sum() / 2 == f["expect_knn_offdiag_pairs"]
    assert knn[3, 3] == 1.0
    for pr in f["prune_values"]:
        got = sd.sparseSNN(idx, edges="shared", prune=pr).toarray()
        assert np.array_equal(got, np.where(m > pr, m, 0.0))
    assert sd.sparseSNN(idx, edges="shared", prune=1).nnz == 0
    with pytest.warns(sd.SparseDistWarning, match="symmetrise"):
        sd.sparseSNN(idx, edges="shared", symmetrise=False)


def test_snn_reads_kept_device_indices(oracle_lib, gpu_ctx):
    """sparseSNN(nn) on the context's last self-search reads the indices the
    search left on the device (no upload), with degenerate columns (NA rows)
    among them; the graph equals the one built from an uploaded copy and the
    oracle's.  A later self-search replaces the kept indices: the earlier
    result then goes up from the host again, and an idx made writable is
    never read from the device."""
    X = synth.abs_normal(800, 3000, 0.05, 30).tolil()
    for c in (7, 1500, 2999):
        X[:, c] = 0
    X = X.tocsc()
    X.eliminate_zeros()
    nn = sd.sparseKNN(X, k=12, method="cosine", verbose=False)
    assert not nn.idx.flags.writeable
    assert (nn.idx[[7, 1500, 2999]] == NA_INT).all()
    gpu_ctx.stats(reset=True)
    g_dev = sd.sparseSNN(nn, prune=1 / 15)
    assert gpu_ctx.stats(reset=True)["snn_device_inputs"] == 1
    g_host = sd.sparseSNN(nn.idx.copy(), prune=1 / 15)
    assert gpu_ctx.stats(reset=True)["snn_device_inputs"] == 0
    p, ri, x = oracle_lib.snn_graph(np.asfortranarray(nn.idx), include_self=True, prune=1 / 15,
                                    symmetrise=True, nthreads=NT)
    same_graph(g_dev, p, ri, x)
    same_graph(g_host, p, ri, x)
    # another self-search replaces the kept indices
    nn2 = sd.sparseKNN(X, k=9, method="binary", verbose=False)
    g_old = sd.sparseSNN(nn, prune=1 / 15)
    assert gpu_ctx.stats(reset=True)["snn_device_inputs"] == 0
    same_graph(g_old, p, ri, x)
    g2 = sd.sparseSNN(nn2, edges="shared")
    assert gpu_ctx.stats(reset=True)["snn_device_inputs"] == 1
    p2, ri2, x2 = oracle_lib.snn_graph(np.asfortranarray(nn2.idx), include_self=True, prune=0,
                                       symmetrise=True, edges="shared", nthreads=NT)
    same_graph(g2, p2, ri2, x2)
    # an idx made writable again goes up from the host
    nn2.idx.setflags(write=True)
    sd.sparseSNN(nn2, edges="shared")
    assert gpu_ctx.stats(reset=True)["snn_device_inputs"] == 0
