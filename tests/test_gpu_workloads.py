"""GPU parity at the benchmark workloads' own shapes: the rows, density and
value distribution of each workload, with enough columns that the large-search
machinery runs (threshold pre-pass, filtered pass, the symmetric or mirrored
screen), checked against the CPU oracle.

Per-query results do not depend on the other queries, so a seeded sample of
query columns is compared with the oracle bit for bit (idx and dist); every
row of a symmetric self-search is also compared with a full-grid cross search
of the same columns.  Reference: R/sparseKNN.R:110-313,
R/sparseKNNCross.R:33-40, 151-188.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import sparsedist_amd as sd
from conftest import bits, self_rows_from_cross
from sparsedist_amd import synth

pytestmark = pytest.mark.gpu
NT = min(16, os.cpu_count() or 1)


def same(gpu, ora, label):
    gi, gd = gpu
    oi, od = ora[0], ora[1]
    assert gi.shape == oi.shape, label
    bad = np.argwhere(gi != oi)
    assert bad.size == 0, f"{label}: idx differs at {bad[:5].tolist()}"
    db = np.argwhere(bits(gd) != bits(od))
    assert db.size == 0, f"{label}: dist differs at {db[:5].tolist()}"


def self_search_parity(oracle_lib, gpu_ctx, Y, method, k, nsample, seed, margin=0):
    """Self-search of every column of Y in one call (symmetric filtered pass):
    sampled rows against the oracle, every row against a full-grid search."""
    n = Y.shape[1]
    dY = sd.DeviceMatrix(Y, gpu_ctx)
    gpu_ctx.stats(reset=True)
    gpu_ctx.set_timing(True)  # launches are counted while timing is on
    try:
        g = sd.sparseKNN(dY, k=k, method=method, margin=margin, verbose=False, context=gpu_ctx)
    finally:
        gpu_ctx.set_timing(False)
    st = gpu_ctx.stats(reset=True)
    assert st["fused_calls"] == 1 and st["filter_launches"] == 1, st
    assert st["filter_pairs"] == n * (n + 1) // 2, st
    rng = np.random.default_rng(seed)
    qcols = sorted(set(rng.choice(n, nsample, replace=False).tolist()) | {0, n - 1})
    o = oracle_lib.knn(Y[:, qcols], Y, k=k + 1, method=method, nthreads=NT)
    same((g.idx[qcols], g.dist[qcols]), self_rows_from_cross(o[0], o[1], qcols, k, False),
         f"{method}: sampled rows vs oracle")
    c = sd.sparseKNN(dY, Y.copy(), k=k + 1, method=method, margin=margin, verbose=False,
                     context=gpu_ctx)
    same((g.idx, g.dist), self_rows_from_cross(c.idx, c.dist, list(range(n)), k, False),
         f"{method}: symmetric pass vs full grid, every row")
    return st


@pytest.fixture(scope="module")
def count_matrix():
    # the count-matrix workload's rows, density and values, 16640 columns
    # (130 reference tiles of 128: the pre-pass + filtered pass)
    return synth.nb_log1p(20000, 16640, 0.06, 501)


@pytest.mark.parametrize("method", ["binary", "cosine"])
def test_count_matrix_workload_self_search(oracle_lib, gpu_ctx, count_matrix, method):
    """20000 rows, density 0.06, NB/log1p, k = 20: the headline bench's own
    path (FP4 Jaccard or BF16 cosine symmetric pass, FP64 re-rank)."""
    self_search_parity(oracle_lib, gpu_ctx, count_matrix, method, 20, 300, 502)


def test_count_matrix_workload_snn(oracle_lib, gpu_ctx, count_matrix):
    """sparseSNN(prune = 1/15) of the kNN graph at this size, identical to the
    oracle's CSC (structure and weight bits)."""
    nn = sd.sparseKNN(count_matrix, k=20, method="cosine", verbose=False, context=gpu_ctx)
    G = sd.sparseSNN(nn, prune=1 / 15, context=gpu_ctx)
    p, ri, x = oracle_lib.snn_graph(nn.idx, include_self=True, prune=1 / 15, symmetrise=True,
                                    nthreads=NT)
    assert np.array_equal(G.indptr, p) and np.array_equal(G.indices, ri)
    assert np.array_equal(bits(G.data), bits(x))


@pytest.fixture(scope="module")
def large_sparse():
    # the large precision-form workload's rows, density and values, 16640
    # columns (its own row count: 30000 rows, ~900 stored per column)
    return synth.nb_log1p(30000, 16640, 0.03, 511)


@pytest.mark.parametrize("method", ["binary", "cosine", "euclidean"])
def test_large_sparse_workload_self_search(oracle_lib, gpu_ctx, large_sparse, method):
    """30000 rows, density 0.03, NB/log1p, k = 20 with re-rank margin m = 16:
    BF16 slabs with FP32 accumulation and FP64 re-rank for cosine and
    euclidean, FP4 patterns with exact integer counts for Jaccard."""
    st = self_search_parity(oracle_lib, gpu_ctx, large_sparse, method, 20, 250, 512,
                            margin=16 if method != "binary" else 0)
    if method != "binary":
        assert st["last_margin"] == 16


def test_dense_embedding_workload_cross(oracle_lib, gpu_ctx):
    """256-dimensional N(0,1) unit columns, cosine, dist = FALSE, k = 10, both
    directions, with both sides above 16384 columns so one filtered pass
    serves both directions (the reverse search's buffers filled through the
    mirror, on dense signed data).  Both directions equal two independent
    searches on every row and the oracle on sampled rows."""
    E = synth.unit_normal_dense(256, 17000, 521).astype(np.float64)
    Y = synth.unit_normal_dense(256, 20000, 522).astype(np.float64)
    dE, dY = sd.DeviceMatrix(E, gpu_ctx), sd.DeviceMatrix(Y, gpu_ctx)
    gpu_ctx.stats(reset=True)
    gpu_ctx.set_timing(True)
    try:
        cr = sd.sparseKNNCross(dE, dY, k=10, method="cosine", dist=False, verbose=False,
                               context=gpu_ctx)
    finally:
        gpu_ctx.set_timing(False)
    st = gpu_ctx.stats(reset=True)
    assert st["filter_launches"] == 1 and st["filter_pairs"] == 17000 * 20000, st
    fw = sd.sparseKNN(dE, dY, k=10, method="cosine", dist=False, verbose=False, context=gpu_ctx)
    rv = sd.sparseKNN(dY, dE, k=10, method="cosine", dist=False, verbose=False, context=gpu_ctx)
    same((cr["forward"].idx, cr["forward"].dist), (fw.idx, fw.dist), "forward vs own search")
    same((cr["reverse"].idx, cr["reverse"].dist), (rv.idx, rv.dist), "reverse vs own search")
    rng = np.random.default_rng(523)
    sf = np.sort(rng.choice(17000, 200, replace=False))
    sr = np.sort(rng.choice(20000, 200, replace=False))
    Es, Ys = sp.csc_matrix(E), sp.csc_matrix(Y)
    of = oracle_lib.knn(Es[:, sf], Ys, k=10, method="cosine", dist=False, nthreads=NT)
    orv = oracle_lib.knn(Ys[:, sr], Es, k=10, method="cosine", dist=False, nthreads=NT)
    same((cr["forward"].idx[sf], cr["forward"].dist[sf]), of, "forward vs oracle")
    same((cr["reverse"].idx[sr], cr["reverse"].dist[sr]), orv, "reverse vs oracle")


def test_large_query_set_one_split_prepass(oracle_lib, gpu_ctx):
    """A short-row query set of 2048 query tiles (eight per CU on 256 CUs)
    runs the threshold pre-pass as one split per query tile, whose lists need
    no merge: 524,288 256-dimensional query columns against 16,384 references
    (128 reference tiles, so the search takes the two-pass path).  Sampled
    rows equal the oracle's bit for bit."""
    nq, nr = 2048 * 256, 16384
    Q = synth.unit_normal_dense(256, nq, 531).astype(np.float64)
    Y = synth.unit_normal_dense(256, nr, 532).astype(np.float64)
    dQ, dY = sd.DeviceMatrix(Q, gpu_ctx), sd.DeviceMatrix(Y, gpu_ctx)
    gpu_ctx.stats(reset=True)
    gpu_ctx.set_timing(True)  # launches are counted by the kernel timers
    try:
        nn = sd.sparseKNN(dQ, dY, k=10, method="cosine", verbose=False, context=gpu_ctx)
    finally:
        gpu_ctx.set_timing(False)
    st = gpu_ctx.stats(reset=True)
    assert st["filter_launches"] == 1 and st["certified_queries"] == nq, st
    rng = np.random.default_rng(533)
    s = np.sort(np.concatenate([rng.choice(nq, 298, replace=False), [0, nq - 1]]))
    s = np.unique(s)
    o = oracle_lib.knn(sp.csc_matrix(Q[:, s]), sp.csc_matrix(Y), k=10, method="cosine",
                       nthreads=NT)
    same((nn.idx[s], nn.dist[s]), o, "one-split pre-pass vs oracle")
