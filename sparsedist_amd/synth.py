"""Seeded synthetic inputs for the benchmark workloads and tests.

- ``nb_log1p``: count-like sparse matrix.  round(d * r * c) non-zero cells
  placed uniformly over the matrix (column counts multinomial over columns,
  rows drawn without replacement within a column, so no cell repeats and no
  explicit zero is stored), values NB(mu=2, size=0.5) truncated at zero, then
  log1p(1e4 * x / colsum) per column (the workload described by the
  reference's bench generator, bench/R/03-datasets.R:186-235).
- ``abs_normal``: the same placement with |N(0, 1)| values (the reference
  tests' ``abs(rsparsematrix(...))`` shape, tests/testthat/test-sparseKNN.R:7-12).
- ``unit_normal_dense``: dense N(0, 1) columns normalised to unit length in FP32.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def _positions(rng, nrow, ncol, density):
    nnz = int(round(density * nrow * ncol))
    counts = rng.multinomial(nnz, np.full(ncol, 1.0 / ncol)) if ncol > 0 else np.zeros(0, int)
    counts = np.minimum(counts, nrow)
    indptr = np.zeros(ncol + 1, dtype=np.int64)
    np.cumsum(counts, out=indptr[1:])
    rows = np.empty(int(indptr[-1]), dtype=np.int32)
    for j in range(ncol):
        c = int(counts[j])
        if c:
            r = rng.choice(nrow, size=c, replace=False)
            r.sort()
            rows[indptr[j]:indptr[j + 1]] = r
    return indptr, rows


def _nb_truncated(rng, n, mu=2.0, size=0.5):
    p = size / (size + mu)
    x = rng.negative_binomial(size, p, size=n).astype(np.float64)
    zero = np.flatnonzero(x == 0)
    while zero.size:
        x[zero] = rng.negative_binomial(size, p, size=zero.size)
        zero = zero[x[zero] == 0]
    return x


def nb_log1p(nrow, ncol, density, seed):
    rng = np.random.default_rng(seed)
    indptr, rows = _positions(rng, nrow, ncol, density)
    x = _nb_truncated(rng, rows.size)
    colsum = np.add.reduceat(x, indptr[:-1]) if x.size else np.zeros(ncol)
    colsum = np.where(np.diff(indptr) > 0, colsum, 1.0)
    scale = np.repeat(1e4 / colsum, np.diff(indptr))
    vals = np.log1p(x * scale)
    return sp.csc_matrix((vals, rows, indptr), shape=(nrow, ncol))


def abs_normal(nrow, ncol, density, seed):
    rng = np.random.default_rng(seed)
    indptr, rows = _positions(rng, nrow, ncol, density)
    vals = np.abs(rng.standard_normal(rows.size))
    vals[vals == 0] = 1.0
    return sp.csc_matrix((vals, rows, indptr), shape=(nrow, ncol))


def unit_normal_dense(nrow, ncol, seed):
    rng = np.random.default_rng(seed)
    E = rng.standard_normal((nrow, ncol), dtype=np.float32)
    E /= np.linalg.norm(E, axis=0, keepdims=True)
    return E
