"""Seeded synthetic inputs for the benchmark workloads and tests.

- ``nb_log1p``: count-like sparse matrix.  round(d * r * c) non-zero cells
  placed uniformly over the matrix (column counts multinomial over columns,
  rows drawn without replacement within a column, so no cell repeats and no
  explicit zero is stored), values NB(mu=2, size=0.5) truncated at zero, then
  log1p(1e4 * x / colsum) per column (the workload described by the
  reference's bench generator, bench/R/03-datasets.R:186-235).
- ``nb_log1p_parallel``: the same distribution generated by worker processes
  (for the 20000 x 500000 serving reference).
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


def _nb_block(args):
    """Worker of nb_log1p_parallel: columns [c0, c1) written into shared memory."""
    from multiprocessing import shared_memory
    (nrow, c0, c1, counts, seed, shm_rows, shm_vals, nnz_total, off) = args
    rng = np.random.default_rng(seed)
    n = c1 - c0
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=indptr[1:])
    rows = np.empty(int(indptr[-1]), dtype=np.int32)
    for b0 in range(0, n, 1024):
        b1 = min(n, b0 + 1024)
        need = counts[b0:b1].astype(np.int64)
        keys = np.empty(0, dtype=np.int64)
        while True:
            cols = np.repeat(np.arange(b0, b1, dtype=np.int64), need)
            r = rng.integers(0, nrow, size=cols.size, dtype=np.int64)
            keys = np.unique(np.concatenate((keys, cols * nrow + r)))
            got = np.bincount(keys // nrow - b0, minlength=b1 - b0)
            need = counts[b0:b1] - got
            if not need.any():
                break
        rows[indptr[b0]:indptr[b1]] = (keys % nrow).astype(np.int32)
    x = _nb_truncated(rng, rows.size)
    colsum = np.add.reduceat(x, indptr[:-1]) if x.size else np.zeros(n)
    colsum = np.where(np.diff(indptr) > 0, colsum, 1.0)
    vals = np.log1p(x * np.repeat(1e4 / colsum, np.diff(indptr)))
    a = shared_memory.SharedMemory(name=shm_rows)
    b = shared_memory.SharedMemory(name=shm_vals)
    np.ndarray(nnz_total, dtype=np.int32, buffer=a.buf)[off:off + rows.size] = rows
    np.ndarray(nnz_total, dtype=np.float64, buffer=b.buf)[off:off + rows.size] = vals
    a.close()
    b.close()
    return rows.size


def nb_log1p_parallel(nrow, ncol, density, seed, workers=16):
    """nb_log1p's distribution for very large matrices, generated by worker
    processes over column ranges (within a column: rows drawn with
    replacement, duplicates dropped and redrawn, which leaves a uniformly
    random subset of the column's count).  Deterministic for a given
    (seed, workers).  Call it before the process initialises the GPU: the
    workers are forked."""
    import multiprocessing as mp
    from multiprocessing import shared_memory
    ss = np.random.SeedSequence(seed)
    rng = np.random.default_rng(ss.spawn(1)[0])
    nnz = int(round(density * nrow * ncol))
    counts = np.minimum(rng.multinomial(nnz, np.full(ncol, 1.0 / ncol)), nrow).astype(np.int64)
    indptr = np.zeros(ncol + 1, dtype=np.int64)
    np.cumsum(counts, out=indptr[1:])
    total = int(indptr[-1])
    a = shared_memory.SharedMemory(create=True, size=max(total * 4, 8))
    b = shared_memory.SharedMemory(create=True, size=max(total * 8, 8))
    try:
        bounds = np.linspace(0, ncol, workers + 1).astype(np.int64)
        seeds = ss.spawn(workers + 1)[1:]
        jobs = [(nrow, int(bounds[w]), int(bounds[w + 1]), counts[bounds[w]:bounds[w + 1]],
                 seeds[w], a.name, b.name, total, int(indptr[bounds[w]]))
                for w in range(workers) if bounds[w + 1] > bounds[w]]
        with mp.get_context("fork").Pool(len(jobs)) as pool:
            pool.map(_nb_block, jobs)
        rows = np.ndarray(total, dtype=np.int32, buffer=a.buf).copy()
        vals = np.ndarray(total, dtype=np.float64, buffer=b.buf).copy()
    finally:
        a.close()
        a.unlink()
        b.close()
        b.unlink()
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
