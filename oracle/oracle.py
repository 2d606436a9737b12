"""ctypes bindings for the CPU oracle (oracle/_build/libsd_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the reported
host-CPU baseline.  Only the test suite, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module; the product path
(``sparsedist_amd``) never does.

Every wrapper follows the C function of the same name in ``sd_oracle.c``; see
that file for the reference file:line each one restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libsd_oracle.so")

BINARY, COSINE, EUCLIDEAN, MANHATTAN, PEARSON, COVARIANCE, JS = range(7)
METHODS = {"binary": BINARY, "jaccard": BINARY, "cosine": COSINE, "euclidean": EUCLIDEAN,
           "manhattan": MANHATTAN, "pearson": PEARSON, "covariance": COVARIANCE, "js": JS}
NA_INTEGER = np.iinfo(np.int32).min
NA_REAL_BITS = 0x7FF00000000007A2

OR_OK, OR_EINVAL, OR_ENONFINITE, OR_ENOMEM, OR_ENOCAND, OR_ECAP = range(6)


class OracleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class _Csc(C.Structure):
    _fields_ = [("nrow", C.c_int64), ("ncol", C.c_int64), ("p", C.c_void_p),
                ("i", C.c_void_p), ("x", C.c_void_p)]


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _check(rc):
    if rc != OR_OK:
        raise OracleError(rc, lib().or_last_error().decode())


def _csc(m):
    """Accepts a scipy.sparse matrix (converted to CSC with sorted indices) or
    a (p, i, x, nrow) tuple; returns (struct, keepalive)."""
    if isinstance(m, tuple):
        p, i, x, nrow = m
    else:
        m = m.tocsc()
        m.sort_indices()
        p, i, x, nrow = m.indptr, m.indices, m.data, m.shape[0]
    p = np.ascontiguousarray(p, dtype=np.int64)
    i = np.ascontiguousarray(i, dtype=np.int32)
    x = np.ascontiguousarray(x, dtype=np.float64)
    s = _Csc(int(nrow), len(p) - 1, p.ctypes.data, i.ctypes.data, x.ctypes.data)
    return s, (p, i, x)


def slab(method, m, m2, dist=True, nthreads=1):
    a, ka = _csc(m)
    b, kb = _csc(m2)
    D = np.empty((a.ncol, b.ncol), dtype=np.float64, order="F")
    _check(lib().or_slab(C.c_int(METHODS[method]), C.byref(a), C.byref(b), C.c_int(int(dist)),
                         C.c_int(nthreads), D.ctypes.data_as(C.c_void_p)))
    return D


def topk_block(D, k, decreasing, self_row, nthreads=1):
    D = np.asfortranarray(D, dtype=np.float64)
    nrow, ncol = D.shape
    sr = np.ascontiguousarray(self_row, dtype=np.int32)
    if sr.size != ncol:
        raise OracleError(OR_EINVAL, "self_row must have one entry per column of D.")
    idx = np.empty((ncol, max(k, 0)), dtype=np.int32, order="F")
    val = np.empty((ncol, max(k, 0)), dtype=np.float64, order="F")
    _check(lib().or_topk_block(D.ctypes.data_as(C.c_void_p), C.c_int64(nrow), C.c_int64(ncol),
                               C.c_int(k), C.c_int(int(decreasing)), sr.ctypes.data_as(C.c_void_p),
                               C.c_int(nthreads), idx.ctypes.data_as(C.c_void_p),
                               val.ctypes.data_as(C.c_void_p)))
    return idx, val


def topk_rows_accum(D, k, decreasing, best_idx, best_val, col_offset, self_col, nthreads=1,
                    tile_rows=64):
    """Updates best_idx / best_val (Fortran-ordered nrow x k) in place."""
    D = np.asfortranarray(D, dtype=np.float64)
    nrow, ncol = D.shape
    sc = np.ascontiguousarray(self_col, dtype=np.int32)
    assert best_idx.flags.f_contiguous and best_val.flags.f_contiguous
    _check(lib().or_topk_rows_accum(D.ctypes.data_as(C.c_void_p), C.c_int64(nrow), C.c_int64(ncol),
                                    C.c_int(k), C.c_int(int(decreasing)),
                                    best_idx.ctypes.data_as(C.c_void_p),
                                    best_val.ctypes.data_as(C.c_void_p), C.c_int64(col_offset),
                                    sc.ctypes.data_as(C.c_void_p), C.c_int(nthreads),
                                    C.c_int(tile_rows)))


def knn(X, Y=None, k=10, method="binary", dist=True, include_self=False, drop_degenerate=True,
        block_size=256, nthreads=1):
    """Returns (idx, dist, warned): idx int32 / dist float64, nq x k_eff
    (Fortran order), NA_integer_ / NA_real_ padding."""
    a, ka = _csc(X)
    yb = None
    if Y is not None:
        b, kb = _csc(Y)
        yb = C.byref(b)
    nq = a.ncol
    idx = np.empty(nq * k, dtype=np.int32)
    dst = np.empty(nq * k, dtype=np.float64)
    keff = C.c_int(0)
    warn = C.c_int(0)
    _check(lib().or_knn(C.byref(a), yb, C.c_int(k), C.c_int(METHODS[method]), C.c_int(int(dist)),
                        C.c_int(int(include_self)), C.c_int(int(drop_degenerate)),
                        C.c_int64(block_size), C.c_int(nthreads), idx.ctypes.data_as(C.c_void_p),
                        dst.ctypes.data_as(C.c_void_p), C.byref(keff), C.byref(warn)))
    ke = keff.value
    return (idx[: nq * ke].reshape((nq, ke), order="F"),
            dst[: nq * ke].reshape((nq, ke), order="F"), bool(warn.value))


def last_phases():
    """(slab seconds, top-k seconds) of the last knn() call on this thread."""
    a, b = C.c_double(0.0), C.c_double(0.0)
    lib().or_last_phases(C.byref(a), C.byref(b))
    return a.value, b.value


def knn_cross(X, Y, k=10, method="binary", dist=True, drop_degenerate=True, block_size=256,
              nthreads=1):
    a, ka = _csc(X)
    b, kb = _csc(Y)
    nq, nr = a.ncol, b.ncol
    fi = np.empty(nq * k, dtype=np.int32)
    fd = np.empty(nq * k, dtype=np.float64)
    ri = np.empty(nr * k, dtype=np.int32)
    rd = np.empty(nr * k, dtype=np.float64)
    kf, kr, warn = C.c_int(0), C.c_int(0), C.c_int(0)
    _check(lib().or_knn_cross(C.byref(a), C.byref(b), C.c_int(k), C.c_int(METHODS[method]),
                              C.c_int(int(dist)), C.c_int(int(drop_degenerate)),
                              C.c_int64(block_size), C.c_int(nthreads),
                              fi.ctypes.data_as(C.c_void_p), fd.ctypes.data_as(C.c_void_p),
                              C.byref(kf), ri.ctypes.data_as(C.c_void_p),
                              rd.ctypes.data_as(C.c_void_p), C.byref(kr), C.byref(warn)))
    f = (fi[: nq * kf.value].reshape((nq, kf.value), order="F"),
         fd[: nq * kf.value].reshape((nq, kf.value), order="F"))
    r = (ri[: nr * kr.value].reshape((nr, kr.value), order="F"),
         rd[: nr * kr.value].reshape((nr, kr.value), order="F"))
    return f, r, warn.value


def snn_jaccard(idx, include_self=True, nthreads=1):
    idx = np.asfortranarray(idx, dtype=np.int32)
    n, k = idx.shape
    w = np.empty((n, k), dtype=np.float64, order="F")
    _check(lib().or_snn_jaccard(idx.ctypes.data_as(C.c_void_p), C.c_int(n), C.c_int(k),
                                C.c_int(int(include_self)), C.c_int(nthreads),
                                w.ctypes.data_as(C.c_void_p)))
    return w


def snn_graph(idx, include_self=True, prune=0.0, symmetrise=True, edges="knn", nthreads=1):
    """Returns (colptr int64[n+1], rowind int32[nnz], val float64[nnz])."""
    idx = np.asfortranarray(idx, dtype=np.int32)
    n, k = idx.shape
    cap = 2 * n * k + 16
    while True:
        p = np.empty(n + 1, dtype=np.int64)
        ri = np.empty(cap, dtype=np.int32)
        x = np.empty(cap, dtype=np.float64)
        nnz = C.c_int64(0)
        rc = lib().or_snn_graph(idx.ctypes.data_as(C.c_void_p), C.c_int(n), C.c_int(k),
                                C.c_int(int(include_self)), C.c_double(prune),
                                C.c_int(int(symmetrise)), C.c_int(0 if edges == "knn" else 1),
                                C.c_int(nthreads), p.ctypes.data_as(C.c_void_p),
                                ri.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p),
                                C.c_int64(cap), C.byref(nnz))
        if rc == OR_ECAP:
            cap = int(nnz.value) + 16
            continue
        _check(rc)
        m = int(nnz.value)
        return p, ri[:m].copy(), x[:m].copy()
