"""Host mirror of sparseDist's kNN / SNN R API on top of libsd_hip.so.

``sparseKNN``, ``sparseKNNCross`` and ``sparseSNN`` keep the reference's
argument names, defaults, validation order, error and warning text and
result layout (R/sparseKNN.R:110-313, R/sparseKNNCross.R:44-199,
R/sparseSNN.R:120-251).  All computation happens on the GPU through the C
ABI; there is no CPU path.

Result conventions follow R's memory representation: ``idx`` is an int32
``nq x k`` array (Fortran order) with ``NA_INTEGER`` (INT32_MIN) padding,
``dist`` a float64 array with ``NA_REAL`` (R's NA_real_ bit pattern).
"""
from __future__ import annotations

import ctypes as C
import math
import numbers
import warnings

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import SparseDistError, check

NA_INTEGER = np.int32(np.iinfo(np.int32).min)
NA_REAL = np.array([0x7FF00000000007A2], dtype=np.uint64).view(np.float64)[0]
INT_MAX = 2147483647

_METHODS = ("binary", "jaccard", "cosine", "euclidean", "manhattan", "pearson", "js",
            "covariance")
_METHOD_CODE = {"binary": _lib.SD_BINARY, "cosine": _lib.SD_COSINE,
                "euclidean": _lib.SD_EUCLIDEAN, "manhattan": _lib.SD_MANHATTAN,
                "pearson": _lib.SD_PEARSON, "covariance": _lib.SD_COVARIANCE,
                "js": _lib.SD_JS}
_PRECISION = {"auto": _lib.SD_PREC_AUTO, "exact": _lib.SD_PREC_EXACT,
              "fp32": _lib.SD_PREC_FP32, "bf16": _lib.SD_PREC_BF16}


class SparseDistWarning(UserWarning):
    pass


def is_na_real(x):
    """True where x holds R's NA_real_ (not any NaN)."""
    a = np.asarray(x, dtype=np.float64)
    return a.view(np.uint64) == np.uint64(0x7FF00000000007A2)


# ---------------------------------------------------------------------------
# context and resident matrices
# ---------------------------------------------------------------------------
class Context:
    """A device, its stream and scratch memory (sd_context)."""

    def __init__(self, device: int = 0):
        L = _lib.load()
        h = C.c_void_p()
        check(L.sd_context_create(device, C.byref(h)))
        self._h = h
        self.device = device

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            _lib.load().sd_context_destroy(self._h)
            self._h = C.c_void_p()

    def cancel(self):
        check(_lib.load().sd_cancel(self._h))

    def set_timing(self, on: bool):
        check(_lib.load().sd_context_set_timing(self._h, int(bool(on))))

    def set_buffer_budget(self, nbytes: int):
        """Device bytes one fused search may hold in candidate buffers (0: half
        the free memory); more query columns than fit run in blocks."""
        check(_lib.load().sd_context_set_buffer_budget(self._h, int(nbytes)))

    def stats(self, reset: bool = False) -> dict:
        s = _lib.Stats()
        check(_lib.load().sd_context_stats(self._h, C.byref(s), int(bool(reset))))
        return {f: getattr(s, f) for f, _ in s._fields_}

    def stream(self) -> int:
        return _lib.load().sd_context_stream(self._h) or 0

    def autotune(self, measure: bool = True) -> dict:
        """Measure (or, with measure=False, report) the rates of the path
        chooser's cost model on this device; see sd_autotune."""
        t = _lib.TuneInfo()
        check(_lib.load().sd_autotune(self._h, int(bool(measure)), C.byref(t)))
        return {f: getattr(t, f) for f, _ in t._fields_ if f != "reserved"}

    def set_tuning(self, on: bool):
        """Tune the filtered search's threshold sample online (default) or use
        the fixed rule; see sd_context_set_tuning."""
        check(_lib.load().sd_context_set_tuning(self._h, int(bool(on))))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _as_csc(M):
    """Coerce dense arrays and other sparse formats to CSC with sorted
    indices (reference .as_csparse, R/sparseDist.R:198-202)."""
    if sp.issparse(M):
        M = M.tocsc()
    else:
        M = sp.csc_matrix(np.asarray(M, dtype=np.float64))
    if not M.has_sorted_indices:
        M = M.sorted_indices()
    return M


class DeviceMatrix:
    """A CSC matrix resident in device memory (sd_matrix).  Explicit zeros
    are dropped at upload; derived layouts are cached with it."""

    def __init__(self, M, context: Context | None = None):
        self.ctx = context or default_context()
        M = _as_csc(M)
        self.shape = M.shape
        p = np.ascontiguousarray(M.indptr, dtype=np.int64)
        i = np.ascontiguousarray(M.indices, dtype=np.int32)
        x = np.ascontiguousarray(M.data, dtype=np.float64)
        h = C.c_void_p()
        check(_lib.load().sd_matrix_upload_csc(self.ctx.handle, M.shape[0], M.shape[1],
                                               p.ctypes.data, i.ctypes.data, x.ctypes.data,
                                               C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    @property
    def nrow(self):
        return self.shape[0]

    @property
    def ncol(self):
        return self.shape[1]

    def drop_cache(self):
        check(_lib.load().sd_matrix_drop_cache(self._h))

    def free(self):
        if self._h:
            _lib.load().sd_matrix_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# validators (reference R/utils-validate.R:17-42)
# ---------------------------------------------------------------------------
def _chk_flag(value, name):
    if not isinstance(value, (bool, np.bool_)):
        raise ValueError(f"'{name}' must be TRUE or FALSE.")


def _is_number(value):
    return isinstance(value, numbers.Real) and not isinstance(value, (bool, np.bool_))


def _chk_count(value, name, allow_zero=False):
    lower = 0 if allow_zero else 1
    if isinstance(value, np.ndarray) and value.size == 1 and value.ndim <= 1:
        value = value.reshape(()).item() if value.ndim else value.item()
    ok = _is_number(value) and not (isinstance(value, float) and math.isnan(value))
    if ok:
        v = float(value)
        ok = math.isfinite(v) and v == math.floor(v) and v >= lower and v <= INT_MAX
    if not ok:
        kind = "non-negative" if allow_zero else "positive"
        raise ValueError(
            f"'{name}' must be a {kind} integer no larger than .Machine$integer.max.")
    return int(value)


def _match_method(method):
    """match.arg over the reference's method list (exact, then unique prefix)."""
    if isinstance(method, (list, tuple)):
        method = method[0]
    if not isinstance(method, str):
        raise ValueError("'arg' must be NULL or a character vector")
    if method in _METHODS:
        m = method
    else:
        hits = [c for c in _METHODS if c.startswith(method)] if method else []
        if len(hits) != 1:
            raise ValueError("'arg' should be one of " + ", ".join(f"“{c}”" for c in _METHODS))
        m = hits[0]
    return "binary" if m == "jaccard" else m


def _direction(method, dist):
    """Ranking direction from the method (R/sparseKNN.R:168-181)."""
    if method in ("euclidean", "manhattan", "js") and not dist:
        raise ValueError(
            "'dist = FALSE' is only meaningful for methods \"binary\", \"cosine\" and "
            f"\"pearson\"; {method} has no similarity form.")
    return {"binary": not dist, "cosine": not dist, "pearson": not dist, "covariance": True,
            "euclidean": False, "manhattan": False, "js": False}[method]


def detect_cores() -> int:
    """Usable cores minus one, at least 1 (R/sparseDist.R:221-224)."""
    return int(_lib.load().sd_detect_cores())


# ---------------------------------------------------------------------------
# kNN
# ---------------------------------------------------------------------------
def _colnames(M):
    """Column names of a labelled input (a pandas DataFrame), else None: the
    reference names the result rows after colnames(X) (R/sparseKNN.R:303-312)."""
    cols = getattr(M, "columns", None)
    return None if cols is None else [str(c) for c in cols]


class KNNResult(dict):
    """``list(idx, dist)`` with the ``self_search`` attribute; ``names`` holds
    the row names (the query matrix's column names, when it has any)."""

    def __init__(self, idx, dist, self_search, names=None, kept=None):
        super().__init__(idx=idx, dist=dist)
        self.self_search = self_search
        self.names = names
        # (context, serial) when the context still may hold these indices on
        # the device (a self-search's idx, returned read-only)
        self._kept = kept

    @property
    def idx(self):
        return self["idx"]

    @property
    def dist(self):
        return self["dist"]


def _params(k, method, dist, include_self, drop_degenerate, block_size, precision, margin):
    p = _lib.KnnParams()
    p.k = k
    p.method = _METHOD_CODE[method]
    p.dist = int(bool(dist))
    p.include_self = int(bool(include_self))
    p.drop_degenerate = int(bool(drop_degenerate))
    if precision not in _PRECISION:
        raise ValueError(f"unknown precision {precision!r}")
    p.precision = _PRECISION[precision]
    p.margin = int(margin)
    p.block_size = int(block_size)
    return p


def _device(M, ctx):
    if isinstance(M, DeviceMatrix):
        return M, None
    dm = DeviceMatrix(M, ctx)
    return dm, dm


def _ncol(M):
    return M.ncol if isinstance(M, DeviceMatrix) else M.shape[1]


def _nrow(M):
    return M.nrow if isinstance(M, DeviceMatrix) else M.shape[0]


def sparseKNN(X, Y=None, k=10, method="binary", dist=True, include_self=False,
              drop_degenerate=True, block_size=256, ncores=1, verbose=True, *,
              precision="auto", margin=0, context: Context | None = None):
    """Blocked exact k-nearest neighbours of the columns of X (reference
    sparseKNN, R/sparseKNN.R:110-313).  ``block_size`` and ``ncores`` are
    validated as in the reference; results never depend on them."""
    if X is None:
        raise ValueError("'X' must be provided.")
    _chk_flag(dist, "dist")
    _chk_flag(include_self, "include_self")
    _chk_flag(drop_degenerate, "drop_degenerate")
    _chk_flag(verbose, "verbose")
    ncores = _chk_count(ncores, "ncores", allow_zero=True)
    k = _chk_count(k, "k")
    block_size = _chk_count(block_size, "block_size")
    method = _match_method(method)
    _direction(method, dist)
    nq = _ncol(X)
    if nq < 1:
        raise ValueError("'X' must have at least one column.")
    selfmode = Y is None
    if not selfmode and _nrow(X) != _nrow(Y):
        raise ValueError("'X' and 'Y' must have the same number of rows.")
    ctx = context or default_context()
    dx, _kx = _device(X, ctx)
    dy = None
    if not selfmode:
        dy, _ky = _device(Y, ctx)
    params = _params(k, method, dist, include_self, drop_degenerate, block_size, precision, margin)
    idx = np.empty(nq * k, dtype=np.int32)
    dst = np.empty(nq * k, dtype=np.float64)
    keff, warn = C.c_int32(0), C.c_int32(0)
    check(_lib.load().sd_knn(ctx.handle, dx.handle, dy.handle if dy else None, C.byref(params),
                             idx.ctypes.data, dst.ctypes.data, C.byref(keff), C.byref(warn)))
    ke = keff.value
    if warn.value:
        warnings.warn(f"k ({k}) exceeds the {ke} available neighbours; returning {ke}.",
                      SparseDistWarning, stacklevel=2)
    ix = idx[: nq * ke].reshape((nq, ke), order="F")
    kept = None
    if selfmode:
        # the context keeps this idx on the device for sparseSNN(result); the
        # array is read-only so that the two copies cannot drift apart
        serial = _kept_serial(ctx)
        if serial > 0:
            ix.setflags(write=False)
            kept = (ctx, serial)
    return KNNResult(ix, dst[: nq * ke].reshape((nq, ke), order="F"), selfmode, _colnames(X),
                     kept)


def _kept_info(ctx):
    """(device pointer, n, k, serial) of the context's kept self-search idx."""
    d, n, k, serial = C.c_void_p(), C.c_int32(0), C.c_int32(0), C.c_int64(0)
    check(_lib.load().sd_knn_kept(ctx.handle, C.byref(d), C.byref(n), C.byref(k),
                                  C.byref(serial)))
    return d.value, n.value, k.value, serial.value


def _kept_serial(ctx):
    return _kept_info(ctx)[3]


def sparseKNNCross(X, Y=None, k=10, method="binary", dist=True, drop_degenerate=True,
                   block_size=256, ncores=1, verbose=True, *, precision="auto", margin=0,
                   context: Context | None = None):
    """Both directions (reference sparseKNNCross, R/sparseKNNCross.R:44-199):
    ``forward`` equals sparseKNN(X, Y) and ``reverse`` equals sparseKNN(Y, X)
    bit for bit."""
    if Y is None:
        raise ValueError("'Y' is required. For a single matrix use sparseKNN(), which already "
                         "exploits symmetry via the full/diag arguments.")
    _chk_flag(dist, "dist")
    _chk_flag(drop_degenerate, "drop_degenerate")
    _chk_flag(verbose, "verbose")
    ncores = _chk_count(ncores, "ncores", allow_zero=True)
    k = _chk_count(k, "k")
    block_size = _chk_count(block_size, "block_size")
    method = _match_method(method)
    _direction(method, dist)
    if _nrow(X) != _nrow(Y):
        raise ValueError("'X' and 'Y' must have the same number of rows.")
    nq, nr = _ncol(X), _ncol(Y)
    if nq < 1:
        raise ValueError("'X' must have at least one column.")
    if nr < 1:
        raise ValueError("'Y' must have at least one column.")
    ctx = context or default_context()
    dx, _kx = _device(X, ctx)
    dy, _ky = _device(Y, ctx)
    params = _params(k, method, dist, False, drop_degenerate, block_size, precision, margin)
    fi = np.empty(nq * k, dtype=np.int32)
    fd = np.empty(nq * k, dtype=np.float64)
    ri = np.empty(nr * k, dtype=np.int32)
    rd = np.empty(nr * k, dtype=np.float64)
    kf, kr, warn = C.c_int32(0), C.c_int32(0), C.c_int32(0)
    check(_lib.load().sd_knn_cross(ctx.handle, dx.handle, dy.handle, C.byref(params),
                                   fi.ctypes.data, fd.ctypes.data, C.byref(kf), ri.ctypes.data,
                                   rd.ctypes.data, C.byref(kr), C.byref(warn)))
    if warn.value & 1:
        warnings.warn(f"k ({k}) exceeds the {kf.value} available neighbours for the forward "
                      f"direction; returning {kf.value}.", SparseDistWarning, stacklevel=2)
    if warn.value & 2:
        warnings.warn(f"k ({k}) exceeds the {kr.value} available neighbours for the reverse "
                      f"direction; returning {kr.value}.", SparseDistWarning, stacklevel=2)
    fwd = KNNResult(fi[: nq * kf.value].reshape((nq, kf.value), order="F"),
                    fd[: nq * kf.value].reshape((nq, kf.value), order="F"), False, _colnames(X))
    rev = KNNResult(ri[: nr * kr.value].reshape((nr, kr.value), order="F"),
                    rd[: nr * kr.value].reshape((nr, kr.value), order="F"), False, _colnames(Y))
    return {"forward": fwd, "reverse": rev}


# ---------------------------------------------------------------------------
# SNN
# ---------------------------------------------------------------------------
def _idx_matrix(idx):
    """Validate an index matrix as R/sparseSNN.R:125-166 does and convert it
    to int32 with NA_INTEGER padding."""
    if isinstance(idx, KNNResult):
        ss = idx.self_search
        if ss is False:
            raise ValueError(
                "sparseSNN() requires a self-search sparseKNN() result; the indices from "
                "sparseKNN(X, Y) refer to the columns of Y, which are a different node set.")
        if ss is not True:
            raise ValueError(
                "'idx' must be a matrix, or a list produced by sparseKNN(). This list carries "
                "no 'self_search' marker -- it may have been dropped by subsetting -- so it "
                "cannot be confirmed to describe a single node set. Pass the index matrix "
                "itself if that is intended.")
        idx = idx.idx
    elif isinstance(idx, dict):
        raise ValueError(
            "'idx' must be a matrix, or a list produced by sparseKNN(). This list carries no "
            "'self_search' marker -- it may have been dropped by subsetting -- so it cannot be "
            "confirmed to describe a single node set. Pass the index matrix itself if that is "
            "intended.")
    if not isinstance(idx, np.ndarray) or idx.ndim != 2:
        raise ValueError("'idx' must be a matrix or a sparseKNN() result.")
    n, k = idx.shape
    if n < 1 or k < 1:
        raise ValueError("'idx' must have at least one row and one column.")
    if idx.dtype.kind not in "iuf":
        raise ValueError("'idx' must be a numeric or integer matrix.")
    if idx.dtype == np.int32 and idx.flags["F_CONTIGUOUS"] and idx.min() >= 1 \
            and idx.max() <= n:
        return idx  # all entries are row indices (no NA): nothing to convert
    if idx.dtype.kind == "f":
        a = idx.astype(np.float64)
        na = np.isnan(a)
        with np.errstate(invalid="ignore"):
            bad = ~na & (~np.isfinite(a) | (a != np.floor(a)) | (a < 1) | (a > n))
        if bad.any():
            raise ValueError("'idx' entries must be NA or finite, whole-number, 1-based row "
                             "indices in [1, nrow(idx)].")
        out = np.where(na, NA_INTEGER, np.where(na, 0, a)).astype(np.int32)
    else:
        a = idx.astype(np.int64)
        na = a == int(NA_INTEGER)
        bad = ~na & ((a < 1) | (a > n))
        if bad.any():
            raise ValueError("'idx' entries must be NA or finite, whole-number, 1-based row "
                             "indices in [1, nrow(idx)].")
        out = a.astype(np.int32)
    return np.asfortranarray(out)


def tune_samples(reset: bool = False) -> list:
    """The process-wide threshold-sample table (sd_tune_samples): one dict per
    class of searches with its candidates' (stride, rank), measured ns per
    10^6 scored pairs and timed runs.  reset=True clears the table."""
    L = _lib.load()
    n = C.c_int32(0)
    check(L.sd_tune_samples(None, 0, C.byref(n), 0))
    buf = (_lib.TuneSample * max(n.value, 1))()
    check(L.sd_tune_samples(buf, n.value, C.byref(n), int(bool(reset))))
    out = []
    for e in buf[: n.value]:
        out.append({"method": e.method, "layout": e.layout, "density_bucket": e.density_bucket,
                    "row_bytes_log2": e.row_bytes_log2, "list_len": e.list_len,
                    "symmetric": bool(e.symmetric), "rule": e.rule, "last_run": e.last_run,
                    "candidates": [(e.stride[c], e.rank[c]) for c in range(4)],
                    "ns_per_mpair": list(e.ns_per_mpair), "runs": list(e.runs)})
    return out


class _Missing:
    pass


_MISSING = _Missing()


def sparseSNN(idx, include_self=True, prune=0, symmetrise=_MISSING, edges="knn", ncores=1,
              verbose=True, *, context: Context | None = None):
    """Shared-nearest-neighbour graph (reference sparseSNN,
    R/sparseSNN.R:120-251).  Returns an n x n scipy CSC matrix (the
    dgCMatrix the reference returns)."""
    sym_missing = symmetrise is _MISSING
    if sym_missing:
        symmetrise = True
    idx32 = _idx_matrix(idx)
    _chk_flag(include_self, "include_self")
    _chk_flag(symmetrise, "symmetrise")
    _chk_flag(verbose, "verbose")
    if isinstance(edges, (list, tuple)):
        edges = edges[0]
    if edges not in ("knn", "shared"):
        hits = [e for e in ("knn", "shared") if isinstance(edges, str) and edges
                and e.startswith(edges)]
        if len(hits) != 1:
            raise ValueError("'arg' should be one of “knn”, “shared”")
        edges = hits[0]
    if edges == "shared" and not sym_missing and not symmetrise:
        warnings.warn("'symmetrise' is ignored when edges = \"shared\": that edge set is "
                      "symmetric by construction.", SparseDistWarning, stacklevel=2)
    if (not _is_number(prune) or (isinstance(prune, float) and math.isnan(prune))
            or not math.isfinite(float(prune)) or prune < 0 or prune > 1):
        raise ValueError("'prune' must be a finite number between 0 and 1.")
    _chk_count(ncores, "ncores", allow_zero=True)
    n, k = idx32.shape
    ctx = context or default_context()
    # a self-search result whose indices the context still holds on the
    # device (same search, array left read-only) is not uploaded again
    src, on_device = idx32.ctypes.data, 0
    kept = getattr(idx, "_kept", None) if isinstance(idx, KNNResult) else None
    if kept is not None and kept[0] is ctx and not idx.idx.flags.writeable:
        d, kn, kk, serial = _kept_info(ctx)
        if d and serial == kept[1] and (kn, kk) == (n, k):
            src, on_device = d, 1
    g = C.c_void_p()
    check(_lib.load().sd_snn(ctx.handle, src, n, k, int(include_self), float(prune),
                             int(symmetrise), 0 if edges == "knn" else 1, on_device,
                             C.byref(g)))
    return _graph_to_csc(g, n)


def _graph_to_csc(g, n_rows_cols):
    """Copy an sd_graph to a scipy CSC matrix and free it."""
    n = n_rows_cols
    try:
        nn, nnz = C.c_int64(0), C.c_int64(0)
        check(_lib.load().sd_graph_info(g, C.byref(nn), C.byref(nnz)))
        colptr = np.empty(n + 1, dtype=np.int64)
        rows = np.empty(max(nnz.value, 1), dtype=np.int32)
        vals = np.empty(max(nnz.value, 1), dtype=np.float64)
        check(_lib.load().sd_graph_copy(g, colptr.ctypes.data, rows.ctypes.data,
                                        vals.ctypes.data))
    finally:
        _lib.load().sd_graph_free(g)
    m = nnz.value
    return sp.csc_matrix((vals[:m], rows[:m], colptr), shape=(n, n))


# ---------------------------------------------------------------------------
# serving: a request loop over a resident reference (sd_server)
# ---------------------------------------------------------------------------
class Server:
    """Serves kNN (+ optional SNN) requests against a resident reference Y.
    With ``paused=True`` the workers wait for ``resume()``, so requests
    submitted meanwhile are scheduled as one batch.

    Each request equals ``sparseKNN(X, Y, k, method, dist, drop_degenerate)``
    bit for bit; with ``snn=True`` it also carries the shared-neighbour graph
    of its query columns (queries joined when their neighbour sets in Y
    intersect, Jaccard-weighted, kept when the weight exceeds ``prune``).
    Requests are admitted against ``budget_bytes`` of device memory and cut
    into work items of ``chunk_cols`` columns that ``slots`` workers, each on
    its own stream, take smallest-request-first.  With ``graphs=True`` items
    of at most 256 columns (cosine or binary) replay a HIP graph captured per
    search parameters (off by default: ordinary searches measured faster).
    """

    def __init__(self, Y, slots=4, chunk_cols=4096, budget_bytes=0, max_pending=0,
                 paused=False, graphs=False, *, context: Context | None = None):
        self.ctx = context or default_context()
        self.Y = Y if isinstance(Y, DeviceMatrix) else DeviceMatrix(Y, self.ctx)
        prm = _lib.ServerParams(int(slots), int(chunk_cols), int(budget_bytes), int(max_pending),
                                int(bool(paused)), int(bool(graphs)))
        h = C.c_void_p()
        check(_lib.load().sd_server_create(self.ctx.handle, self.Y.handle, C.byref(prm),
                                           C.byref(h)))
        self._h = h
        self._keep = {}

    def submit(self, X, k=10, method="binary", dist=True, drop_degenerate=True, snn=False,
               prune=0.0, precision="auto", margin=0, borrow=False) -> int:
        """Queue a request; returns its ticket.  Raises SparseDistError with
        code SD_EBUDGET when the server refuses it.  borrow=True lends X's
        arrays to the server instead of copying them (the server keeps a
        reference until the request is released)."""
        _chk_flag(dist, "dist")
        _chk_flag(drop_degenerate, "drop_degenerate")
        k = _chk_count(k, "k")
        method = _match_method(method)
        M = _as_csc(X)
        return self.submit_csc(M.shape[0], M.indptr, M.indices, M.data, k=k, method=method,
                               dist=dist, drop_degenerate=drop_degenerate, snn=snn, prune=prune,
                               precision=precision, margin=margin, borrow=borrow)

    def submit_csc(self, nrow, indptr, indices, data, k=10, method="binary", dist=True,
                   drop_degenerate=True, snn=False, prune=0.0, precision="auto", margin=0,
                   borrow=False) -> int:
        """submit() for a CSC matrix given as arrays (indptr[0] == 0, rows
        ascending within each column)."""
        k = _chk_count(k, "k")
        method = _match_method(method)
        p = np.ascontiguousarray(indptr, dtype=np.int64)
        i = np.ascontiguousarray(indices, dtype=np.int32)
        x = np.ascontiguousarray(data, dtype=np.float64)
        r = _lib.Request()
        r.nrow, r.ncol = int(nrow), p.size - 1
        r.colptr, r.rowind, r.values = p.ctypes.data, i.ctypes.data, x.ctypes.data
        r.knn = _params(k, method, dist, False, drop_degenerate, 256, precision, margin)
        r.snn = int(bool(snn))
        r.borrow = int(bool(borrow))
        r.prune = float(prune)
        t = C.c_int64(0)
        check(_lib.load().sd_server_submit(self._h, C.byref(r), C.byref(t)))
        self._keep[t.value] = (p.size - 1, bool(snn), (p, i, x) if borrow else None)
        return t.value

    def wait(self, ticket, timeout=None) -> dict:
        """Wait for a request (timeout in seconds, None = forever); returns
        its info: done, status, k_eff, warn, queued_ms, run_ms, total_ms."""
        info = _lib.RequestInfo()
        check(_lib.load().sd_server_wait(self._h, ticket, -1.0 if timeout is None else
                                         float(timeout), C.byref(info)))
        return {f: getattr(info, f) for f, _ in info._fields_}

    def result(self, ticket, release=True):
        """The finished request's KNNResult (and SNN graph as ``["snn"]``)."""
        info = self.wait(ticket)
        if info["status"] != _lib.SD_OK:
            msg = _lib.load().sd_server_error(self._h, ticket).decode()
            if release:
                self.release(ticket)
            raise SparseDistError(info["status"], msg)
        nq, want_snn, _ = self._keep[ticket]
        ke = info["k_eff"]
        idx = np.empty(max(nq * ke, 1), dtype=np.int32)
        dst = np.empty(max(nq * ke, 1), dtype=np.float64)
        g = C.c_void_p()
        check(_lib.load().sd_server_result(self._h, ticket, idx.ctypes.data, dst.ctypes.data,
                                           C.byref(g)))
        res = KNNResult(idx[: nq * ke].reshape((nq, ke), order="F"),
                        dst[: nq * ke].reshape((nq, ke), order="F"), False)
        if want_snn and g.value:
            res["snn"] = _graph_to_csc(g, nq)
        res["info"] = info
        if release:
            self.release(ticket)
        return res

    def release(self, ticket):
        check(_lib.load().sd_server_release(self._h, ticket))
        self._keep.pop(ticket, None)

    def resume(self):
        """Let the workers of a server created with ``paused=True`` start."""
        check(_lib.load().sd_server_resume(self._h))

    def stats(self) -> dict:
        s = _lib.ServerStats()
        check(_lib.load().sd_server_stats_get(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in s._fields_}

    def close(self):
        if self._h:
            _lib.load().sd_server_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# one-to-one helpers (the reference's internal natives)
# ---------------------------------------------------------------------------
def sparseDist2(X, Y, method="binary", dist=True, *, context: Context | None = None):
    """Dense slab between the columns of X (rows of the result) and Y
    (reference fastJacc2 / fastCosine2 / fastEuclidean2 / fastManhattan2 /
    fastCorr2 / fastCov2 with m = X, m2 = Y)."""
    method = _match_method(method)
    ctx = context or default_context()
    dx, _kx = _device(X, ctx)
    dy, _ky = _device(Y, ctx)
    D = np.empty((dx.ncol, dy.ncol), dtype=np.float64, order="F")
    check(_lib.load().sd_slab(ctx.handle, _METHOD_CODE[method], dx.handle, dy.handle,
                              int(bool(dist)), D.ctypes.data))
    return D


def topKBlock(D, k, decreasing, self_row, ncores=1, *, context: Context | None = None):
    """Reference topKBlock (src/topK.cpp:67-149): returns (idx, dist)."""
    D = np.asfortranarray(D, dtype=np.float64)
    nrow, ncol = D.shape
    sr = np.ascontiguousarray(self_row, dtype=np.int32).ravel()
    if sr.size != ncol:
        raise ValueError("self_row must have one entry per column of D.")
    if k < 1:
        raise ValueError("k must be >= 1.")
    ctx = context or default_context()
    idx = np.empty((ncol, k), dtype=np.int32, order="F")
    val = np.empty((ncol, k), dtype=np.float64, order="F")
    check(_lib.load().sd_topk_block(ctx.handle, D.ctypes.data, nrow, ncol, int(k),
                                    int(bool(decreasing)), sr.ctypes.data, idx.ctypes.data,
                                    val.ctypes.data))
    return idx, val


def topKRowsAccum(D, k, decreasing, best_idx, best_val, col_offset, self_col, ncores=1,
                  tile_rows=64, *, context: Context | None = None):
    """Reference topKRowsAccum (src/topKRows.cpp:104-248): updates the
    Fortran-ordered nrow x k arrays best_idx / best_val in place."""
    D = np.asfortranarray(D, dtype=np.float64)
    nrow, ncol = D.shape
    sc = np.ascontiguousarray(self_col, dtype=np.int32).ravel()
    if best_idx.shape != (nrow, k) or best_val.shape != (nrow, k):
        raise ValueError("best_idx and best_val must both be nrow(D) x k.")
    if sc.size != nrow:
        raise ValueError("self_col must have one entry per row of D.")
    if not (best_idx.flags.f_contiguous and best_val.flags.f_contiguous
            and best_idx.dtype == np.int32 and best_val.dtype == np.float64):
        raise ValueError("best_idx / best_val must be Fortran-ordered int32 / float64 arrays")
    ctx = context or default_context()
    check(_lib.load().sd_topk_rows_accum(ctx.handle, D.ctypes.data, nrow, ncol, int(k),
                                         int(bool(decreasing)), best_idx.ctypes.data,
                                         best_val.ctypes.data, int(col_offset), sc.ctypes.data))


def snnJaccard(idx, include_self=True, ncores=1, verbose=False, *,
               context: Context | None = None):
    """Reference snnJaccard (src/snn.cpp:198-250): n x k weights aligned with
    idx, NA_REAL where undefined."""
    a = np.asarray(idx)
    if a.ndim != 2 or a.shape[0] < 1 or a.shape[1] < 1:
        raise ValueError("'idx' must have at least one row and column.")
    idx32 = np.asfortranarray(a.astype(np.int32) if a.dtype.kind in "iu" else
                              np.where(np.isnan(a), NA_INTEGER, np.nan_to_num(a)).astype(np.int32))
    n, k = idx32.shape
    ctx = context or default_context()
    w = np.empty((n, k), dtype=np.float64, order="F")
    check(_lib.load().sd_snn_jaccard(ctx.handle, idx32.ctypes.data, n, k, int(bool(include_self)),
                                     w.ctypes.data))
    return w
