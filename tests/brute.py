"""Independent brute-force definitions used to pin the oracle and the GPU.

These follow the reference test suite's own executable definitions
(ref_topk, tests/testthat/test-sparseKNN.R:16-36; ref_snn,
tests/testthat/test-sparseSNN.R:14-35; ref_shared,
tests/testthat/test-sparseSNN-shared.R:54-69) and base-R / numpy closed forms
(tests/testthat/test-sparseDist.R:155-247), written with numpy on dense
arrays -- no merge walks, no heaps -- so a slip in the oracle's loops shows up
as a disagreement here.
"""
import numpy as np
import scipy.sparse as sp
from scipy.spatial.distance import cdist

NA_INT = np.iinfo(np.int32).min


def dense(M):
    return M.toarray() if sp.issparse(M) else np.asarray(M, dtype=np.float64)


def jaccard(A, B, dist=True):
    """Exact: integer intersection and union counts, one FP64 division."""
    a = (dense(A) != 0).astype(np.int64)
    b = (dense(B) != 0).astype(np.int64)
    inter = a.T @ b
    union = a.sum(0)[:, None] + b.sum(0)[None, :] - inter
    with np.errstate(invalid="ignore", divide="ignore"):
        sim = np.where(union > 0, inter / np.where(union > 0, union, 1), 1.0)
    return 1.0 - sim if dist else sim


def cosine(A, B, dist=True):
    a, b = dense(A), dense(B)
    na, nb = np.sqrt((a * a).sum(0)), np.sqrt((b * b).sum(0))
    with np.errstate(invalid="ignore", divide="ignore"):
        c = (a.T @ b) / np.outer(na, nb)
    ea, eb = na == 0, nb == 0
    c = np.where(ea[:, None] | eb[None, :], np.where(ea[:, None] & eb[None, :], 1.0, 0.0), c)
    c = np.clip(np.nan_to_num(c, nan=0.0), -1, 1)
    return 1.0 - c if dist else c


def euclidean(A, B):
    return cdist(dense(A).T, dense(B).T, "euclidean")


def manhattan(A, B):
    return cdist(dense(A).T, dense(B).T, "cityblock")


def pearson(A, B, dist=True):
    a, b = dense(A), dense(B)
    ac, bc = a - a.mean(0), b - b.mean(0)
    with np.errstate(invalid="ignore", divide="ignore"):
        c = (ac.T @ bc) / np.outer(np.sqrt((ac * ac).sum(0)), np.sqrt((bc * bc).sum(0)))
    c = np.clip(np.where(np.isfinite(c), c, 0.0), -1, 1)
    return 1.0 - c if dist else c


def covariance(A, B):
    a, b = dense(A), dense(B)
    n = a.shape[0]
    return ((a - a.mean(0)).T @ (b - b.mean(0))) / max(n - 1, 1)


def slab(method, A, B, dist=True):
    return {"binary": lambda: jaccard(A, B, dist), "cosine": lambda: cosine(A, B, dist),
            "euclidean": lambda: euclidean(A, B), "manhattan": lambda: manhattan(A, B),
            "pearson": lambda: pearson(A, B, dist),
            "covariance": lambda: covariance(A, B)}[method]()


def ref_topk(M, k, decreasing, exclude_self):
    """Reference ref_topk: candidates are rows of M, one result row per
    column; finite only; order by (value, index), lower index first."""
    nref, nq = M.shape
    idx = np.full((nq, k), NA_INT, dtype=np.int32)
    val = np.full((nq, k), np.nan)
    for q in range(nq):
        v = M[:, q]
        cand = np.arange(nref)
        if exclude_self:
            cand = cand[cand != q]
        cand = cand[np.isfinite(v[cand])]
        key = -v[cand] if decreasing else v[cand]
        order = cand[np.lexsort((cand, key))]
        take = order[:k]
        idx[q, :take.size] = take + 1
        val[q, :take.size] = v[take]
    return idx, val


def neighbour_sets(idx, include_self=True):
    n = idx.shape[0]
    sets = []
    for i in range(n):
        s = [int(v) for v in idx[i] if v != NA_INT]
        s = [v for v in s if v != i + 1]
        if include_self:
            s.append(i + 1)
        sets.append(sorted(set(s)))
    return sets


def ref_snn(idx, include_self=True):
    n, k = idx.shape
    sets = neighbour_sets(idx, include_self)
    w = np.full((n, k), np.nan)
    for i in range(n):
        if not sets[i]:
            continue
        for t in range(k):
            v = idx[i, t]
            if v == NA_INT or not sets[v - 1]:
                continue
            a, b = set(sets[i]), set(sets[v - 1])
            w[i, t] = len(a & b) / len(a | b)
    return w


def ref_knn_graph(idx, include_self=True, prune=0.0, symmetrise=True):
    """Dense n x n graph: kNN edge weights, strict prune, union symmetrise."""
    n, k = idx.shape
    w = ref_snn(idx, include_self)
    g = np.zeros((n, n))
    for i in range(n):
        for t in range(k):
            if idx[i, t] != NA_INT and not np.isnan(w[i, t]) and w[i, t] > prune:
                g[i, idx[i, t] - 1] = w[i, t]
    if symmetrise:
        g = np.where(g > 0, g, g.T)
    return g


def ref_shared(idx, include_self=True, prune=0.0):
    n = idx.shape[0]
    sets = neighbour_sets(idx, include_self)
    out = np.zeros((n, n))
    for i in range(n - 1):
        for j in range(i + 1, n):
            a, b = set(sets[i]), set(sets[j])
            if not a or not b:
                continue
            inter = len(a & b)
            if inter == 0:
                continue
            w = inter / len(a | b)
            if w > prune:
                out[i, j] = out[j, i] = w
    return out
