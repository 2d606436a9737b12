"""Argument validation of the host mirror: the reference's error and warning
text (R/utils-validate.R:17-42, R/sparseKNN.R:114-190,
R/sparseKNNCross.R:56-99, R/sparseSNN.R:125-197).  All checks raise before
any device work, so these run without a GPU."""
import numpy as np
import pytest
import scipy.sparse as sp

from sparsedist_amd import KNNResult, sparseKNN, sparseKNNCross, sparseSNN

X = sp.random(20, 8, density=0.3, random_state=1, format="csc")
Y = sp.random(20, 5, density=0.3, random_state=2, format="csc")


@pytest.mark.parametrize("kw,pattern", [
    (dict(k=0), "positive"), (dict(k=-1), "positive"), (dict(block_size=0), "positive"),
    (dict(k=2.5), "positive integer"), (dict(block_size=4.5), "positive integer"),
    (dict(ncores=1.7), "non-negative integer"), (dict(ncores=float("inf")), "non-negative integer"),
    (dict(k=float("inf")), "positive integer"), (dict(k=[2, 3]), "positive integer"),
    (dict(k="3"), "positive integer"), (dict(k=3e9), "integer.max"),
    (dict(block_size=3e9), "integer.max"), (dict(ncores=3e9), "integer.max"),
    (dict(ncores=float("nan")), "non-negative integer"), (dict(k=True), "positive integer"),
])
def test_counts(kw, pattern):
    with pytest.raises(ValueError, match=pattern):
        sparseKNN(X, **kw, verbose=False)


@pytest.mark.parametrize("bad", [None, 1, "TRUE", [True, False], np.array([])])
@pytest.mark.parametrize("name", ["dist", "include_self", "drop_degenerate", "verbose"])
def test_flags(bad, name):
    with pytest.raises(ValueError, match="TRUE or FALSE"):
        sparseKNN(X, **{name: bad})


@pytest.mark.parametrize("m", ["euclidean", "manhattan", "js"])
def test_dist_false_only_meaningful(m):
    with pytest.raises(ValueError, match="only meaningful"):
        sparseKNN(X, k=3, method=m, dist=False, verbose=False)


def test_rows_must_match():
    with pytest.raises(ValueError, match="same number of rows"):
        sparseKNN(X, sp.random(19, 4, density=0.3, format="csc"), k=3, verbose=False)
    with pytest.raises(ValueError, match="same number of rows"):
        sparseKNNCross(X, sp.random(19, 4, density=0.3, format="csc"), k=3, verbose=False)


def test_cross_requires_y_and_validates():
    with pytest.raises(ValueError, match="'Y' is required"):
        sparseKNNCross(X, None)
    for kw, pat in [(dict(k=2.5), "'k' must be a positive integer"),
                    (dict(block_size=0), "'block_size' must be a positive integer"),
                    (dict(ncores=-1), "'ncores' must be a non-negative integer"),
                    (dict(dist=1), "'dist' must be TRUE or FALSE"),
                    (dict(drop_degenerate=1), "'drop_degenerate' must be TRUE or FALSE"),
                    (dict(verbose=1), "'verbose' must be TRUE or FALSE")]:
        with pytest.raises(ValueError, match=pat):
            sparseKNNCross(X, Y, **kw)


def test_snn_input_contract():
    idx = np.array([[2, 3], [1, 3], [1, 2]], dtype=np.int32)
    with pytest.raises(ValueError, match="self-search"):
        sparseSNN(KNNResult(idx, np.zeros((3, 2)), False))
    with pytest.raises(ValueError, match="self_search"):
        sparseSNN(KNNResult(idx, np.zeros((3, 2)), None))
    with pytest.raises(ValueError, match="self_search"):
        sparseSNN({"idx": idx})
    with pytest.raises(ValueError, match="matrix"):
        sparseSNN("nope")
    with pytest.raises(ValueError, match="whole-number"):
        sparseSNN(np.array([[1.5]]))
    with pytest.raises(ValueError, match="finite"):
        sparseSNN(np.array([[np.inf]]))
    with pytest.raises(ValueError, match="numeric or integer"):
        sparseSNN(np.array([["1"]]))
    with pytest.raises(ValueError, match="1-based"):
        sparseSNN(np.array([[0]], dtype=np.int32))
    with pytest.raises(ValueError, match="1-based"):
        sparseSNN(np.array([[2]], dtype=np.int32))
    for bad in (-0.1, 1.5, float("inf"), float("-inf"), float("nan"), None):
        with pytest.raises(ValueError, match="between 0 and 1"):
            sparseSNN(idx, prune=bad)
    with pytest.raises(ValueError, match="TRUE or FALSE"):
        sparseSNN(idx, include_self=None)
    with pytest.raises(ValueError, match="TRUE or FALSE"):
        sparseSNN(idx, symmetrise=1)
    with pytest.raises(ValueError, match="non-negative integer"):
        sparseSNN(idx, ncores=1.5)
