import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libsd_hip.so kernels)")


def bits(a):
    """uint64 view of float64 data, so NA_real_, NaN and -0.0 compare exactly."""
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64)).view(np.uint64)


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def gpu_ctx():
    import sparsedist_amd as sd
    return sd.default_context()


def self_rows_from_cross(oi, od, qcols, k, include_self):
    """Self-search rows (distances, ascending) of the columns qcols of Y, from
    a cross search of Y[:, qcols] against Y with k + 1 neighbours: the
    query's own column leaves the list, and with include_self it comes back
    at its restored value 0 in (value, index) order (R/sparseKNN.R:270-279)."""
    gi = np.empty((len(qcols), k), np.int32)
    gd = np.empty((len(qcols), k), np.float64)
    na_int = np.iinfo(np.int32).min
    for t, q in enumerate(qcols):
        if (oi[t] == na_int).all():  # degenerate query: an NA row either way
            gi[t] = oi[t][:k]
            gd[t] = od[t][:k]
            continue
        pairs = [(d, i) for i, d in zip(oi[t].tolist(), od[t].tolist()) if i != q + 1]
        if include_self:
            pairs.append((0.0, q + 1))
            pairs.sort()
        gi[t] = [i for _, i in pairs[:k]]
        gd[t] = [d for d, _ in pairs[:k]]
    return gi, gd
