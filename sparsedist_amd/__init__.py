"""sparsedist_amd -- sparseDist's blocked exact-kNN and SNN path on MI355X.

The compute lives in libsd_hip.so (hand-written HIP for gfx950) behind the C
ABI of include/sd.h; this package binds it with ctypes and mirrors the
reference's R interface (sparseKNN, sparseKNNCross, sparseSNN).
"""
from ._lib import SparseDistError, load  # noqa: F401
from .api import (  # noqa: F401
    NA_INTEGER,
    NA_REAL,
    Context,
    DeviceMatrix,
    KNNResult,
    Server,
    SparseDistWarning,
    default_context,
    detect_cores,
    is_na_real,
    snnJaccard,
    sparseDist2,
    sparseKNN,
    sparseKNNCross,
    sparseSNN,
    topKBlock,
    topKRowsAccum,
    tune_samples,
)

__version__ = "0.1.0"
