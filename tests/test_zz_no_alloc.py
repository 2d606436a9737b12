"""sd_snn_jaccard_noalloc (the allocation-free form of sd_snn_jaccard, whose
neighbour sets live in a static device array) against sd_snn_jaccard, the
entry point that allocates its scratch: bit-identical weights on two calls in
a row, the second with a smaller input; oversized and invalid inputs are
rejected.  The test allocates the input and output device buffers itself
with the HIP runtime the library links (torch bundles a runtime of its own,
which cannot share the device with it in one process)."""
import ctypes as C

import numpy as np
import pytest

import sparsedist_amd as sd
from sparsedist_amd import _lib
from conftest import bits

pytestmark = pytest.mark.gpu
NA_INT = np.iinfo(np.int32).min
H2D, D2H = 1, 2  # hipMemcpyKind


class DeviceArray:
    """A device copy of a numpy array (hipMalloc / hipMemcpy / hipFree)."""

    def __init__(self, host):
        _lib.load()  # the library's HIP runtime is then the process's
        self.hip = C.CDLL("libamdhip64.so.7")
        self.host = np.ascontiguousarray(host)
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(self.host.nbytes)) == 0
        assert self.hip.hipMemcpy(self.ptr, self.host.ctypes.data_as(C.c_void_p),
                                  C.c_size_t(self.host.nbytes), H2D) == 0

    def get(self):
        out = np.empty_like(self.host)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr,
                                  C.c_size_t(out.nbytes), D2H) == 0
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def index_matrix(n, k, seed):
    """n x k (Fortran order) neighbour lists with NA entries, self entries and
    repeats."""
    rng = np.random.default_rng(seed)
    idx = rng.integers(1, n + 1, size=(n, k)).astype(np.int32)
    idx[rng.random(idx.shape) < 0.05] = NA_INT
    idx[::11, 0] = np.arange(1, n + 1, 11)
    idx[::7, -1] = idx[::7, 0]
    return np.asfortranarray(idx)


def noalloc(ctx, d_idx, n, k, include_self, d_w):
    return _lib.load().sd_snn_jaccard_noalloc(ctx.handle, d_idx.ptr, n, k, int(include_self),
                                             d_w.ptr)


def test_noalloc_weights_match_allocating_entry_point(gpu_ctx):
    ctx = sd.Context(0)
    for n, k, seed in ((60000, 20, 1), (3000, 7, 2)):  # the second call is smaller
        idx = index_matrix(n, k, seed)
        d_idx = DeviceArray(idx.ravel(order="F"))
        for include_self in (True, False):
            want = sd.snnJaccard(idx, include_self=include_self)
            d_w = DeviceArray(np.full(n * k, 123.0))
            rc = noalloc(ctx, d_idx, n, k, include_self, d_w)
            assert rc == _lib.SD_OK, _lib.load().sd_last_error().decode()
            got = d_w.get().reshape((n, k), order="F")
            d_w.free()
            assert np.array_equal(bits(got), bits(want)), (n, k, include_self)
        d_idx.free()
    ctx.close()


def test_noalloc_rejects_oversized_and_invalid_input(gpu_ctx):
    ctx = sd.Context(0)
    d_idx = DeviceArray(np.ones(64, np.int32))
    d_w = DeviceArray(np.zeros(64))
    # 4M rows of 24-entry sets need 400 MB: refused before any work
    rc = noalloc(ctx, d_idx, 1 << 22, 20, True, d_w)
    assert rc == _lib.SD_EINVAL
    assert "bytes of scratch" in _lib.load().sd_last_error().decode()
    # an index outside [1, n] is reported like sd_snn_jaccard reports it
    bad = DeviceArray(np.array([1, 2, 9, 1], np.int32))
    rc = noalloc(ctx, bad, 2, 2, True, d_w)
    assert rc == _lib.SD_EINVAL
    assert "1-based row indices" in _lib.load().sd_last_error().decode()
    for a in (d_idx, d_w, bad):
        a.free()
    ctx.close()
