"""The C ABI library loads and exports every function include/sd.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import sparsedist_amd
from sparsedist_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "sd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sd_[a-z_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_version_and_error_string_without_gpu():
    L = _lib.load()
    assert b"gfx950" in L.sd_version()
    assert isinstance(L.sd_last_error(), bytes)
    assert L.sd_detect_cores() >= 1


def test_package_exports_reference_api():
    for name in ("sparseKNN", "sparseKNNCross", "sparseSNN", "topKBlock", "topKRowsAccum",
                 "snnJaccard", "sparseDist2"):
        assert callable(getattr(sparsedist_amd, name))
