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


# C types of include/sd.h's struct fields and their sizes on x86-64 / gfx950 hosts
C_SIZES = {"int32_t": 4, "int64_t": 8, "double": 8, "float": 4}
STRUCTS = {"sd_stats": "Stats", "sd_tune_info": "TuneInfo", "sd_knn_params": "KnnParams",
           "sd_server_params": "ServerParams", "sd_request": "Request",
           "sd_request_info": "RequestInfo", "sd_server_stats": "ServerStats",
           "sd_tune_sample": "TuneSample"}


def header_structs():
    """{name: [(field, size)]} for the typedef'd structs of include/sd.h."""
    text = open(os.path.join(ROOT, "include", "sd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for body, name in re.findall(r"typedef struct \{(.*?)\}\s*(sd_[a-z_]+);", text, flags=re.S):
        fields = []
        for decl in body.split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            m = re.match(r"(const )?([a-z_0-9]+)(\*?) (.+)", decl)
            assert m, decl
            ctype, ptr, names = m.group(2), m.group(3), m.group(4)
            for nm in names.split(","):
                nm = nm.strip()
                if nm.startswith("*"):
                    ptr, nm = "*", nm[1:]
                arr = re.match(r"([a-z_0-9]+)\[(\d+)\]$", nm)  # fixed-size array field
                count = 1
                if arr:
                    nm, count = arr.group(1), int(arr.group(2))
                size = 8 if ptr else C_SIZES.get(ctype)
                if size is None:  # a nested struct declared in the same header
                    size = sum(sz for _, sz in out[ctype]) if ctype in out else None
                fields.append((nm, None if size is None else size * count))
        out[name] = fields
    return out


def test_struct_mirrors_match_the_header():
    """Every struct the Python binding passes through the ABI has the
    header's fields in the header's order with the header's sizes, so the
    layouts agree (a drift would corrupt data silently)."""
    hs = header_structs()
    for cname, pyname in STRUCTS.items():
        cls = getattr(_lib, pyname)
        py = [(f, ctypes.sizeof(t)) for f, t in cls._fields_]
        assert [f for f, _ in py] == [f for f, _ in hs[cname]], cname
        for (f, psz), (_, csz) in zip(py, hs[cname]):
            if csz is not None:
                assert psz == csz, (cname, f, psz, csz)

