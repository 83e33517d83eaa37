"""The C ABI library loads without a GPU and exports every symbol that
include/*.h declares; the ctypes signatures cover the whole device API."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_symbols():
    names = set()
    for fn in sorted(os.listdir(INCLUDE)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        text = re.sub(r"typedef struct[^;]*?\{.*?\}\s*\w+\s*;", "", text, flags=re.S)
        text = re.sub(r"enum\s*\{.*?\}\s*;", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(\w+)\s*\([^;{]*\)\s*;", text, flags=re.M):
            names.add(m.group(1))
        for m in re.finditer(r"^extern\s+[\w\s\*]+?\b(\w+)\s*;", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_headers_declare_the_expected_surface():
    names = declared_symbols()
    for must in ("ams_spgemm", "ams_spgemm_symbolic", "ams_spgemm_numeric", "ams_batched_run",
                 "ams_spgemm_masked_bwd", "ams_autotune", "hybaccspgemm", "pack_ams_pack",
                 "ams_hybacc_numc", "flp_populate", "scan", "free_csr", "ams_ht_load"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from ams_hip import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libams_hip.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_signatures_cover_the_device_header():
    from ams_hip import _lib
    text = open(os.path.join(INCLUDE, "ams_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    dev = {m.group(1) for m in re.finditer(r"\b(ams_\w+)\s*\(", text)}
    assert dev <= set(_lib.SIGNATURES), dev - set(_lib.SIGNATURES)
    _lib.lib()  # binds every signature


def test_params_struct_layout_matches_c():
    from ams_hip import _lib
    L = _lib.lib()
    p = _lib.AmsParams()
    L.ams_default_params(ctypes.byref(p))
    assert p.mode == 0 and p.acc_thres == 2.0 and p.dense_width % 64 == 0
    assert 1 <= p.sym_hash_max <= 5461 and 1 <= p.num_hash_max <= 5461
