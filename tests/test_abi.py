"""The C-ABI library loads and exports every entry point declared in include/rs3d.h
(no compute call: this runs without a GPU)."""
import ctypes
import os
import re

from rayspace3d_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "rs3d.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b(\w+)\s*\(", txt, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while", "sizeof")))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    for must in ("rs3d_bvh_build", "rs3d_trace_points", "rs3d_filter_refine", "rs3d_overlap_join",
                 "merge_and_deduplicate_pairs_gpu", "compact_hash_table_pairs",
                 "exclusive_scan_gpu", "count_hits_gpu", "compact_hits_gpu"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the same set
    assert set(declared_functions()) <= set(_lib.SIGNATURES)
    _lib.load()


def test_error_path_without_device():
    lib = _lib.load()
    handle = ctypes.c_void_p()
    rc = lib.rs3d_ctx_create(0, None)
    assert rc == _lib.RS3D_E_ARG
    assert b"out is null" in lib.rs3d_last_error()
    assert handle.value is None
