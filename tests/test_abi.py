"""The C ABI library: builds, loads and exports every entry point include/pdgpu.h
declares (no device calls; runs without a GPU)."""
import ctypes
import os
import re

from pd_amd import engine as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "pdgpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pdgpu_\w+)\s*\(", src)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(G.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(G.LIB_PATH)
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_version_string():
    lib = G.load()
    assert b"gfx950" in lib.pdgpu_version()


def test_null_handle_is_rejected():
    lib = G.load()
    assert lib.pdgpu_destroy(None) == -1
    assert lib.pdgpu_build_list(None) == -1
