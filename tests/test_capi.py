"""The C-ABI library loads and exports every entry point include/motex.h declares.

No compute calls: this runs on machines without a GPU.
"""
import ctypes
import os
import re

import motex_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "motex.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(motex_[a-z_]+)\s*\(", hdr)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ["motex_ctx_create", "motex_ctx_destroy", "motex_extract_all_pairs",
              "motex_extract_pair", "motex_submit", "motex_poll", "motex_last_error"]:
        assert s in syms


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(motex_amd.LIB_PATH)
    for s in declared_symbols():
        assert hasattr(L, s), s


def test_binding_covers_every_declared_symbol():
    bound = {name for name, _, _ in motex_amd.SIGNATURES}
    assert set(declared_symbols()) == bound
    motex_amd.lib()  # all argtypes resolve


def test_missing_library_fails_loudly(monkeypatch):
    import importlib
    monkeypatch.setattr(motex_amd, "_lib", None)
    monkeypatch.setattr(motex_amd, "LIB_PATH", "/nonexistent/libmotex.so")
    try:
        motex_amd.lib()
    except ImportError as e:
        assert "missing" in str(e)
    else:
        raise AssertionError("expected ImportError")
