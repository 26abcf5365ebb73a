"""The C ABI library loads and exports every function include/marbles_amd.h
declares (no GPU calls)."""
import ctypes
import os
import re

from marbles_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "marbles_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mrb_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("mrb_level_create", "mrb_stream", "mrb_collide", "mrb_fill_ghosts",
                 "mrb_stream_collide", "mrb_reduce_conservation"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert set(declared_functions()) <= set(_lib.SIGNATURES)


def test_errors_are_reported_without_a_gpu():
    lib = _lib.load()
    # invalid descriptors are rejected before any device call
    h = ctypes.c_void_p()
    desc = _lib.LevelDesc()
    desc.nbox = 0
    assert lib.mrb_level_create(ctypes.byref(desc), ctypes.byref(h)) == -1
    assert b"box" in lib.mrb_last_error()
