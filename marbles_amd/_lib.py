"""ctypes binding of libmarbles_amd.so (the C ABI declared in include/marbles_amd.h)."""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int, c_int64, c_void_p

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmarbles_amd.so")

MRB_LAYOUT_FAB = 0
MRB_LAYOUT_COMPACT = 1
MRB_STORAGE_F64 = 0
MRB_STORAGE_F32_SHIFTED = 1
MRB_STORAGE_F16_SCALED = 2
MRB_PLOT_STREAMING = 1
MRB_PLOT_DERIVED = 2
MRB_TAG_GREATER = 0
MRB_TAG_LESS = 1
MRB_TAG_ADJACENT_DIFFERENCE = 2
MRB_ARR_F = 0
MRB_ARR_G = 1
MRB_ARR_ISFLUID = 2

BC_PERIODIC = 0
BC_NOSLIP = 1
BC_VELOCITY = 2
BC_PRESSURE = 3
BC_OUTFLOW = 5
BC_SLIPX = 6
BC_SLIPY = 7
BC_SLIPZ = 8


class LevelDesc(ctypes.Structure):
    _fields_ = [
        ("ncell", c_int * 3),
        ("periodic", c_int * 3),
        ("prob_lo", c_double * 3),
        ("prob_hi", c_double * 3),
        ("nbox", c_int),
        ("box_lo", POINTER(c_int)),
        ("box_hi", POINTER(c_int)),
        ("ng", c_int),
        ("layout", c_int),
        ("nlattice", c_int),
        ("storage", c_int),
    ]


class Phys(ctypes.Structure):
    _fields_ = [
        ("nu", c_double),
        ("alpha", c_double),
        ("adiabatic_exponent", c_double),
        ("m_bar", c_double),
        ("dt", c_double),
        ("mesh_speed", c_double),
    ]


class BC(ctypes.Structure):
    _fields_ = [
        ("bc_type", c_int * 6),
        ("velbc_kind", c_int),
        ("vel_dir", c_int),
        ("mach_ref", c_double),
        ("density", c_double),
        ("temperature", c_double),
        ("adiabatic_exponent", c_double),
        ("m_bar", c_double),
        ("normal_dir", c_int),
    ]


# name -> (restype, argtypes)
SIGNATURES = {
    "mrb_level_create": (c_int, [POINTER(LevelDesc), POINTER(c_void_p)]),
    "mrb_level_destroy": (c_int, [c_void_p]),
    "mrb_set_phys": (c_int, [c_void_p, POINTER(Phys)]),
    "mrb_set_bc": (c_int, [c_void_p, POINTER(BC)]),
    "mrb_set_shift_state": (c_int, [c_void_p, c_double, c_double]),
    "mrb_fab_size": (c_int64, [c_void_p, c_int, c_int]),
    "mrb_upload": (c_int, [c_void_p, c_int, c_void_p]),
    "mrb_write_checkpoint": (c_int, [c_char_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "mrb_checkpoint_header": (c_int, [c_char_p, POINTER(c_int), c_void_p, c_void_p, c_void_p, c_int]),
    "mrb_checkpoint_boxes": (c_int, [c_char_p, c_int, POINTER(c_int), c_void_p, c_void_p, c_int]),
    "mrb_read_checkpoint": (c_int, [c_char_p, c_void_p, c_int]),
    "mrb_write_plotfile": (c_int, [c_char_p, c_void_p, c_double, c_int, c_int]),
    "mrb_write_plotfile_levels": (c_int, [c_char_p, c_void_p, c_int, c_double, c_void_p, c_int]),
    "mrb_plot_nvars": (c_int, [c_int]),
    "mrb_forces_file_open": (c_int, [c_char_p, c_int]),
    "mrb_forces_file_append": (c_int, [c_char_p, c_double, c_void_p]),
    "mrb_stencil": (c_int, [c_void_p, c_void_p, c_void_p]),
    "mrb_cluster_tags": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_double, POINTER(c_int), c_void_p,
                                 c_void_p, c_int]),
    "mrb_amr_regrid_fill": (c_int, [c_void_p, c_void_p, c_void_p]),
    "mrb_tag_level": (c_int, [c_void_p, c_char_p, c_int, c_double, c_void_p]),
    "mrb_plot_fields": (c_int, [c_void_p, c_void_p, c_int]),
    "mrb_download": (c_int, [c_void_p, c_int, c_void_p]),
    "mrb_compute_flags_sphere": (c_int, [c_void_p, POINTER(c_double), c_double, c_int]),
    "mrb_compute_flags_cylinder": (c_int, [c_void_p, POINTER(c_double), c_double, c_double, c_int, c_int]),
    "mrb_finalize_flags": (c_int, [c_void_p]),
    "mrb_init_taylorgreen": (c_int, [c_void_p, c_double, c_double, POINTER(c_double)]),
    "mrb_init_constant": (
        c_int, [c_void_p, c_double, POINTER(c_double), c_double, c_double, c_double]),
    "mrb_fill_ghosts": (c_int, [c_void_p, c_int, c_void_p]),
    "mrb_physbc": (c_int, [c_void_p, c_int, c_void_p]),
    "mrb_stream": (c_int, [c_void_p, c_int, c_void_p]),
    "mrb_collide": (c_int, [c_void_p, c_void_p]),
    "mrb_stream_collide": (c_int, [c_void_p, c_void_p]),
    "mrb_advance": (c_int, [c_void_p, c_int, c_void_p]),
    "mrb_stream_compact": (c_int, [c_void_p, c_void_p]),
    "mrb_advance_timed": (c_int, [c_void_p, c_int, c_void_p, POINTER(c_double), POINTER(c_int)]),
    "mrb_hot_kernel_count": (c_int, []),
    "mrb_hot_kernel_variant_name": (c_char_p, [c_int]),
    "mrb_set_hot_kernel": (c_int, [c_void_p, c_int]),
    "mrb_autotune": (c_int, [c_void_p, c_int, c_void_p, POINTER(c_double), POINTER(c_int)]),
    "mrb_reduce_conservation": (c_int, [c_void_p, POINTER(c_double), c_void_p]),
    "mrb_reduce_conservation_noalloc": (c_int, [c_void_p, c_void_p, c_void_p]),
    "mrb_stream_copy": (c_int, [c_void_p, c_void_p, ctypes.c_int64, c_int, c_void_p]),
    "mrb_last_error": (c_char_p, []),
    "mrb_hot_kernel_name": (c_char_p, [c_void_p]),
    "mrb_population_ptr": (c_int, [c_void_p, c_int, POINTER(c_void_p), POINTER(c_int64)]),
    "mrb_storage_saturations": (c_int, [c_void_p, POINTER(c_int64)]),
    "mrb_storage_shift": (c_int, [c_void_p, c_int, POINTER(c_double)]),
    "mrb_macro_fields": (c_int, [c_void_p, c_void_p, c_void_p]),
    "mrb_vorticity": (c_int, [c_void_p, c_void_p, c_void_p]),
    "mrb_tag_cells": (c_int, [c_void_p, c_void_p, c_int, c_double, c_void_p, c_void_p]),
    "mrb_init_fields": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_double,
                                c_double]),
    "mrb_amr_link": (c_int, [c_void_p, c_void_p]),
    "mrb_coalesce": (c_int, [c_void_p, c_void_p, c_void_p]),
    "mrb_explode": (c_int, [c_void_p, c_void_p, c_int, c_void_p]),
    "mrb_amr_advance": (c_int, [c_void_p, c_void_p, c_int, c_void_p]),
    "mrb_amr_advance_graph": (c_int, [c_void_p, c_void_p, c_int, c_void_p]),
    "mrb_autotune_count": (c_int, [c_void_p]),
    "mrb_eb_force": (c_int, [c_void_p, POINTER(c_double), c_void_p]),
    "mrb_iso_create": (c_int, [POINTER(c_int), c_double, c_double, c_double, POINTER(c_void_p)]),
    "mrb_iso_destroy": (c_int, [c_void_p]),
    "mrb_iso_forward": (c_int, [c_void_p, c_void_p, c_double, c_void_p, c_int, POINTER(c_double), c_void_p]),
    "mrb_iso_gradient": (c_int, [c_void_p, c_void_p, c_double, c_void_p, c_int, c_int, POINTER(c_double),
                                 POINTER(c_double), c_void_p, c_void_p, c_void_p]),
    "mrb_iso_record": (c_int, [c_void_p, c_void_p, c_double, c_void_p, c_int, c_int, POINTER(c_double), c_void_p]),
    "mrb_iso_reverse": (c_int, [c_void_p, POINTER(c_double), c_void_p, c_void_p, c_void_p]),
    "mrb_iso_record_interval": (c_int, [c_void_p, POINTER(c_int)]),
    "mrb_iso_state": (c_int, [c_void_p, POINTER(c_double)]),
    "mrb_autotune_name": (ctypes.c_char_p, [c_void_p, c_int]),
}

_lib = None


MRB_E_UNSUPPORTED = -3


class MrbError(RuntimeError):
    """An error status of the C ABI; `code` is the MRB_E_* value (None when
    the library could not be loaded)."""

    def __init__(self, msg: str, code: int | None = None):
        super().__init__(msg)
        self.code = code


def load() -> ctypes.CDLL:
    """Loads the HIP engine; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MrbError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    # torch ships its own libamdhip64.so.7; in one process the HIP runtime
    # loaded first serves every library that needs that soname, and torch
    # finds no GPU on any runtime but its own.  Loading torch first lets the
    # engine and torch tensors share torch's runtime in any call order.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != 0:
        msg = load().mrb_last_error().decode()
        raise MrbError(f"marbles_amd error {rc}: {msg}", rc)
