"""Files of the reference solver for engine levels, written and read by
libmarbles_amd.so: checkpoints (LBM::write_checkpoint_file and
LBM::read_checkpoint_file, Source/LBM.cpp:1684-1860: AMReX's checkpoint
Header plus the f and g MultiFabs in VisMF format) and the forces file
(LBM::open_forces_file / output_forces_file, LBM.cpp:1925-1969)."""
from __future__ import annotations

import ctypes
import os
from ctypes import c_double, c_int, c_void_p
from typing import Sequence

from . import _lib
from ._lib import check


def _levels_array(levels):
    return (c_void_p * len(levels))(*[L.h.value if isinstance(L.h, c_void_p) else L.h for L in levels])


def write_checkpoint(path: str, levels: Sequence, steps: Sequence[int], dts: Sequence[float],
                     times: Sequence[float]) -> None:
    """Writes <path>/Header and <path>/Level_<l>/{f,g}_00_{H,D_00000}
    (LBM::write_checkpoint_file) for the levels, coarsest first."""
    lib = _lib.load()
    n = len(levels)
    check(lib.mrb_write_checkpoint(os.fsencode(path), _levels_array(levels), n, (c_int * n)(*steps),
                                   (c_double * n)(*dts), (c_double * n)(*times)))


def checkpoint_info(path: str) -> dict:
    """The checkpoint's level count, per level step, dt and time, and the
    BoxArrays (lists of (lo, hi) cell triples): what a restart needs to
    create its levels before read_checkpoint."""
    lib = _lib.load()
    p = os.fsencode(path)
    nl = c_int()
    check(lib.mrb_checkpoint_header(p, ctypes.byref(nl), None, None, None, 0))
    n = nl.value
    steps, dts, times = (c_int * n)(), (c_double * n)(), (c_double * n)()
    check(lib.mrb_checkpoint_header(p, ctypes.byref(nl), steps, dts, times, n))
    boxes = []
    for lev in range(n):
        nb = c_int()
        check(lib.mrb_checkpoint_boxes(p, lev, ctypes.byref(nb), None, None, 0))
        lo, hi = (c_int * (3 * nb.value))(), (c_int * (3 * nb.value))()
        check(lib.mrb_checkpoint_boxes(p, lev, ctypes.byref(nb), lo, hi, nb.value))
        boxes.append([(tuple(lo[3 * b:3 * b + 3]), tuple(hi[3 * b:3 * b + 3])) for b in range(nb.value)])
    return {"nlevels": n, "steps": list(steps), "dts": list(dts), "times": list(times), "boxes": boxes}


def read_checkpoint(path: str, levels: Sequence) -> None:
    """Loads f and g of every level (created with the checkpoint's boxes,
    flags already set; solid cells get 0 as fill_f_inside_eb gives them)."""
    lib = _lib.load()
    check(lib.mrb_read_checkpoint(os.fsencode(path), _levels_array(levels), len(levels)))


def forces_file_open(path: str, initialize: bool = True) -> None:
    """Creates the forces file with its header (or keeps an existing one when
    initialize is False, as a restart does)."""
    check(_lib.load().mrb_forces_file_open(os.fsencode(path), int(bool(initialize))))


def forces_file_append(path: str, time: float, force) -> None:
    """Appends one line: the time and the EB force components."""
    f = (c_double * 3)(*[float(x) for x in force])
    check(_lib.load().mrb_forces_file_append(os.fsencode(path), float(time), f))
