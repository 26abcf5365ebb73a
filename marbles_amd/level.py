"""Host-side level object over the C ABI.

Mirrors the per-level surface of the reference solver (NREL/marbles
Source/LBM.H): a level owns f and g with ghost cells, the is_fluid flags and
the boundary functor; `stream`, `collide` and the FillPatchOps calls
`fillpatch`/`physbc` run the same stages as LBM::stream (LBM.cpp:558),
LBM::collide (LBM.cpp:607) and FillPatchOps (FillPatchOps.H:66-159).  On the
compact layout `step`/`advance` run the fused hot path.
"""
from __future__ import annotations

import ctypes
from ctypes import c_double, c_int, c_void_p
from typing import Iterable, Sequence

import numpy as np

from . import _lib
from ._lib import BC, LevelDesc, Phys, check

NQ = 27
AIR_MOLAR_MASS = 28.96

_LAYOUTS = {"fab": _lib.MRB_LAYOUT_FAB, "compact": _lib.MRB_LAYOUT_COMPACT}
_STORAGES = {
    "f64": _lib.MRB_STORAGE_F64,
    "f32s": _lib.MRB_STORAGE_F32_SHIFTED,
    "f16s": _lib.MRB_STORAGE_F16_SCALED,
}
_ARRS = {"f": _lib.MRB_ARR_F, "g": _lib.MRB_ARR_G, "is_fluid": _lib.MRB_ARR_ISFLUID}


def _d3(v) -> ctypes.Array:
    return (c_double * 3)(*[float(x) for x in v])


class Level:
    """One AMR level of f and g populations on the GPU."""

    def __init__(
        self,
        ncell: Sequence[int],
        periodic: Sequence[int],
        prob_lo: Sequence[float] | None = None,
        prob_hi: Sequence[float] | None = None,
        boxes: Iterable[tuple[Sequence[int], Sequence[int]]] | None = None,
        ng: int = 0,
        layout: str = "compact",
        storage: str = "f64",
    ):
        self.lib = _lib.load()
        self.ncell = tuple(int(n) for n in ncell)
        self.periodic = tuple(int(bool(p)) for p in periodic)
        self.prob_lo = tuple(float(x) for x in (prob_lo if prob_lo is not None else (0, 0, 0)))
        self.prob_hi = tuple(float(x) for x in (prob_hi if prob_hi is not None else self.ncell))
        if boxes is None:
            boxes = [((0, 0, 0), tuple(n - 1 for n in self.ncell))]
        self.boxes = [(tuple(int(a) for a in lo), tuple(int(a) for a in hi)) for lo, hi in boxes]
        self.ng = int(ng)
        self.layout = layout
        self.storage = storage
        lo = (c_int * (3 * len(self.boxes)))(*[a for b in self.boxes for a in b[0]])
        hi = (c_int * (3 * len(self.boxes)))(*[a for b in self.boxes for a in b[1]])
        desc = LevelDesc(
            (c_int * 3)(*self.ncell), (c_int * 3)(*self.periodic), _d3(self.prob_lo), _d3(self.prob_hi),
            len(self.boxes), lo, hi, self.ng, _LAYOUTS[layout], 2, _STORAGES[storage])
        h = c_void_p()
        check(self.lib.mrb_level_create(ctypes.byref(desc), ctypes.byref(h)))
        self.h = h
        self.fab_shapes = [
            tuple(hi[d] - lo[d] + 1 + 2 * self.ng for d in (2, 1, 0)) for lo, hi in self.boxes
        ]

    # -- lifetime -----------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.mrb_level_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- parameters ---------------------------------------------------------
    def set_phys(self, nu, alpha=None, adiabatic_exponent=5.0 / 3.0, m_bar=AIR_MOLAR_MASS, dt=1.0,
                 mesh_speed=1.0) -> None:
        p = Phys(nu, nu if alpha is None else alpha, adiabatic_exponent, m_bar, dt, mesh_speed)
        check(self.lib.mrb_set_phys(self.h, ctypes.byref(p)))

    def set_bc(self, bc_lo: Sequence[int], bc_hi: Sequence[int], velocity_bc: dict | None = None) -> None:
        """velocity_bc: {"kind": "constant"|"channel"|"parabolic"|"noop", "dir" (constant),
        "normal_dir", "tangential_dir" (parabolic), "Mach_ref", "initial_density",
        "initial_temperature", "adiabatic_exponent", "mean_molecular_mass"} (VelocityBC.cpp:6-54;
        Mach_ref defaults 0.005, parabolic 0.05)."""
        v = dict(velocity_bc or {})
        kind = {"noop": 0, "constant": 1, "channel": 2, "parabolic": 3}[v.get("kind", "noop")]
        mach = 0.05 if kind == 3 else 0.005
        b = BC((c_int * 6)(*list(bc_lo), *list(bc_hi)), kind,
               int(v.get("tangential_dir" if kind == 3 else "dir", 1)),
               float(v.get("Mach_ref", mach)), float(v.get("initial_density", 1.0)),
               float(v.get("initial_temperature", 1.0 / 3.0)), float(v.get("adiabatic_exponent", 5.0 / 3.0)),
               float(v.get("mean_molecular_mass", AIR_MOLAR_MASS)), int(v.get("normal_dir", 0)))
        check(self.lib.mrb_set_bc(self.h, ctypes.byref(b)))

    def set_shift_state(self, rho0: float, T0: float) -> None:
        check(self.lib.mrb_set_shift_state(self.h, rho0, T0))

    # -- flags and initial state -------------------------------------------
    def flags_sphere(self, center, radius, has_fluid_inside=False) -> None:
        check(self.lib.mrb_compute_flags_sphere(self.h, _d3(center), float(radius), int(has_fluid_inside)))

    def flags_cylinder(self, center, radius, height, direction, has_fluid_inside=False) -> None:
        """EB2 "cylinder" along axis `direction` (eb2.cylinder_* inputs)."""
        check(self.lib.mrb_compute_flags_cylinder(self.h, _d3(center), float(radius), float(height),
                                                  int(direction), int(has_fluid_inside)))

    def flags_all_regular(self) -> None:
        isf = np.zeros(self.fab_elems("is_fluid"), dtype=np.int32)
        off = 0
        for shp in self.fab_shapes:
            n = int(np.prod(shp))
            isf[off:off + n] = 1
            off += 2 * n
        self.upload("is_fluid", isf)
        if self.layout == "fab":
            check(self.lib.mrb_finalize_flags(self.h))
        else:
            check(self.lib.mrb_finalize_flags(self.h))

    def flags_upload(self, isf: np.ndarray) -> None:
        self.upload("is_fluid", np.ascontiguousarray(isf, dtype=np.int32))
        check(self.lib.mrb_finalize_flags(self.h))

    def init_fields(self, rho, u, v, w, T, adiabatic_exponent=5.0 / 3.0, m_bar=AIR_MOLAR_MASS) -> None:
        """Initial state from per-cell fields over every fab cell (box order,
        x fastest), like an ic::InitializerBase plugin (IC.H:427-531)."""
        n = sum(int(np.prod(shp)) for shp in self.fab_shapes)
        arrs = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (rho, u, v, w, T)]
        for a in arrs:
            if a.size != n:
                raise ValueError(f"field has {a.size} cells, the level's fabs hold {n}")
        check(self.lib.mrb_init_fields(self.h, *[a.ctypes.data_as(c_void_p) for a in arrs],
                                       float(adiabatic_exponent), float(m_bar)))

    def init_taylorgreen(self, rho0=1.0, v0=1.0, omega=(1.0, 1.0, 1.0)) -> None:
        check(self.lib.mrb_init_taylorgreen(self.h, rho0, v0, _d3(omega)))

    def init_constant(self, density=1.0, mach=(0.0, 0.0, 0.0), temperature=1.0 / 3.0,
                      adiabatic_exponent=5.0 / 3.0, m_bar=AIR_MOLAR_MASS) -> None:
        check(self.lib.mrb_init_constant(self.h, density, _d3(mach), temperature, adiabatic_exponent, m_bar))

    # -- data ---------------------------------------------------------------
    def fab_elems(self, which: str) -> int:
        per = 2 if which == "is_fluid" else NQ
        return sum(per * int(np.prod(s)) for s in self.fab_shapes)

    def upload(self, which: str, arr: np.ndarray) -> None:
        dt = np.int32 if which == "is_fluid" else np.float64
        a = np.ascontiguousarray(arr, dtype=dt).ravel()
        if a.size != self.fab_elems(which):
            raise ValueError(f"{which}: expected {self.fab_elems(which)} elements, got {a.size}")
        check(self.lib.mrb_upload(self.h, _ARRS[which], a.ctypes.data_as(c_void_p)))

    def download(self, which: str) -> np.ndarray:
        dt = np.int32 if which == "is_fluid" else np.float64
        out = np.empty(self.fab_elems(which), dtype=dt)
        check(self.lib.mrb_download(self.h, _ARRS[which], out.ctypes.data_as(c_void_p)))
        return out

    def fabs(self, which: str) -> list[np.ndarray]:
        """Per-box arrays shaped (ncomp, nz, ny, nx) including ghost cells."""
        flat = self.download(which)
        ncomp = 2 if which == "is_fluid" else NQ
        out, off = [], 0
        for shp in self.fab_shapes:
            n = ncomp * int(np.prod(shp))
            out.append(flat[off:off + n].reshape((ncomp,) + shp))
            off += n
        return out

    def valid(self, which: str) -> np.ndarray:
        """Valid cells assembled over the domain, shaped (ncomp, nz, ny, nx)."""
        ncomp = 2 if which == "is_fluid" else NQ
        dom = np.zeros((ncomp,) + tuple(reversed(self.ncell)),
                       dtype=np.int32 if which == "is_fluid" else np.float64)
        g = self.ng
        for (lo, hi), fab in zip(self.boxes, self.fabs(which)):
            dom[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = \
                fab[:, g:fab.shape[1] - g, g:fab.shape[2] - g, g:fab.shape[3] - g]
        return dom

    # -- stages -------------------------------------------------------------
    def fillpatch(self, lattice: int, stream=None) -> None:
        check(self.lib.mrb_fill_ghosts(self.h, lattice, stream))

    def physbc(self, lattice: int, stream=None) -> None:
        check(self.lib.mrb_physbc(self.h, lattice, stream))

    def stream(self, lattice: int, stream=None) -> None:
        check(self.lib.mrb_stream(self.h, lattice, stream))

    def collide(self, stream=None) -> None:
        check(self.lib.mrb_collide(self.h, stream))

    def step(self, stream=None) -> None:
        check(self.lib.mrb_stream_collide(self.h, stream))

    def advance(self, nsteps: int, stream=None) -> None:
        check(self.lib.mrb_advance(self.h, int(nsteps), stream))

    def advance_timed(self, nsteps: int, stream=None) -> dict:
        """Fused steps with HIP events around each kernel; returns {kernel: total ms}."""
        ms = (c_double * 4)()
        ns = c_int()
        check(self.lib.mrb_advance_timed(self.h, int(nsteps), stream, ms, ctypes.byref(ns)))
        names = self.hot_kernel_name().split("+")
        return {names[k] if k < len(names) else f"k{k}": ms[k] for k in range(ns.value)}

    @staticmethod
    def hot_kernel_variants() -> list[str]:
        lib = _lib.load()
        return [lib.mrb_hot_kernel_variant_name(i).decode() for i in range(lib.mrb_hot_kernel_count())]

    def set_hot_kernel(self, name_or_index) -> None:
        idx = name_or_index if isinstance(name_or_index, int) else self.hot_kernel_variants().index(name_or_index)
        check(self.lib.mrb_set_hot_kernel(self.h, idx))

    def autotune(self, reps: int = 5, stream=None) -> dict:
        """Times every kernel variant without advancing the state; selects the fastest.
        Returns {variant name: mean ms per step}."""
        names = [self.lib.mrb_autotune_name(self.h, i).decode()
                 for i in range(self.lib.mrb_autotune_count(self.h))]
        ms = (c_double * len(names))()
        best = c_int()
        check(self.lib.mrb_autotune(self.h, int(reps), stream, ms, ctypes.byref(best)))
        return {n: ms[i] for i, n in enumerate(names)}

    def stream_compact(self, stream=None) -> None:
        check(self.lib.mrb_stream_compact(self.h, stream))

    def conservation(self, stream=None) -> np.ndarray:
        out = (c_double * 5)()
        check(self.lib.mrb_reduce_conservation(self.h, out, stream))
        return np.array(out[:])

    def eb_force(self, stream=None) -> np.ndarray:
        """Momentum-exchange force on the embedded boundary (LBM::compute_eb_forces)."""
        out = (c_double * 3)()
        check(self.lib.mrb_eb_force(self.h, out, stream))
        return np.array(out[:])

    def macro_fields(self, out, stream=None) -> None:
        """Density, velocity and temperature per cell into a device tensor of
        5 * ncells doubles (compact layout)."""
        n = int(np.prod(self.ncell))
        if out.numel() != 5 * n or out.dtype.itemsize != 8 or not out.is_cuda:
            raise ValueError("out must be a device float64 tensor of 5 * ncells elements")
        check(self.lib.mrb_macro_fields(self.h, c_void_p(out.data_ptr()), stream))

    def vorticity(self, out, stream=None) -> None:
        """Vorticity x, y, z and magnitude per cell (LBM::compute_derived) into a
        device tensor of 4 * ncells doubles (compact layout)."""
        n = int(np.prod(self.ncell))
        if out.numel() != 4 * n or out.dtype.itemsize != 8 or not out.is_cuda:
            raise ValueError("out must be a device float64 tensor of 4 * ncells elements")
        check(self.lib.mrb_vorticity(self.h, c_void_p(out.data_ptr()), stream))

    def plot_fields(self, streaming: bool = False, derived: bool = False) -> np.ndarray:
        """The plotfile variables of LBM::plot_file_mf as a host array (nvars, nz,
        ny, nx) over the domain: macrodata of the last step's post-stream
        state, optionally f and g, optionally vorticity and dQCorr, then
        is_fluid and eb_boundary."""
        flags = (_lib.MRB_PLOT_STREAMING if streaming else 0) | (_lib.MRB_PLOT_DERIVED if derived else 0)
        nv = self.lib.mrb_plot_nvars(flags)
        shapes = [tuple(h - l + 1 for l, h in zip(reversed(lo), reversed(hi))) for lo, hi in self.boxes]
        flat = np.empty(nv * sum(int(np.prod(s)) for s in shapes))
        check(self.lib.mrb_plot_fields(self.h, flat.ctypes.data_as(c_void_p), flags))
        out = np.zeros((nv,) + tuple(reversed(self.ncell)))
        off = 0
        for (lo, hi), sh in zip(self.boxes, shapes):
            n = int(np.prod(sh))
            out[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = flat[off:off + nv * n].reshape((nv,) + sh)
            off += nv * n
        return out

    def write_plotfile(self, path: str, time: float, step: int, streaming: bool = False,
                       derived: bool = False) -> None:
        """LBM::write_plot_file: an AMReX plotfile of this level at `path`."""
        import os

        flags = (_lib.MRB_PLOT_STREAMING if streaming else 0) | (_lib.MRB_PLOT_DERIVED if derived else 0)
        check(self.lib.mrb_write_plotfile(os.fsencode(path), self.h, float(time), int(step), flags))

    def tag(self, field: str, criterion: str, value: float) -> np.ndarray:
        """LBM::ErrorEst: uint8 tags (nz, ny, nx) over the domain where the plot
        variable `field` (e.g. "vel_mag", "vort_mag") meets the criterion
        ('greater', 'less' or 'adjacent_difference'); the boxes must cover the
        domain."""
        crit = {"greater": _lib.MRB_TAG_GREATER, "less": _lib.MRB_TAG_LESS,
                "adjacent_difference": _lib.MRB_TAG_ADJACENT_DIFFERENCE}[criterion]
        out = np.zeros(tuple(reversed(self.ncell)), dtype=np.uint8)
        check(self.lib.mrb_tag_level(self.h, field.encode(), crit, float(value), out.ctypes.data_as(c_void_p)))
        return out

    def tag_cells(self, field, criterion: str, value: float, tags, stream=None) -> None:
        """Refinement tags (uint8 device tensor of ncells) from a device float64
        field: criterion 'greater', 'less' or 'adjacent_difference'."""
        n = int(np.prod(self.ncell))
        if field.numel() != n or tags.numel() != n or tags.dtype.itemsize != 1:
            raise ValueError("field: ncells doubles, tags: ncells bytes")
        crit = {"greater": _lib.MRB_TAG_GREATER, "less": _lib.MRB_TAG_LESS,
                "adjacent_difference": _lib.MRB_TAG_ADJACENT_DIFFERENCE}[criterion]
        check(self.lib.mrb_tag_cells(self.h, c_void_p(field.data_ptr()), crit, float(value),
                                     c_void_p(tags.data_ptr()), stream))

    def storage_saturations(self) -> int:
        """FP16 storage: values clamped to the FP16 maximum so far (0 otherwise)."""
        n = ctypes.c_int64()
        check(self.lib.mrb_storage_saturations(self.h, ctypes.byref(n)))
        return int(n.value)

    def shift(self, which: str) -> np.ndarray:
        """Rest-state values subtracted by the reduced-precision storages."""
        out = (c_double * NQ)()
        check(self.lib.mrb_storage_shift(self.h, _ARRS[which], out))
        return np.array(out[:])

    def hot_kernel_name(self) -> str:
        return self.lib.mrb_hot_kernel_name(self.h).decode()

    def population_ptr(self, lattice: int) -> tuple[int, int]:
        p = c_void_p()
        n = ctypes.c_int64()
        check(self.lib.mrb_population_ptr(self.h, lattice, ctypes.byref(p), ctypes.byref(n)))
        return int(p.value or 0), int(n.value)
