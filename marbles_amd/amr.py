"""Two-level hierarchy over the C ABI.

Mirrors the level loop of the reference solver (NREL/marbles
Source/LBM.cpp:398-544, evolve -> time_step -> advance) for refinement
ratio 2: the fine level runs two substeps with dt/2 per coarse step, fine
ghost cells are filled from the coarse level once per coarse step
(FillPatchOps.H:118-129) and the coarse level takes the masked average of
the fine one after its stream (LBM.cpp:1571-1584, Utilities.cpp:5-28).
"""
from __future__ import annotations

from . import _lib
from ._lib import check
from .level import Level


class Hierarchy:
    """Coarse level covering the domain plus one fine level (ratio 2)."""

    def __init__(self, cfg: dict, ng: int = 3):
        from . import configs

        self.cfg = cfg
        nc = tuple(cfg["ncell"])
        nf = tuple(2 * n for n in nc)
        self.crse = Level(nc, cfg["periodic"], cfg["prob_lo"], cfg["prob_hi"], boxes=cfg["coarse_boxes"],
                          ng=ng, layout="fab")
        self.fine = Level(nf, cfg["periodic"], cfg["prob_lo"], cfg["prob_hi"], boxes=cfg["fine_boxes"],
                          ng=ng, layout="fab")
        gamma = cfg["ic"].get("adiabatic_exponent", 5.0 / 3.0)
        for lvl, dt in ((self.crse, 1.0), (self.fine, 0.5)):
            lvl.set_phys(cfg["nu"], cfg.get("alpha"), dt=dt)
            lvl.set_bc(cfg["bc_lo"], cfg["bc_hi"], cfg.get("velocity_bc"))
            configs.set_eb_flags(lvl, cfg.get("eb", {"type": "all_regular"}))
            lvl.init_fields(*configs.level_fields(cfg, lvl.ncell, lvl.boxes, ng), adiabatic_exponent=gamma)
        check(_lib.load().mrb_amr_link(self.crse.h, self.fine.h))

    @property
    def levels(self):
        return (self.crse, self.fine)

    def regrid(self, tags=None, boxes=None, n_error_buf: int = 1, blocking_factor: int = 8,
               max_grid_size: int = 32, grid_eff: float = 0.7) -> list:
        """AmrCore::regrid -> LBM::RemakeLevel for the fine level: new fine
        boxes from refinement tags on the coarse domain (uint8 array
        (nz, ny, nx); cluster_tags) or given directly, the new level filled
        from the old one and the coarse interpolation, then linked; returns
        the new boxes."""
        from . import configs

        if boxes is None:
            boxes = cluster_tags(tags, n_error_buf, blocking_factor, max_grid_size, grid_eff)
        if not boxes:
            raise ValueError("regrid: no fine boxes")
        cfg, old = self.cfg, self.fine
        new = Level(old.ncell, cfg["periodic"], cfg["prob_lo"], cfg["prob_hi"], boxes=boxes, ng=old.ng,
                    layout="fab")
        new.set_phys(cfg["nu"], cfg.get("alpha"), dt=0.5)
        new.set_bc(cfg["bc_lo"], cfg["bc_hi"], cfg.get("velocity_bc"))
        configs.set_eb_flags(new, cfg.get("eb", {"type": "all_regular"}))
        lib = _lib.load()
        check(lib.mrb_amr_regrid_fill(self.crse.h, old.h, new.h))
        check(lib.mrb_amr_link(self.crse.h, new.h))
        old.close()
        self.fine = new
        return new.boxes

    def updates_per_step(self) -> int:
        """Lattice updates per coarse step: N_coarse + 2 N_fine."""
        def cells(lvl):
            return sum((hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1) for lo, hi in lvl.boxes)
        return cells(self.crse) + 2 * cells(self.fine)

    def explode(self, lattice: int, stream=None) -> None:
        check(_lib.load().mrb_explode(self.crse.h, self.fine.h, int(lattice), stream))

    def coalesce(self, stream=None) -> None:
        check(_lib.load().mrb_coalesce(self.fine.h, self.crse.h, stream))

    def write_plotfile(self, path: str, time: float, steps, streaming: bool = False, derived: bool = False) -> None:
        """LBM::write_plot_file of both levels (AMReX multi-level plotfile);
        steps = (coarse steps, fine steps)."""
        import ctypes
        import os

        flags = (_lib.MRB_PLOT_STREAMING if streaming else 0) | (_lib.MRB_PLOT_DERIVED if derived else 0)
        lv = (ctypes.c_void_p * 2)(self.crse.h.value, self.fine.h.value)
        st = (ctypes.c_int * 2)(*[int(x) for x in steps])
        check(_lib.load().mrb_write_plotfile_levels(os.fsencode(path), lv, 2, float(time), st, flags))

    def evolve(self, nsteps: int, regrid_int: int = 0, tagging=(), n_error_buf: int = 1,
               blocking_factor: int = 8, max_grid_size: int = 32, grid_eff: float = 0.7, stream=None) -> None:
        """LBM::evolve -> time_step(0) (LBM.cpp:398-512): before a coarse step
        whose completed-step count is a positive multiple of regrid_int the
        fine level is regridded (LBM.cpp:455-480) from the refinement
        indicators `tagging` -- dicts with the reference's keys field_name
        and one of value_greater, value_less, adjacent_difference_greater
        (read_tagging_parameters, LBM.cpp:302-396), their tags OR-ed as
        ErrorEst sets them.  With no tag at all the fine level is kept."""
        import numpy as np

        crit = {"value_greater": "greater", "value_less": "less", "adjacent_difference_greater": "adjacent_difference"}
        for _ in range(int(nsteps)):
            step = getattr(self, "step", 0)
            if regrid_int > 0 and step > getattr(self, "last_regrid", 0) and step % regrid_int == 0:
                tags = np.zeros(tuple(reversed(self.crse.ncell)), dtype=np.uint8)
                for ind in tagging:
                    (key, value), = [(k, v) for k, v in ind.items() if k in crit]
                    tags |= self.crse.tag(ind["field_name"], crit[key], float(value))
                if tags.any():
                    self.regrid(tags=tags, n_error_buf=n_error_buf, blocking_factor=blocking_factor,
                                max_grid_size=max_grid_size, grid_eff=grid_eff)
                self.last_regrid = step
            self.advance(1, stream=stream)
            self.step = step + 1

    def advance(self, nsteps: int, stream=None, graph: bool = False) -> None:
        """nsteps coarse steps; graph=True replays a captured hipGraph per step."""
        fn = _lib.load().mrb_amr_advance_graph if graph else _lib.load().mrb_amr_advance
        check(fn(self.crse.h, self.fine.h, int(nsteps), stream))

    def close(self) -> None:
        # the fine level holds maps into the coarse one: release it first
        self.fine.close()
        self.crse.close()


def cluster_tags(tags, n_error_buf: int = 1, blocking_factor: int = 8, max_grid_size: int = 32,
                 grid_eff: float = 0.7) -> list:
    """Fine boxes (ratio 2; (lo, hi) cell triples) covering the tagged cells
    of a coarse-domain uint8 array (nz, ny, nx) (mrb_cluster_tags)."""
    import ctypes

    import numpy as np

    t = np.ascontiguousarray(tags, dtype=np.uint8)
    ncell = (ctypes.c_int * 3)(t.shape[2], t.shape[1], t.shape[0])
    lib = _lib.load()
    nb = ctypes.c_int()
    args = (int(n_error_buf), int(blocking_factor), int(max_grid_size), float(grid_eff))
    check(lib.mrb_cluster_tags(t.ctypes.data_as(ctypes.c_void_p), ncell, *args, ctypes.byref(nb), None, None, 0))
    lo, hi = (ctypes.c_int * (3 * nb.value))(), (ctypes.c_int * (3 * nb.value))()
    check(lib.mrb_cluster_tags(t.ctypes.data_as(ctypes.c_void_p), ncell, *args, ctypes.byref(nb), lo, hi, nb.value))
    return [(tuple(lo[3 * b:3 * b + 3]), tuple(hi[3 * b:3 * b + 3])) for b in range(nb.value)]
