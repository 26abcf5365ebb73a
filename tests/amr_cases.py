"""Two-level hierarchies shared by the AMR tests (oracle and engine)."""
from __future__ import annotations

import numpy as np

from marbles_amd import configs


def small_amr(n: int = 16, seed: int = 5, fraction: float = 0.3, nbox: int = 24, **kw) -> dict:
    """Small two-level workload: n^3 coarse cells in boxes with edges 4 and 8."""
    return configs.amr_two_level(ncell=(n, n, n), edges=(4, 8), nbox=nbox, fine_fraction=fraction,
                                 seed=seed, **kw)


def with_sphere_across_fine_edge(cfg: dict, radius: float = 3.5) -> dict:
    """The workload with an EB sphere centred on the high-x face of the first
    fine box's coarse footprint: solid cells under the footprint, in the
    coalescence ring and in the explosion shell, on both levels."""
    cfg = dict(cfg)
    lo, hi = cfg["fine_boxes"][0]
    centre = ((hi[0] + 1) / 2.0, (lo[1] + hi[1] + 1) / 4.0, (lo[2] + hi[2] + 1) / 4.0)
    cfg["eb"] = {"type": "sphere", "center": centre, "radius": radius, "has_fluid_inside": 0}
    return cfg


def fine_ncell(cfg):
    return tuple(2 * n for n in cfg["ncell"])


def oracle_hierarchy(cfg: dict, ng: int = 3, fine_boxes=None):
    """(coarse, fine) oracle levels initialised like MakeNewLevelFromScratch."""
    from oracle.pyoracle import OracleLevel

    fb = cfg["fine_boxes"] if fine_boxes is None else fine_boxes
    C = OracleLevel(cfg["ncell"], cfg["periodic"], cfg["prob_lo"], cfg["prob_hi"],
                    boxes=cfg["coarse_boxes"], ng=ng)
    F = OracleLevel(fine_ncell(cfg), cfg["periodic"], cfg["prob_lo"], cfg["prob_hi"], boxes=fb, ng=ng)
    gamma = cfg["ic"]["adiabatic_exponent"]
    for L, dt in ((C, 1.0), (F, 0.5)):
        L.set_phys(cfg["nu"], cfg.get("alpha"), dt=dt)
        L.set_bc(cfg["bc_lo"], cfg["bc_hi"], cfg.get("velocity_bc"))
        configs.set_eb_flags(L, cfg.get("eb", {"type": "all_regular"}))
        L.ic_fields(*configs.level_fields(cfg, L.ncell, L.boxes, ng), adiabatic_exponent=gamma)
        L.init_finish()
    return C, F


def masked_mean(children: np.ndarray) -> np.ndarray:
    """Mean over axis 0 of the values that are not the -1 sentinel (NaN when
    all are)."""
    keep = np.abs(children + 1.0) > 2.220446049250313e-16 * 1e10
    s = np.where(keep, children, 0.0).sum(axis=0)
    n = keep.sum(axis=0)
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(n > 0, s / np.maximum(n, 1), np.nan)
