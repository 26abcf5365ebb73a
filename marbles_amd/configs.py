"""Standard workloads, expressed with the reference's input-deck keys.

Each function returns a dict that both the engine (`build_level`) and the
oracle harness understand.  Sizes are parameters so that the same workload
can be run at its full size (benchmarks) and scaled down (parity tests).
"""
from __future__ import annotations

import math

AIR_MOLAR_MASS = 28.96
# Mach number whose constant-velocity inflow is u = 0.05 at T = 1/3,
# gamma = 5/3 (u = Mach * sqrt(gamma R T)), as written in an input deck.
MACH_U005 = 0.0670820393249937


def taylor_green(n: int = 64, tau: float = 0.6, v0: float = 0.05, max_box: int | None = None) -> dict:
    """Periodic Taylor-Green vortex (Tests/test_files/tg deck family):
    u0 = v0, rho0 = 1, wavenumber 2*pi/n per axis (ic_taylorgreen.omega = 2/n because
    IC.H fixes L = 1/pi), start at T = 1/3 (fixed by IC.H:80), nu = (tau - 1/2)/3.
    In the reference's thermal model this state is linearly unstable for
    tau < 1 (a grid-scale energy mode grows ~1.7x per step at tau = 0.6 and
    the run diverges near step 55 at n = 64); tau >= 1 is stable."""
    return {
        "name": f"taylor_green_{n}",
        "ncell": (n, n, n),
        "periodic": (1, 1, 1),
        "prob_lo": (0.0, 0.0, 0.0),
        "prob_hi": (float(n),) * 3,
        "bc_lo": (0, 0, 0),
        "bc_hi": (0, 0, 0),
        "velocity_bc": None,
        "nu": (tau - 0.5) / 3.0,
        "tau": tau,
        "T0": 1.0 / 3.0,
        "ic": {"type": "taylorgreen", "rho0": 1.0, "v0": v0, "omega": (2.0 / n,) * 3},
        "eb": {"type": "all_regular"},
        "max_box": max_box,
    }


def channel_sphere(ncell=(256, 256, 256), radius: float = 24.0, center=(64.0, 128.0, 128.0),
                   tau: float = 0.55, u: float = 0.05, temperature: float = 0.2,
                   max_box: int | None = None) -> dict:
    """Channel: velocity inlet at x-lo (u = 0.05), zero-gradient outflow at x-hi,
    periodic y/z, EB sphere (fluid outside) with halfway bounce-back
    (channel_cylinder deck family with eb2.geom_type = sphere).

    The thermal model's relaxation time is tau = nu / (R T dt) + 1/2
    (LBM.cpp:675-677), so nu = (tau - 1/2) R T0.  T0 = 0.2 (the reference
    decks' lbm.initial_temperature = 0.20) keeps the energy lattice linearly
    stable at tau = 0.55; at the default T0 = 1/3 the reference model
    amplifies a grid-scale energy mode for tau < 1 and diverges.  The inlet
    Mach number is u / sqrt(gamma R T0) (VelocityBC.cpp:12-15)."""
    gamma = 5.0 / 3.0
    mach = u / math.sqrt(gamma * temperature)
    return {
        "name": "channel_sphere_{}x{}x{}".format(*ncell),
        "ncell": tuple(ncell),
        "periodic": (0, 1, 1),
        "prob_lo": (0.0, 0.0, 0.0),
        "prob_hi": tuple(float(n) for n in ncell),
        "bc_lo": (2, 0, 0),
        "bc_hi": (5, 0, 0),
        "velocity_bc": {"kind": "constant", "dir": 0, "Mach_ref": mach, "initial_density": 1.0,
                        "initial_temperature": temperature, "adiabatic_exponent": gamma,
                        "mean_molecular_mass": AIR_MOLAR_MASS},
        "nu": (tau - 0.5) * temperature,
        "tau": tau,
        "T0": temperature,
        "ic": {"type": "constant", "density": 1.0, "mach": (mach, 0.0, 0.0),
               "temperature": temperature, "adiabatic_exponent": gamma},
        "eb": {"type": "sphere", "center": tuple(center), "radius": radius, "has_fluid_inside": 0},
        "max_box": max_box,
    }


def chop_boxes(ncell, max_box: int | None):
    """Domain chopped into boxes of at most max_box cells per side (x fastest)."""
    if not max_box:
        return [((0, 0, 0), tuple(n - 1 for n in ncell))]
    out = []
    for z in range(0, ncell[2], max_box):
        for y in range(0, ncell[1], max_box):
            for x in range(0, ncell[0], max_box):
                lo = (x, y, z)
                hi = tuple(min(a + max_box, n) - 1 for a, n in zip(lo, ncell))
                out.append((lo, hi))
    return out


def build_level(cfg: dict, layout: str = "compact", storage: str = "f64", ng: int | None = None):
    """Creates and initialises an engine Level for a workload dict."""
    from .level import Level

    if layout == "compact":
        boxes, ngv = None, 0
    else:
        boxes, ngv = chop_boxes(cfg["ncell"], cfg.get("max_box")), (3 if ng is None else ng)
    lvl = Level(cfg["ncell"], cfg["periodic"], cfg["prob_lo"], cfg["prob_hi"], boxes=boxes, ng=ngv,
                layout=layout, storage=storage)
    lvl.set_phys(cfg["nu"], cfg.get("alpha"))
    lvl.set_bc(cfg["bc_lo"], cfg["bc_hi"], cfg.get("velocity_bc"))
    eb = cfg["eb"]
    if eb["type"] == "sphere":
        lvl.flags_sphere(eb["center"], eb["radius"], eb.get("has_fluid_inside", 0))
    else:
        lvl.flags_all_regular()
    ic = cfg["ic"]
    if ic["type"] == "taylorgreen":
        lvl.init_taylorgreen(ic["rho0"], ic["v0"], ic["omega"])
    else:
        lvl.init_constant(ic["density"], ic["mach"], ic["temperature"], ic["adiabatic_exponent"])
    return lvl


def mlups(cells: int, steps: int, seconds: float) -> float:
    return cells * steps / seconds / 1e6


def bytes_per_update(storage: str = "f64", lattices: int = 2) -> int:
    """Algorithmic HBM bytes per lattice update: read + write of every stored
    population once (flags not counted)."""
    es = {"f64": 8, "f32s": 4, "f16s": 2}[storage]
    return 2 * 27 * lattices * es


def taylor_green_decay(t: float, nu: float, k: float) -> float:
    """Velocity amplitude factor of a 2D Taylor-Green mode with wavenumber k per axis."""
    return math.exp(-2.0 * nu * k * k * t)
