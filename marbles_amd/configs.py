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


def channel_profile(ncell=(64, 32, 32), kind: str = "channel", tau: float = 0.55, u: float = 0.05,
                    temperature: float = 0.2, max_box: int | None = None) -> dict:
    """Walled channel with a velocity-profile inlet at x-lo and zero-gradient
    outflow at x-hi (Channel / Parabolic velocity operators, VelocityBC.H:
    90-192).  kind "channel": no-slip walls in y and z, u_x = 16 u_ref
    y (ny-1-y) z (nz-1-z) / ny^4 on the ghost cells; kind "parabolic":
    no-slip walls in y, periodic z, u_x = 4 u_m y (H - y) / H^2 at the ghost
    cell centres (normal_dir 1, tangential_dir 0).  u_ref / u_m = u; the
    domain starts at rest."""
    gamma = 5.0 / 3.0
    mach = u / math.sqrt(gamma * temperature)
    vbc = {"kind": kind, "Mach_ref": mach, "initial_density": 1.0, "initial_temperature": temperature,
           "adiabatic_exponent": gamma, "mean_molecular_mass": AIR_MOLAR_MASS}
    if kind == "parabolic":
        vbc.update({"normal_dir": 1, "tangential_dir": 0})
        periodic, bc_lo, bc_hi = (0, 0, 1), (2, 1, 0), (5, 1, 0)
    else:
        periodic, bc_lo, bc_hi = (0, 0, 0), (2, 1, 1), (5, 1, 1)
    return {
        "name": "channel_{}_{}x{}x{}".format(kind, *ncell),
        "ncell": tuple(ncell),
        "periodic": periodic,
        "prob_lo": (0.0, 0.0, 0.0),
        "prob_hi": tuple(float(n) for n in ncell),
        "bc_lo": bc_lo,
        "bc_hi": bc_hi,
        "velocity_bc": vbc,
        "nu": (tau - 0.5) * temperature,
        "tau": tau,
        "T0": temperature,
        "ic": {"type": "constant", "density": 1.0, "mach": (0.0, 0.0, 0.0),
               "temperature": temperature, "adiabatic_exponent": gamma},
        "eb": {"type": "all_regular"},
        "max_box": max_box,
    }


def slip_channel(ncell=(24, 40, 16), tau: float = 0.6, u: float = 0.04, temperature: float = 0.2,
                 sphere: bool = True, max_box: int | None = None) -> dict:
    """Slip walls (Tests/test_files/single_rotated_box deck: lbm.bc_lo =
    6 2 0, lbm.bc_hi = 6 5 0): SLIPWALLXNORMAL on both x faces, a velocity
    inlet along y at y-lo, zero-gradient outflow at y-hi, periodic z; a
    sphere stands in for the deck's rotated box.  Starts at rest."""
    gamma = 5.0 / 3.0
    mach = u / math.sqrt(gamma * temperature)
    vbc = {"kind": "constant", "dir": 1, "Mach_ref": mach, "initial_density": 1.0,
           "initial_temperature": temperature, "adiabatic_exponent": gamma, "mean_molecular_mass": AIR_MOLAR_MASS}
    eb = {"type": "all_regular"}
    if sphere:
        eb = {"type": "sphere", "center": (ncell[0] / 2, ncell[1] / 3, ncell[2] / 2), "radius": 4.0,
              "has_fluid_inside": 0}
    return {
        "name": "slip_channel_{}x{}x{}".format(*ncell),
        "ncell": tuple(ncell),
        "periodic": (0, 0, 1),
        "prob_lo": (0.0, 0.0, 0.0),
        "prob_hi": tuple(float(n) for n in ncell),
        "bc_lo": (6, 2, 0),
        "bc_hi": (6, 5, 0),
        "velocity_bc": vbc,
        "nu": (tau - 0.5) * temperature,
        "tau": tau,
        "T0": temperature,
        "ic": {"type": "constant", "density": 1.0, "mach": (0.0, 0.0, 0.0),
               "temperature": temperature, "adiabatic_exponent": gamma},
        "eb": eb,
        "max_box": max_box,
    }


def pressure_case(kind: str, ncell=(40, 20, 24), tau: float = 0.6, u: float = 0.04, temperature: float = 0.2,
                  max_box: int | None = None) -> dict:
    """Workloads with PRESSURE faces (pressure_bc_f / pressure_bc_g,
    BC.H:222-321); the boundary density and temperature come from the
    velocity operator (1 and 1/3 with none).
      "outlet":  velocity inlet u at x-lo, pressure outlet at x-hi (rho_bc 1),
                 y/z periodic, a sphere upstream of the outlet; starts at rest.
      "outlet_body": the same with the sphere reaching into the two layers
                 next to the outlet: the outlet's sums then read the solid
                 cells' contents (0, then the stream's -1), and the state
                 blows up within a few steps, as in the reference.
      "walled":  velocity inlet u along z at z-lo, pressure outlet at z-hi,
                 no-slip x walls, y periodic; starts at rest.
      "both":    pressure at both x faces (no operator: rho_bc 1, T 1/3),
                 no-slip y walls, z periodic; starts at rho 1.01, T0.
      "channel": the Channel profile inlet at x-lo, pressure outlet at
                 x-hi, no-slip y and z walls."""
    gamma = 5.0 / 3.0
    mach = u / math.sqrt(gamma * temperature)
    vbc = {"kind": "constant", "dir": 0, "Mach_ref": mach, "initial_density": 1.0,
           "initial_temperature": temperature, "adiabatic_exponent": gamma, "mean_molecular_mass": AIR_MOLAR_MASS}
    eb = {"type": "all_regular"}
    density = 1.0
    if kind in ("outlet", "outlet_body"):
        periodic, bc_lo, bc_hi = (0, 1, 1), (2, 0, 0), (3, 0, 0)
        cx = ncell[0] - (2.5 if kind == "outlet_body" else 10.0)
        eb = {"type": "sphere", "center": (cx, ncell[1] / 2, ncell[2] / 2), "radius": 4.0, "has_fluid_inside": 0}
    elif kind == "walled":
        periodic, bc_lo, bc_hi = (0, 1, 0), (1, 0, 2), (1, 0, 3)
        vbc["dir"] = 2
    elif kind == "both":
        periodic, bc_lo, bc_hi = (0, 0, 1), (3, 1, 0), (3, 1, 0)
        vbc = None
        density = 1.01
    elif kind == "channel":
        periodic, bc_lo, bc_hi = (0, 0, 0), (2, 1, 1), (3, 1, 1)
        vbc = dict(vbc, kind="channel")
    else:
        raise ValueError(kind)
    return {
        "name": "pressure_{}_{}x{}x{}".format(kind, *ncell),
        "ncell": tuple(ncell),
        "periodic": periodic,
        "prob_lo": (0.0, 0.0, 0.0),
        "prob_hi": tuple(float(n) for n in ncell),
        "bc_lo": bc_lo,
        "bc_hi": bc_hi,
        "velocity_bc": vbc,
        "nu": (tau - 0.5) * temperature,
        "tau": tau,
        "T0": temperature,
        "ic": {"type": "constant", "density": density, "mach": (0.0, 0.0, 0.0),
               "temperature": temperature, "adiabatic_exponent": gamma},
        "eb": eb,
        "max_box": max_box,
    }


def cylinder_deck(kind: str = "turek_2d2", scale: int = 1, reynolds: float | None = None) -> dict:
    """The reference's EB-cylinder decks, optionally coarsened by `scale`
    (cells, radius and centre divided by it; same Mach numbers and
    temperature, viscosity divided by it to keep the Reynolds number).

    "turek_2d2" (Tests/test_files/cylinder_turek_2d2/cylinder_turek_2d2.inp):
    the Schaefer-Turek 2D-2 channel, 660 x 123 x 4 cells (z periodic),
    parabolic inflow at x-lo (Mach 0.1 at the centreline, T = 0.03333,
    gamma = 1.4), no-slip walls in y, outflow at x-hi, cylinder r = 15 along
    z at (60, 60), nu = alpha = 0.00648 (Re = u_max D / nu = 100); starts
    from uniform flow at Mach 0.1.
    "single_cylinder" (Tests/test_files/single_cylinder/single_cylinder.inp):
    246 x 576 x 4 cells, parabolic inflow at y-lo (Mach 0.015, T = 0.03,
    gamma = 5/3), no-slip walls in x, outflow at y-hi, cylinder r = 30 along
    z at (120, 120), nu = 0.006; starts at rest.
    `reynolds` (optional) replaces the deck's viscosity (and diffusivity)
    by u_mean D / Re, the Schaefer-Turek definition of Re on the mean inflow
    velocity."""
    if kind == "turek_2d2":
        n, gamma, T, mach, r, c = (660, 123, 4), 1.4, 0.03333, 0.1, 15.0, (60.0, 60.0)
        nu, bc_lo, bc_hi, normal, tangential = 0.00648, (2, 1, 0), (5, 1, 0), 1, 0
        ic_mach = (0.1, 0.0, 0.0)
    elif kind == "single_cylinder":
        n, gamma, T, mach, r, c = (246, 576, 4), 5.0 / 3.0, 0.03, 0.015, 30.0, (120.0, 120.0)
        nu, bc_lo, bc_hi, normal, tangential = 0.006, (1, 2, 0), (1, 5, 0), 0, 1
        ic_mach = (0.0, 0.0, 0.0)
    else:
        raise ValueError("unknown cylinder deck " + kind)
    if scale != 1:
        n = (n[0] // scale, n[1] // scale, n[2])
        r, c, nu = r / scale, (c[0] / scale, c[1] / scale), nu / scale
    vbc = {"kind": "parabolic", "Mach_ref": mach, "normal_dir": normal, "tangential_dir": tangential,
           "initial_density": 1.0, "initial_temperature": T, "adiabatic_exponent": gamma,
           "mean_molecular_mass": AIR_MOLAR_MASS}
    u_max = mach * math.sqrt(gamma * T)
    name = f"cylinder_{kind}_{n[0]}x{n[1]}x{n[2]}"
    if reynolds is not None:
        nu = (2.0 * u_max / 3.0) * (2.0 * r) / reynolds
        name += f"_Re{reynolds:g}"
    return {
        "name": name,
        "ncell": n,
        "periodic": (0, 0, 1),
        "prob_lo": (0.0, 0.0, -n[2] / 2.0),
        "prob_hi": (float(n[0]), float(n[1]), n[2] / 2.0),
        "bc_lo": bc_lo,
        "bc_hi": bc_hi,
        "velocity_bc": vbc,
        "nu": nu,
        "alpha": nu,
        "adiabatic_exponent": gamma,
        "T0": T,
        "ic": {"type": "constant", "density": 1.0, "mach": ic_mach, "temperature": T, "adiabatic_exponent": gamma},
        "eb": {"type": "cylinder", "center": (c[0], c[1], 0.0), "radius": r, "height": 1000.0, "direction": 2,
               "has_fluid_inside": 0},
        # force coefficients: 2 F / (rho U^2 D L_z) with U the mean inflow
        # velocity, 2/3 of the parabola's maximum (the Schaefer-Turek
        # normalisation), St = f D / U
        "u_max": u_max,
        "u_mean": 2.0 * u_max / 3.0,
        "diameter": 2.0 * r,
        "flow_dir": tangential,
        "lift_dir": normal,
        "max_box": 32,
    }


def plugin_case(kind: str, ncell, params: dict, nu: float, alpha: float | None = None,
                periodic=(1, 1, 1), bc_lo=(0, 0, 0), bc_hi=(0, 0, 0), prob_lo=(0.0, 0.0, 0.0),
                prob_hi=None, max_box: int | None = None) -> dict:
    """Workload initialised by one of the reference's IC plugins
    (marbles_amd.ic: viscosity_test, thermaldiffusivity_test, sod), regular
    geometry; the level physics take the plugin's adiabatic exponent and
    molecular mass, as the decks set lbm.* to the ic_*.* values."""
    gamma = params.get("adiabatic_exponent", 5.0 / 3.0)
    m_bar = params.get("mean_molecular_mass", AIR_MOLAR_MASS)
    return {
        "name": "{}_{}x{}x{}".format(kind, *ncell),
        "ncell": tuple(ncell),
        "periodic": tuple(periodic),
        "prob_lo": tuple(prob_lo),
        "prob_hi": tuple(prob_hi) if prob_hi else tuple(float(n) for n in ncell),
        "bc_lo": tuple(bc_lo),
        "bc_hi": tuple(bc_hi),
        "velocity_bc": None,
        "nu": nu,
        "alpha": alpha,
        "adiabatic_exponent": gamma,
        "m_bar": m_bar,
        "ic": {"type": "plugin", "kind": kind, "params": dict(params), "adiabatic_exponent": gamma,
               "m_bar": m_bar},
        "eb": {"type": "all_regular"},
        "max_box": max_box,
    }


def shear_wave(ny: int = 64, tau: float = 0.8, temperature: float = 0.2) -> dict:
    """ViscosityTest plugin on a periodic 4 x ny x 4 box: u_x = 0.01 c_s
    sin(2 pi y / ny) decays as exp(-nu k^2 t), nu = (tau - 1/2) R T0."""
    return plugin_case("viscosity_test", (4, ny, 4),
                       {"wave_length": float(ny), "initial_temperature": temperature},
                       nu=(tau - 0.5) * temperature)


def sod_tube(nx: int = 300, tau: float = 0.55, temperature: float = 0.2) -> dict:
    """The sod deck (Tests/test_files/sod/sod.inp) scaled to nx cells:
    density 0.5 and ratio 4, temperature ratio 1/8 across x = nx/2,
    gamma = 2, zero-gradient outflow at both x faces, periodic y and z."""
    nu = (tau - 0.5) * temperature
    return plugin_case("sod", (nx, 2, 2),
                       {"density": 0.5, "x_discontinuity": nx / 2.0, "initial_temperature": temperature,
                        "adiabatic_exponent": 2.0, "density_ratio": 4.0, "temperature_ratio": 0.125},
                       nu=nu, alpha=nu, periodic=(0, 1, 1), bc_lo=(5, 0, 0), bc_hi=(5, 0, 0))


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
    lvl.set_phys(cfg["nu"], cfg.get("alpha"), adiabatic_exponent=cfg.get("adiabatic_exponent", 5.0 / 3.0),
                 m_bar=cfg.get("m_bar", AIR_MOLAR_MASS))
    lvl.set_bc(cfg["bc_lo"], cfg["bc_hi"], cfg.get("velocity_bc"))
    if storage != "f64":
        # the shifted storages subtract the rest equilibrium of this state
        lvl.set_shift_state(*reference_state(cfg))
    set_eb_flags(lvl, cfg["eb"])
    ic = cfg["ic"]
    if ic["type"] == "taylorgreen":
        lvl.init_taylorgreen(ic["rho0"], ic["v0"], ic["omega"])
    elif ic["type"] == "plugin":
        from .ic import fab_fields

        fields = fab_fields(ic["kind"], ic["params"], cfg["ncell"], cfg["prob_lo"], cfg["prob_hi"],
                            boxes or [((0, 0, 0), tuple(n - 1 for n in cfg["ncell"]))], ngv)
        lvl.init_fields(*fields, adiabatic_exponent=ic["adiabatic_exponent"], m_bar=ic["m_bar"])
    elif ic["type"] == "spectral":
        if layout != "compact":
            raise ValueError("spectral initial condition: compact layout only")
        lvl.init_fields(*compact_fields(cfg), adiabatic_exponent=ic["adiabatic_exponent"])
    else:
        lvl.init_constant(ic["density"], ic["mach"], ic["temperature"], ic["adiabatic_exponent"])
    return lvl


def reference_state(cfg: dict) -> tuple[float, float]:
    """(rho0, T0) of a workload's initial state: the rest state the shifted
    population storages subtract."""
    ic = cfg["ic"]
    if ic["type"] == "taylorgreen":
        return float(ic["rho0"]), 1.0 / 3.0
    if ic["type"] == "plugin":
        p = ic["params"]
        return float(p.get("density", 1.0)), float(p.get("initial_temperature", 1.0 / 3.0))
    return float(ic.get("density", 1.0)), float(ic.get("temperature", cfg.get("T0", 1.0 / 3.0)))


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


# ---------------------------------------------------------------------------
# Two-level workload
# ---------------------------------------------------------------------------

def random_boxes(ncell, edges=(8, 16, 32, 64), nbox: int | None = None, seed: int = 0):
    """Domain chopped into boxes whose edges are powers of two in `edges`:
    start from max(edges)-sized blocks and split a random box in half along a
    random splittable axis until there are `nbox` boxes (default: as many as
    boxes of edge 16 would give).  Deterministic for a seed."""
    import numpy as np

    emax, emin = max(edges), min(edges)
    for n in ncell:
        if n % emax:
            raise ValueError("domain extents must be multiples of the largest box edge")
    boxes = [((x, y, z), (x + emax - 1, y + emax - 1, z + emax - 1))
             for z in range(0, ncell[2], emax) for y in range(0, ncell[1], emax)
             for x in range(0, ncell[0], emax)]
    if nbox is None:
        nbox = int(np.prod(ncell)) // (16 ** 3)
    rng = np.random.default_rng(seed)
    while len(boxes) < nbox:
        splittable = [i for i, (lo, hi) in enumerate(boxes) if any(h - l + 1 > emin for l, h in zip(lo, hi))]
        if not splittable:
            break
        i = splittable[int(rng.integers(len(splittable)))]
        lo, hi = boxes[i]
        axes = [d for d in range(3) if hi[d] - lo[d] + 1 > emin]
        d = axes[int(rng.integers(len(axes)))]
        half = (hi[d] - lo[d] + 1) // 2
        hi1 = list(hi)
        hi1[d] = lo[d] + half - 1
        lo2 = list(lo)
        lo2[d] = lo[d] + half
        boxes[i] = (lo, tuple(hi1))
        boxes.append((tuple(lo2), hi))
    return boxes


def refine_boxes(coarse_boxes, fraction: float, tol: float = 0.005, max_box: int = 64, seed: int = 0):
    """Fine-level boxes (refinement ratio 2) over a seeded random subset of
    the coarse boxes whose volume is fraction +- tol of the union of all of
    them; refined boxes are chopped to at most max_box per side."""
    import numpy as np

    vols = [int(np.prod([h - l + 1 for l, h in zip(lo, hi)])) for lo, hi in coarse_boxes]
    total = sum(vols)
    rng = np.random.default_rng(seed)
    chosen, acc = [], 0
    for i in rng.permutation(len(coarse_boxes)):
        if acc >= (fraction - tol) * total:
            break
        if acc + vols[i] <= (fraction + tol) * total:
            chosen.append(int(i))
            acc += vols[i]
    fine = []
    for i in sorted(chosen):
        lo, hi = coarse_boxes[i]
        flo = [2 * a for a in lo]
        fhi = [2 * h + 1 for h in hi]
        for z in range(flo[2], fhi[2] + 1, max_box):
            for y in range(flo[1], fhi[1] + 1, max_box):
                for x in range(flo[0], fhi[0] + 1, max_box):
                    blo = (x, y, z)
                    bhi = tuple(min(a + max_box - 1, h) for a, h in zip(blo, fhi))
                    fine.append((blo, bhi))
    return fine, acc / total


def amr_two_level(ncell=(256, 256, 256), edges=(8, 16, 32, 64), nbox: int | None = None,
                  fine_fraction: float = 0.2, tau: float = 0.6, prandtl: float = 0.71,
                  temperature: float = 0.2, noise: float = 0.05, seed: int = 3, eb: dict | None = None) -> dict:
    """Two levels, ratio 2, periodic: seeded uneven coarse boxes with edges in
    `edges`, a fine level over a seeded 20% of them, thermal f+g with
    nu = (tau - 1/2) R T0 and alpha = nu / Pr, at rest with temperature
    T0 (1 + noise * n(x)), n a seeded smooth field with max |n| <= 1.
    `eb` (e.g. a sphere in physical coordinates) classifies the cells of
    both levels, each at its own resolution (LBM::initialize_is_fluid per
    level).

    T0 = 0.2: the reference thermal model is linearly unstable at T0 = 0.3
    with tau = 0.6 (a grid-scale energy mode grows), stable at T0 = 0.2 on
    both levels (coarse tau = 0.6, fine tau = 2 (tau - 1/2) + 1/2 = 0.7)."""
    coarse = random_boxes(ncell, edges, nbox, seed)
    fine, frac = refine_boxes(coarse, fine_fraction, seed=seed + 1)
    nu = (tau - 0.5) * temperature
    return {
        "name": "amr2_{}x{}x{}".format(*ncell),
        "ncell": tuple(ncell),
        "periodic": (1, 1, 1),
        "prob_lo": (0.0, 0.0, 0.0),
        "prob_hi": tuple(float(n) for n in ncell),
        "bc_lo": (0, 0, 0),
        "bc_hi": (0, 0, 0),
        "velocity_bc": None,
        "nu": nu,
        "alpha": nu / prandtl,
        "tau": tau,
        "T0": temperature,
        "ic": {"type": "thermal_noise", "density": 1.0, "temperature": temperature, "noise": noise,
               "seed": seed + 2, "adiabatic_exponent": 5.0 / 3.0},
        "eb": dict(eb) if eb else {"type": "all_regular"},
        "coarse_boxes": coarse,
        "fine_boxes": fine,
        "fine_fraction": frac,
        "ratio": 2,
    }


def set_eb_flags(lvl, eb: dict) -> None:
    """Cell classification of a level (engine or oracle) from a workload's
    `eb` entry."""
    if eb["type"] == "sphere":
        lvl.flags_sphere(eb["center"], eb["radius"], eb.get("has_fluid_inside", 0))
    elif eb["type"] == "cylinder":
        lvl.flags_cylinder(eb["center"], eb["radius"], eb["height"], eb["direction"], eb.get("has_fluid_inside", 0))
    elif eb["type"] == "all_regular":
        lvl.flags_all_regular()
    else:
        raise ValueError("unsupported eb type " + str(eb["type"]))


def smooth_noise(ic: dict, prob_lo, prob_hi, x, y, z):
    """Seeded smooth field with max |n| <= 1 at physical points (x, y, z):
    a sum of 24 random Fourier modes with integer wavevectors |k_d| <= 3 over
    the domain, normalised by the sum of the amplitudes."""
    import numpy as np

    rng = np.random.default_rng(ic["seed"])
    nm = 24
    k = rng.integers(-3, 4, size=(nm, 3))
    amp = rng.uniform(0.5, 1.0, size=nm)
    phase = rng.uniform(0.0, 2.0 * math.pi, size=nm)
    L = [h - l for l, h in zip(prob_lo, prob_hi)]
    out = np.zeros(np.broadcast(x, y, z).shape)
    for m in range(nm):
        arg = (2.0 * math.pi * (k[m, 0] * (x - prob_lo[0]) / L[0] + k[m, 1] * (y - prob_lo[1]) / L[1]
                                + k[m, 2] * (z - prob_lo[2]) / L[2]) + phase[m])
        out += amp[m] * np.cos(arg)
    return out / amp.sum()


def level_fields(cfg: dict, ncell, boxes, ng: int):
    """(rho, u, v, w, T) of the thermal-noise initial condition at the cell
    centres of every fab cell (ghosts included) of `boxes` on a level with
    `ncell` cells over the workload's physical domain; box order, x fastest."""
    import numpy as np

    ic = cfg["ic"]
    plo, phi = cfg["prob_lo"], cfg["prob_hi"]
    dx = [(h - l) / n for l, h, n in zip(plo, phi, ncell)]
    parts = []
    for lo, hi in boxes:
        ax = [plo[d] + (np.arange(lo[d] - ng, hi[d] + ng + 1) + 0.5) * dx[d] for d in range(3)]
        zz, yy, xx = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
        parts.append(smooth_noise(ic, plo, phi, xx, yy, zz).ravel())
    n = np.concatenate(parts)
    T = ic["temperature"] * (1.0 + ic["noise"] * n)
    rho = np.full_like(T, ic["density"])
    zero = np.zeros_like(T)
    return rho, zero, zero.copy(), zero.copy(), T


# ---------------------------------------------------------------------------
# Precision workload
# ---------------------------------------------------------------------------

def spectral_turbulence(n: int = 512, kmax: int = 32, amplitude: float = 0.04, tau: float = 0.52,
                        temperature: float = 0.2, seed: int = 4) -> dict:
    """Periodic box with a seeded random-phase divergence-free velocity field,
    energy spectrum E(k) ~ k^(-5/3) for 1 <= |k| <= kmax (zero elsewhere),
    max |u| = amplitude, rho = 1, uniform T0.

    T0 = 0.2 rather than the reference default 1/3: at T0 = 1/3 the
    reference thermal model amplifies a grid-scale energy mode for tau < 1
    (tau = 0.52 diverges); at T0 = 0.2 it is stable.  nu = (tau - 1/2) R T0."""
    return {
        "name": f"spectral_{n}",
        "ncell": (n, n, n),
        "periodic": (1, 1, 1),
        "prob_lo": (0.0, 0.0, 0.0),
        "prob_hi": (float(n),) * 3,
        "bc_lo": (0, 0, 0),
        "bc_hi": (0, 0, 0),
        "velocity_bc": None,
        "nu": (tau - 0.5) * temperature,
        "tau": tau,
        "T0": temperature,
        "ic": {"type": "spectral", "kmax": kmax, "amplitude": amplitude, "seed": seed, "density": 1.0,
               "temperature": temperature, "adiabatic_exponent": 5.0 / 3.0},
        "eb": {"type": "all_regular"},
        "max_box": None,
    }


def spectral_velocity(n: int, kmax: int, amplitude: float, seed: int):
    """(u, v, w) on an n^3 periodic grid (arrays indexed [z, y, x]): random
    phases, |u_hat(k)|^2 ~ E(|k|) / (4 pi |k|^2) with E ~ k^(-5/3) on
    1 <= |k| <= kmax, projected onto k . u_hat = 0, real by construction,
    scaled to max |u| = amplitude."""
    import numpy as np

    rng = np.random.default_rng(seed)
    k1 = np.fft.fftfreq(n, d=1.0 / n)
    kz, ky, kx = np.meshgrid(k1, k1, k1[: n // 2 + 1], indexing="ij")
    kk = np.sqrt(kx * kx + ky * ky + kz * kz)
    shell = (kk >= 1.0) & (kk <= kmax)
    amp = np.zeros_like(kk)
    amp[shell] = np.sqrt(kk[shell] ** (-5.0 / 3.0) / (4.0 * np.pi * kk[shell] ** 2))
    uh = []
    for _ in range(3):
        ph = rng.uniform(0.0, 2.0 * np.pi, kk.shape)
        uh.append(amp * np.exp(1j * ph))
    # solenoidal projection u_hat -= k (k . u_hat) / |k|^2
    with np.errstate(invalid="ignore", divide="ignore"):
        kdotu = (kx * uh[0] + ky * uh[1] + kz * uh[2]) / np.where(kk > 0, kk * kk, 1.0)
    uh = [uh[0] - kx * kdotu, uh[1] - ky * kdotu, uh[2] - kz * kdotu]
    u = [np.fft.irfftn(c, s=(n, n, n), axes=(0, 1, 2)) for c in uh]
    scale = amplitude / max(float(np.abs(c).max()) for c in u)
    return [c * scale for c in u]


def spectral_fab_fields(cfg: dict, boxes, ng: int):
    """(rho, u, v, w, T) over every fab cell (ghosts included, box order, x
    fastest) for the spectral initial condition: the periodic field evaluated
    at each fab cell's periodic image."""
    import numpy as np

    ic = cfg["ic"]
    n = cfg["ncell"][0]
    vel = spectral_velocity(n, ic["kmax"], ic["amplitude"], ic["seed"])
    parts = [[], [], []]
    for lo, hi in boxes:
        ax = [np.arange(lo[d] - ng, hi[d] + ng + 1) % n for d in range(3)]
        for c in range(3):
            parts[c].append(vel[c][np.ix_(ax[2], ax[1], ax[0])].ravel())
    u, v, w = (np.concatenate(p) for p in parts)
    return np.full(u.size, ic["density"]), u, v, w, np.full(u.size, ic["temperature"])


def compact_fields(cfg: dict):
    """(rho, u, v, w, T) over the compact layout's cells (x fastest) for a
    field-defined initial condition."""
    import numpy as np

    ic = cfg["ic"]
    n = cfg["ncell"][0]
    if ic["type"] != "spectral":
        raise ValueError("compact_fields: spectral initial conditions only")
    u, v, w = spectral_velocity(n, ic["kmax"], ic["amplitude"], ic["seed"])
    rho = np.full(n ** 3, ic["density"])
    T = np.full(n ** 3, ic["temperature"])
    return rho, u.ravel(), v.ravel(), w.ravel(), T
