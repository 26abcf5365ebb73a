"""End-to-end pins of the oracle: conservation, closed-form decay, exact
fixed points, box-decomposition independence, and the linear stability of
the reference's thermal model."""
import math

import numpy as np
import pytest

from marbles_amd import configs


def test_conservation_periodic(oracle):
    cfg = configs.taylor_green(16, tau=1.0)
    L = oracle.build_oracle_level(cfg)
    c0 = L.conservation()
    L.evolve(20)
    c1 = L.conservation()
    assert abs(c1[0] - c0[0]) <= 1e-12 * c0[0]       # mass
    assert np.all(np.abs(c1[1:4] - c0[1:4]) <= 1e-12 * c0[0])  # momentum
    assert abs(c1[4] - c0[4]) <= 1e-12 * c0[4]       # 2 rho E


def test_taylor_green_decay_closed_form(oracle):
    # u ~ sin(kx)cos(ky)cos(kz): |u|max decays as exp(-3 nu k^2 t) at low Re
    n = 24
    cfg = configs.taylor_green(n, tau=1.0, v0=0.02)
    L = oracle.build_oracle_level(cfg)
    L.f_to_macrodata()
    u0 = np.abs(L.valid("md")[1]).max()
    t = 60
    L.evolve(t)
    u1 = np.abs(L.valid("md")[1]).max()
    k = 2 * math.pi / n
    expect = math.exp(-3 * cfg["nu"] * k * k * t)
    assert u1 / u0 == pytest.approx(expect, rel=0.02)


def test_uniform_channel_is_a_fixed_point(oracle):
    # uniform inflow equal to the initial state, no body: inlet ghosts,
    # outflow copies and the collide leave f and g unchanged
    cfg = configs.channel_sphere((16, 8, 8), radius=0.0, center=(-100.0, -100.0, -100.0))
    cfg["eb"] = {"type": "all_regular"}
    L = oracle.build_oracle_level(cfg)
    f0, g0 = L.valid("f"), L.valid("g")
    L.evolve(10)
    # equal up to the rounding of re-deriving the equilibrium from moments
    assert np.abs(L.valid("f") - f0).max() <= 1e-14
    assert np.abs(L.valid("g") - g0).max() <= 1e-14


@pytest.mark.parametrize("cfg_name", ["tg", "channel"])
def test_box_decomposition_independence(oracle, cfg_name):
    if cfg_name == "tg":
        cfg = configs.taylor_green(16, tau=1.0)
    else:
        cfg = configs.channel_sphere((24, 16, 16), radius=4.0, center=(8.0, 8.0, 8.0))
    a = oracle.build_oracle_level(cfg)
    cfg8 = dict(cfg)
    cfg8["max_box"] = 8
    b = oracle.build_oracle_level(cfg8)
    a.evolve(6)
    b.evolve(6)
    for name in ("f", "g"):
        assert np.array_equal(a.valid(name), b.valid(name))


def test_sphere_flags_bruteforce(oracle):
    cfg = configs.channel_sphere((32, 24, 24), radius=6.0, center=(12.0, 12.0, 12.0))
    L = oracle.build_oracle_level(cfg)
    isf = L.valid("is_fluid")
    z, y, x = np.meshgrid(np.arange(24), np.arange(24), np.arange(32), indexing="ij")
    covered = np.ones_like(x, dtype=bool)
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                d2 = (x + a - 12.0) ** 2 + (y + b - 12.0) ** 2 + (z + c - 12.0) ** 2
                covered &= d2 <= 36.0
    assert np.array_equal(isf[0] == 0, covered)
    # mirror symmetry about the centre plane x = 12 (node coordinates)
    sub = covered[:, :, :24]
    assert np.array_equal(sub, sub[:, :, ::-1])
    # eb_boundary: solid cells with a fluid face neighbour
    fl = ~covered
    ring = covered & (np.roll(fl, 1, 0) | np.roll(fl, -1, 0) | np.roll(fl, 1, 1) | np.roll(fl, -1, 1)
                      | np.roll(fl, 1, 2) | np.roll(fl, -1, 2))
    assert np.array_equal(isf[1] == 1, ring)


def _growth_per_step(oracle, T0, tau, steps=40):
    cfg = configs.taylor_green(12)
    cfg["nu"] = (tau - 0.5) * T0
    cfg["ic"] = {"type": "constant", "density": 1.0, "mach": (0.0, 0.0, 0.0), "temperature": T0,
                 "adiabatic_exponent": 5.0 / 3.0}
    L = oracle.build_oracle_level(cfg)
    rng = np.random.default_rng(0)
    for name in ("f", "g"):
        for b, a in enumerate(L.fabs(name)):
            L.set_fab(name, b, a * (1 + 1e-9 * rng.standard_normal(a.shape)))
    L.fill_boundary("f")
    L.fill_boundary("g")
    L.f_to_macrodata()
    t0 = L.valid("md")[18].std()
    L.evolve(steps)
    t1 = L.valid("md")[18].std()
    return (t1 / t0) ** (1.0 / steps)


def test_thermal_model_linear_stability(oracle):
    # Rest state + 1e-9 noise.  At T0 = 1/3 the energy lattice's linearised
    # equilibrium has d R_xx / d(2 rho E) = 10 R T0 / 3 > 1 and BGK
    # over-relaxation (tau < 1) amplifies a grid-scale mode; at T0 = 0.2 it
    # decays.  This fixes the stable operating point of the workloads.
    assert _growth_per_step(oracle, 1.0 / 3.0, 0.6) > 1.5
    assert _growth_per_step(oracle, 1.0 / 3.0, 1.0) < 1.0
    assert _growth_per_step(oracle, 0.2, 0.55) < 1.0
