"""Initial-condition plugins of the reference (Source/IC.H, IC.cpp) as
per-cell field generators for mrb_init_fields.

Each plugin's thermal_initialize gives, for a fab cell iv, the density,
velocity and temperature from which ic::Initializer::initialize
(IC.H:460-525) sets f = f_eq(rho, u, R T) and the Grad g equilibrium; here
the same per-cell values are evaluated on the host for every fab cell (x
fastest, box order) and handed to the engine, which evaluates the
equilibria on the device.  Parameter names and defaults follow the
constructors in IC.cpp; as there, `mach_components` times the reference
speed of sound replaces `velocity`.
"""
from __future__ import annotations

import math

import numpy as np

AIR_MOLAR_MASS = 28.96


def _common(p: dict) -> dict:
    out = {"density": 1.0, "velocity": (0.0, 0.0, 0.0), "mach_components": (0.0, 0.0, 0.0),
           "initial_temperature": 1.0 / 3.0, "adiabatic_exponent": 5.0 / 3.0,
           "mean_molecular_mass": AIR_MOLAR_MASS}
    out.update(p)
    sos = math.sqrt(out["adiabatic_exponent"] * (AIR_MOLAR_MASS / out["mean_molecular_mass"])
                    * out["initial_temperature"])
    out["speed_of_sound_ref"] = sos
    # IC.cpp: velocity[n] = mach_components[n] * speed_of_sound_ref
    out["velocity"] = tuple(m * sos for m in out["mach_components"])
    return out


def viscosity_test(p: dict, iv, prob_lo, dx):
    """ViscosityTest::thermal_initialize (IC.H:167-249): a shear wave
    u_x = u + 0.01 c_s sin(2 pi y / wave_length) with y at the cell's low
    face (the reference's prob_lo + (iv + 0.5 * 0.0) dx)."""
    q = _common({"wave_length": 1.0, **p})
    y = prob_lo[1] + (iv[1] + 0.5 * 0.0) * dx[1]
    a0 = 0.010 * q["speed_of_sound_ref"]
    rho = np.full(y.shape, q["density"])
    u = q["velocity"][0] + a0 * np.sin(2.0 * math.pi * y / q["wave_length"])
    v = np.full(y.shape, q["velocity"][1])
    w = np.full(y.shape, q["velocity"][2])
    T = np.full(y.shape, q["initial_temperature"])
    return rho, u, v, w, T


def thermal_diffusivity_test(p: dict, iv, prob_lo, dx):
    """ThermalDiffusivityTest::thermal_initialize (IC.H:251-335): a density
    wave rho = rho0 + 0.001 T0 sin(2 pi y / wave_length) at the uniform
    pressure rho0 R T0 (T = p / (rho R))."""
    q = _common({"wave_length": 1.0, **p})
    y = prob_lo[1] + (iv[1] + 0.5 * 0.0) * dx[1]
    a0 = 0.0010 * q["initial_temperature"]
    R = AIR_MOLAR_MASS / q["mean_molecular_mass"]
    pressure = q["density"] * R * q["initial_temperature"]
    rho = q["density"] + a0 * np.sin(2.0 * math.pi * y / q["wave_length"])
    u = np.full(y.shape, q["velocity"][0])
    v = np.full(y.shape, q["velocity"][1])
    w = np.full(y.shape, q["velocity"][2])
    T = pressure / (rho * R)
    return rho, u, v, w, T


def sod_test(p: dict, iv, prob_lo, dx):
    """SodTest::thermal_initialize (IC.H:337-425): density and temperature
    steps smoothed by tanh(3 (x - x_discontinuity)), x at the cell's low
    face."""
    q = _common({"density_ratio": 1.0, "temperature_ratio": 1.0, "x_discontinuity": 10.0, **p})
    x = prob_lo[0] + (iv[0] + 0.5 * 0.0) * dx[0]
    s = 0.5 * (1.0 + np.tanh((x - q["x_discontinuity"]) * 3.0))
    rho = q["density"] + s * (q["density_ratio"] * q["density"] - q["density"])
    T0 = q["initial_temperature"]
    T = T0 + s * (q["temperature_ratio"] * T0 - T0)
    u = np.full(x.shape, q["velocity"][0])
    v = np.full(x.shape, q["velocity"][1])
    w = np.full(x.shape, q["velocity"][2])
    return rho, u, v, w, T


PLUGINS = {"viscosity_test": viscosity_test, "thermaldiffusivity_test": thermal_diffusivity_test,
           "sod": sod_test}


def fab_fields(kind: str, params: dict, ncell, prob_lo, prob_hi, boxes, ng: int):
    """(rho, u, v, w, T) over every fab cell of the boxes (grown by ng; box
    order, x fastest), as ic::Initializer::initialize visits them."""
    dx = [(prob_hi[d] - prob_lo[d]) / ncell[d] for d in range(3)]
    parts = [[] for _ in range(5)]
    for lo, hi in boxes:
        axes = [np.arange(lo[d] - ng, hi[d] + ng + 1) for d in range(3)]
        K, J, I = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
        vals = PLUGINS[kind](params, (I.astype(float), J.astype(float), K.astype(float)), prob_lo, dx)
        for k in range(5):
            parts[k].append(np.broadcast_to(vals[k], I.shape).ravel())
    return tuple(np.concatenate(p) for p in parts)
