"""The force-coefficient post-processing of the cylinder validation
(marbles_amd/validation.py) on synthetic force histories with known
answers: the Schaefer-Turek normalisation, the Strouhal number of a
sinusoidal lift, and no Strouhal number for a steady lift with noise."""
import math

import numpy as np

from marbles_amd import configs, validation


def _history(cfg, cd, cl_amp, period, steps=60000, every=25, noise=0.0, seed=0):
    rho, u, d = 1.0, cfg["u_mean"], cfg["diameter"]
    span = cfg["prob_hi"][2] - cfg["prob_lo"][2]
    q = 0.5 * rho * u * u * d * span
    t = np.arange(every, steps + 1, every, dtype=float)
    f = np.zeros((len(t), 3))
    f[:, cfg["flow_dir"]] = q * (cd + 0.02 * np.sin(4 * math.pi * t / period))
    f[:, cfg["lift_dir"]] = q * cl_amp * np.sin(2 * math.pi * t / period + 0.3)
    if noise:
        f[:, cfg["lift_dir"]] += q * noise * np.random.default_rng(seed).standard_normal(len(t))
    return t, f


def test_shedding_coefficients():
    cfg = configs.cylinder_deck("turek_2d2")
    period = 7321.0
    t, f = _history(cfg, 3.1, 0.8, period)
    r = validation.coefficients(cfg, t, f, window=0.5)
    assert abs(r["cd_mean"] - 3.1) < 2e-3 and abs(r["cd_max"] - 3.12) < 1e-3
    assert abs(r["cl_max"] - 0.8) < 1e-3 and abs(r["cl_amplitude"] - 0.8) < 1e-3
    st = cfg["diameter"] / (period * cfg["u_mean"])
    assert r["periods_in_window"] >= 3 and abs(r["strouhal"] / st - 1.0) < 1e-3, (r["strouhal"], st)


def test_steady_lift_has_no_strouhal_number():
    cfg = configs.cylinder_deck("single_cylinder")
    t, f = _history(cfg, 4.7, 0.0, 5000.0, noise=0.003)
    r = validation.coefficients(cfg, t, f)
    assert r["strouhal"] is None and r["periods_in_window"] == 0
    assert abs(r["cd_mean"] - 4.7) < 2e-3


def test_deck_reynolds_numbers():
    # Re on u_max: 100 for the Turek deck (its inputs' own comment), and the
    # single-cylinder deck's Mach 0.015 at T = 0.03 gives u_max = 0.00335
    assert abs(configs.cylinder_deck("turek_2d2")["u_max"] * 30.0 / 0.00648 - 100.0) < 0.01
    c = configs.cylinder_deck("single_cylinder")
    assert abs(c["u_max"] - 0.015 * math.sqrt(5.0 / 3.0 * 0.03)) < 1e-15
