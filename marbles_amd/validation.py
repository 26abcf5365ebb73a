"""Force coefficients of the EB-cylinder decks (validation against the
reference's published flow-past-a-cylinder numbers, Docs/sphinx/VandV.rst:
10-14: drag 3.096, lift 0.8396, Strouhal 0.2986).

The force history is the momentum-exchange force of LBM::compute_eb_forces
(LBM.cpp:994-1044, written to the forces file by output_forces_file,
LBM.cpp:1955-1970) sampled every few steps; the coefficients follow the
Schaefer-Turek normalisation of the channel-cylinder benchmark the
cylinder_turek_2d2 deck is named after: C = 2 F / (rho U^2 D L_z) with U
the mean inflow velocity (2/3 of the parabola's maximum), D the diameter,
L_z the span, and St = f D / U with f the lift's frequency."""
from __future__ import annotations

import time

import numpy as np

from . import configs


def coefficients(cfg: dict, t: np.ndarray, forces: np.ndarray, window: float = 0.3,
                 min_lift_amplitude: float = 0.02) -> dict:
    """Drag / lift coefficients and Strouhal number from a force history
    (t: step numbers; forces: (n, 3)), over the last `window` of it."""
    rho = cfg["ic"]["density"]
    u, d = cfg["u_mean"], cfg["diameter"]
    span = cfg["prob_hi"][2] - cfg["prob_lo"][2]
    scale = 2.0 / (rho * u * u * d * span)
    cd = scale * forces[:, cfg["flow_dir"]]
    cl = scale * forces[:, cfg["lift_dir"]]
    k0 = int(len(t) * (1.0 - window))
    tw, cdw, clw = t[k0:], cd[k0:], cl[k0:]
    # shedding frequency from the upward zero crossings of the lift about
    # its mean (linear interpolation between samples), counted only after
    # the lift has been below -h since the last one (h: a quarter of the
    # amplitude), so that noise on a steady lift is not taken for shedding;
    # none when the lift swings by less than min_lift_amplitude
    c = clw - clw.mean()
    amp = 0.5 * float(clw.max() - clw.min())
    st, periods = None, 0
    if amp >= min_lift_amplitude:
        h, armed, tc = 0.25 * amp, False, []
        for k in range(len(c) - 1):
            if c[k] < -h:
                armed = True
            if armed and c[k] < 0.0 <= c[k + 1]:
                tc.append(tw[k] + (tw[k + 1] - tw[k]) * (-c[k]) / (c[k + 1] - c[k]))
                armed = False
        if len(tc) >= 3:
            period = float(np.mean(np.diff(tc)))
            periods = len(tc) - 1
            st = d / (period * u)
    return {
        "cd_mean": float(cdw.mean()), "cd_max": float(cdw.max()), "cd_min": float(cdw.min()),
        "cl_max": float(clw.max()), "cl_min": float(clw.min()), "cl_amplitude": float(0.5 * (clw.max() - clw.min())),
        "strouhal": st, "periods_in_window": periods, "window_steps": [int(tw[0]), int(tw[-1])],
    }


def run_cylinder(kind: str = "turek_2d2", steps: int = 200000, sample_every: int = 25, window: float = 0.3,
                 storage: str = "f64", reynolds: float | None = None,
                 log=print) -> tuple[dict, np.ndarray, np.ndarray]:
    """Runs a cylinder deck on the GPU from its initial state for `steps`
    steps, sampling the EB force every `sample_every` steps."""
    cfg = configs.cylinder_deck(kind, reynolds=reynolds)
    lvl = configs.build_level(cfg, layout="compact", storage=storage)
    mass0 = float(lvl.conservation()[0])
    nsamp = steps // sample_every
    t = np.zeros(nsamp)
    forces = np.zeros((nsamp, 3))
    t0 = time.perf_counter()
    for k in range(nsamp):
        lvl.advance(sample_every)
        forces[k] = lvl.eb_force()
        t[k] = (k + 1) * sample_every
        if (k + 1) % max(1, nsamp // 10) == 0:
            log(f"{kind}: step {int(t[k])}, F = {forces[k]}, {time.perf_counter() - t0:.1f}s")
    secs = time.perf_counter() - t0
    mass1 = float(lvl.conservation()[0])
    cells = int(np.prod(cfg["ncell"]))
    res = coefficients(cfg, t, forces, window)
    res.update({"deck": kind, "cfg": cfg["name"], "steps": steps, "seconds": secs,
                "mlups": configs.mlups(cells, steps, secs), "kernel": lvl.hot_kernel_name(),
                "reynolds_umax": cfg["u_max"] * cfg["diameter"] / cfg["nu"],
                "reynolds_umean": cfg["u_mean"] * cfg["diameter"] / cfg["nu"],
                "nu": cfg["nu"], "mass_change": mass1 / mass0 - 1.0,
                "finite": bool(np.isfinite(forces).all())})
    lvl.close()
    return res, t, forces
