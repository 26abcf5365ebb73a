"""Pins of the oracle's Channel and Parabolic velocity operators
(VelocityBC.H:90-192 with VelocityBC.cpp:23-54) inside the boundary fill
(BC.H:344-471).

After the boundary fill, every population of an x-lo ghost cell that
points into the domain is f_eq(rho, u(iv), R T) at the ghost cell's own
velocity.  u(iv) is computed here from the operators' published formulas
(channel: 16 u_ref y (ny-1-y) z (nz-1-z) / ny^4 in x; parabolic:
4 u_m x (H - x) / H^2 along the tangential direction with x the cell
centre along the normal one); a swapped index, a missing factor or the
wrong face would break the equality.  The equilibrium function itself is
pinned elsewhere (tests/test_oracle_stencil.py)."""
import ctypes
import math

import pytest

from marbles_amd import configs

AIR = 28.96


def _expected_velocity(cfg, iv):
    v = cfg["velocity_bc"]
    sos = math.sqrt(v["adiabatic_exponent"] * (AIR / v["mean_molecular_mass"]) * v["initial_temperature"])
    vref = v["Mach_ref"] * sos
    n = cfg["ncell"]
    if v["kind"] == "channel":
        c1 = iv[1] * (n[1] - 1 - iv[1])
        c2 = iv[2] * (n[2] - 1 - iv[2])
        return (16.0 * vref * c1 * c2 / float(n[1]) ** 4, 0.0, 0.0)
    nd, td = v["normal_dir"], v["tangential_dir"]
    lo, hi = cfg["prob_lo"][nd], cfg["prob_hi"][nd]
    h = hi - lo
    x = lo + (iv[nd] + 0.5) * (h / n[nd])
    out = [0.0, 0.0, 0.0]
    out[td] = 4.0 * vref * x * (h - x) / (h * h)
    return tuple(out)


@pytest.mark.parametrize("kind", ["channel", "parabolic"])
def test_inlet_ghosts_hold_the_profile_equilibrium(oracle, kind):
    cfg = configs.channel_profile((8, 10, 12), kind=kind)
    O = oracle.build_oracle_level(cfg)
    O.physbc(0)
    fab = O.fab("f", 0)  # (27, nz+6, ny+6, nx+6), ng = 3
    ng = 3
    ev = oracle.stencil()[0]
    R = AIR / cfg["velocity_bc"]["mean_molecular_mass"]
    RT = R * cfg["velocity_bc"]["initial_temperature"]
    nx, ny, nz = cfg["ncell"]
    checked = 0
    peak = 0.0
    for k in range(nz):
        for j in range(ny):
            iv = (-1, j, k)
            u = _expected_velocity(cfg, iv)
            peak = max(peak, abs(u[0]))
            for q in range(27):
                if ev[q][0] != 1:
                    continue
                dest = (iv[0] + ev[q][0], iv[1] + ev[q][1], iv[2] + ev[q][2])
                if not (0 <= dest[1] < ny and 0 <= dest[2] < nz):
                    continue
                got = fab[q, k + ng, j + ng, iv[0] + ng]
                want = oracle.load().orc_set_equilibrium_value(1.0, (ctypes.c_double * 3)(*u), RT, q)
                assert got == want, (kind, iv, q, got, want)
                checked += 1
    assert checked > 0 and peak > 0.02

