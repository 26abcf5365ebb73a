"""Pins of the oracle's PRESSURE boundary (pressure_bc_f / pressure_bc_g,
BC.H:222-321, called from BCFill::operator(), BC.H:421-452).

f: the ghost feeding the fluid cell ivn takes f_eq(rho_bc, u_n e_n, R T)
with u_n = ndir (1 - S / rho_bc), S = 2 x (the populations ivn receives
travelling outwards, i.e. from the second layer) + (those it receives
tangentially, from its own layer).  On product-form equilibria
f_eq(rho, u e_n, R T) each one-dimensional factor phi_c sums to 1 over
c = -1, 0, 1, with phi_0 = 1 - P and phi_{+-1} = (P +- u) / 2, P = u^2 + R T,
so
    S = rho_1 (P_1 - ndir u_1) + rho_0 (1 - P_0)
with index 0 the layer next to the face and 1 the next one.  A state that
varies along the normal makes the two layers differ, so reading the wrong
layer, dropping the factor 2, swapping the outgoing and tangential sets or
flipping ndir breaks the equality; for rho = rho_bc and a uniform state,
u_n = u (the face passes the flow through).

g: the Grad equilibrium at the normal velocity ndir (1 - 1 / rho_bc), the
caller passing rho_bc_out = 1 (a reproduced quirk): equal to the ghost of a
VELOCITY face whose operator prescribes that velocity.
"""
import ctypes
import math

import numpy as np
import pytest

AIR = 28.96
NG = 3


def _level(oracle, n, periodic, bc_lo, bc_hi, vbc, fields):
    """Single-box oracle level with f, g = equilibria of fields(X, Y, Z)
    (cell-index coordinates of every fab cell, ghosts included)."""
    O = oracle.OracleLevel(n, periodic, ng=NG)
    O.set_phys(0.02)
    O.set_bc(bc_lo, bc_hi, vbc)
    O.flags_all_regular()
    nz, ny, nx = O.fab_shapes[0]
    Z, Y, X = np.meshgrid(*[np.arange(m, dtype=float) - NG for m in (nz, ny, nx)], indexing="ij")
    rho, u, v, w, T = fields(X, Y, Z)
    O.ic_fields(*[np.broadcast_to(a, X.shape).ravel() for a in (rho, u, v, w, T)])
    O.init_finish()
    return O


def _feq(oracle, rho, vel, rt, q):
    return oracle.load().orc_set_equilibrium_value(rho, (ctypes.c_double * 3)(*vel), rt, q)


def test_f_ghosts_follow_the_closed_form(oracle):
    n = (7, 5, 4)
    rho_bc, t_bc = 1.03, 0.25
    vbc = {"kind": "constant", "Mach_ref": 0.0, "initial_density": rho_bc, "initial_temperature": t_bc}

    def rho_of(x):
        return 1.0 + 0.02 * np.sin(0.9 * x + 0.3)

    def u_of(x):
        return 0.03 * np.cos(0.7 * x) + 0.01

    def t_of(x):
        return 0.2 + 0.004 * x

    def fields(X, Y, Z):
        return rho_of(X), u_of(X), 0.0 * X, 0.0 * X, t_of(X)

    O = _level(oracle, n, (0, 1, 1), (3, 0, 0), (3, 0, 0), vbc, fields)
    O.physbc(0)
    fab = O.fab("f", 0)
    ev = oracle.stencil()[0]
    checked = 0
    for face, (ghost_x, x0, x1, ndir) in {"lo": (-1, 0, 1, 1), "hi": (n[0], n[0] - 1, n[0] - 2, -1)}.items():
        P0 = u_of(x0) ** 2 + t_of(x0)
        P1 = u_of(x1) ** 2 + t_of(x1)
        S = rho_of(x1) * (P1 - ndir * u_of(x1)) + rho_of(x0) * (1.0 - P0)
        un = ndir * (1.0 - S / rho_bc)
        assert abs(un) > 1e-3
        for k in range(n[2]):
            for j in range(n[1]):
                for q in range(27):
                    # directions entering the valid box (the others hold the -1
                    # of BC.H:463-466)
                    if ev[q][0] != ndir or not (0 <= j + ev[q][1] < n[1] and 0 <= k + ev[q][2] < n[2]):
                        continue
                    got = fab[q, k + NG, j + NG, ghost_x + NG]
                    want = _feq(oracle, rho_bc, (un, 0.0, 0.0), 1.0 * t_bc, q)
                    assert got == pytest.approx(want, rel=1e-12, abs=0.0), (face, j, k, q)
                    checked += 1
    assert checked == 2 * (3 * n[1] - 2) * (3 * n[2] - 2)


def test_uniform_flow_passes_through_at_the_boundary_density(oracle):
    """rho = rho_bc, uniform u: u_n = u at both faces."""
    n = (6, 4, 4)
    u0, T0 = 0.04, 0.2
    vbc = {"kind": "constant", "Mach_ref": 0.0, "initial_density": 1.0, "initial_temperature": T0}
    O = _level(oracle, n, (0, 1, 1), (3, 0, 0), (3, 0, 0), vbc,
               lambda X, Y, Z: (1.0 + 0 * X, u0 + 0 * X, 0 * X, 0 * X, T0 + 0 * X))
    O.physbc(0)
    fab = O.fab("f", 0)
    ev = oracle.stencil()[0]
    for ghost_x, ndir in ((-1, 1), (n[0], -1)):
        for q in range(27):
            if ev[q][0] != ndir:
                continue
            got = fab[q, NG + 1, NG + 1, ghost_x + NG]
            assert got == pytest.approx(_feq(oracle, 1.0, (u0, 0.0, 0.0), T0, q), rel=1e-13, abs=0.0)


def test_g_ghosts_use_the_unit_outflow_density(oracle):
    n = (5, 4, 4)
    rho_bc, t_bc, gamma = 1.05, 0.3, 1.4
    vbc = {"kind": "constant", "Mach_ref": 0.0, "initial_density": rho_bc, "initial_temperature": t_bc,
           "adiabatic_exponent": gamma}

    def fields(X, Y, Z):
        return 1.0 + 0 * X, 0.02 + 0 * X, 0 * X, 0 * X, 0.2 + 0 * X

    O = _level(oracle, n, (0, 1, 1), (3, 0, 0), (3, 0, 0), vbc, fields)
    O.physbc(1)
    fab = O.fab("g", 0)
    ev = oracle.stencil()[0]
    sos = math.sqrt(gamma * 1.0 * t_bc)
    for ghost_x, ndir, lohi in ((-1, 1, 0), (n[0], -1, 1)):
        ug = ndir * (1.0 - 1.0 / rho_bc)
        # the same state from a VELOCITY face of the constant operator
        vv = dict(vbc, dir=0, Mach_ref=ug / sos)
        bl, bh = [3, 0, 0], [3, 0, 0]
        (bl if lohi == 0 else bh)[0] = 2
        V = _level(oracle, n, (0, 1, 1), bl, bh, vv, fields)
        V.physbc(1)
        ref = V.fab("g", 0)
        for q in range(27):
            if ev[q][0] != ndir:
                continue
            a = fab[q, NG + 1, NG + 1, ghost_x + NG]
            b = ref[q, NG + 1, NG + 1, ghost_x + NG]
            assert a == pytest.approx(b, rel=1e-13, abs=1e-15), (ndir, q)
        # and not the f velocity's state (which would be u_n = 0.02 here)
        W = _level(oracle, n, (0, 1, 1), bl, bh, dict(vbc, dir=0, Mach_ref=0.02 / sos), fields)
        W.physbc(1)
        assert not np.allclose(fab[:, NG + 1, NG + 1, ghost_x + NG], W.fab("g", 0)[:, NG + 1, NG + 1, ghost_x + NG],
                               rtol=1e-6, atol=0)


def test_walled_face_at_rest_including_its_edges(oracle):
    """PRESSURE on z-hi between NOSLIP x walls (y periodic), fluid at rest at
    density rho: at rest f_q = f_qbar, so the wall-reflected populations the
    cells next to a wall receive leave S = rho, and u_n = -(1 - rho / rho_bc)
    on every ghost of the face, the edge ghosts beyond a wall too (the z face
    comes later in BCFill's loop than the x faces and overwrites them)."""
    n = (5, 4, 6)
    rho0, T0, rho_bc = 1.01, 0.2, 0.98
    vbc = {"kind": "constant", "Mach_ref": 0.0, "initial_density": rho_bc, "initial_temperature": T0}
    O = _level(oracle, n, (0, 1, 0), (1, 0, 2), (1, 0, 3), vbc,
               lambda X, Y, Z: (rho0 + 0 * X, 0 * X, 0 * X, 0 * X, T0 + 0 * X))
    O.fillpatch(0)
    fab = O.fab("f", 0)
    ev = oracle.stencil()[0]
    un = -(1.0 - rho0 / rho_bc)
    gz = n[2] + NG
    edges = 0
    for i in range(-1, n[0] + 1):
        for q in range(27):
            if ev[q][2] != -1 or not (0 <= i + ev[q][0] < n[0] and 0 <= 1 + ev[q][1] < n[1]):
                continue
            want = _feq(oracle, rho_bc, (0.0, 0.0, un), T0, q)
            assert fab[q, gz, NG + 1, i + NG] == pytest.approx(want, rel=1e-13, abs=0.0), (i, q)
            edges += i in (-1, n[0])
    assert edges == 6


def test_density_drop_drains_towards_the_face(oracle):
    """Fluid at rest at rho = 1 behind an outlet held at rho_bc = 0.99 (x-lo
    a wall, y, z periodic): after 40 steps the flow runs towards the outlet
    and the density next to it lies between rho_bc and 1."""
    n = (24, 4, 4)
    T0 = 0.2
    vbc = {"kind": "constant", "Mach_ref": 0.0, "initial_density": 0.99, "initial_temperature": T0}
    O = _level(oracle, n, (0, 1, 1), (1, 0, 0), (3, 0, 0), vbc,
               lambda X, Y, Z: (1.0 + 0 * X, 0 * X, 0 * X, 0 * X, T0 + 0 * X))
    O.set_phys(0.05)
    O.evolve(40)
    f = O.valid("f")
    ev = np.array(oracle.stencil()[0], dtype=float)
    rho = f.sum(axis=0)
    ux = np.tensordot(ev[:, 0], f, axes=(0, 0)) / rho
    assert ux[:, :, -1].min() > 1e-4
    assert 0.99 < rho[:, :, -1].mean() < 1.0
    assert rho[:, :, 0].mean() > rho[:, :, -1].mean()
