"""Pins of the oracle's lattice math (Source/Utilities.H) by the identities
the mathematics fixes, independent of how the code spells them."""
import math

import numpy as np
import pytest


def _fsum(xs):
    return math.fsum(xs)


@pytest.fixture(scope="module")
def st(oracle):
    return oracle.stencil()


def test_product_equilibrium_moments(oracle, st):
    # Utilities.H:42-63: f_eq = rho*phi_x*phi_y*phi_z with
    # sum_c phi = 1, sum_c c phi = u, sum_c c^2 phi = P  =>  closed-form moments
    ev, _, _ = st
    rng = np.random.default_rng(1)
    lib = oracle.load()
    for _ in range(20):
        rho = 0.8 + 0.4 * rng.random()
        u = 0.1 * (rng.random(3) - 0.5)
        P = 0.2 + 0.2 * rng.random(3)
        f = [lib.orc_set_extended_equilibrium_value(rho, oracle.d3(u), *P, q) for q in range(27)]
        tol = 4e-16 * rho * 27
        assert abs(_fsum(f) - rho) < tol
        for a in range(3):
            assert abs(_fsum(ev[q][a] * f[q] for q in range(27)) - rho * u[a]) < tol
            assert abs(_fsum(ev[q][a] ** 2 * f[q] for q in range(27)) - rho * P[a]) < tol
        assert abs(_fsum(ev[q][0] * ev[q][1] * f[q] for q in range(27)) - rho * u[0] * u[1]) < tol
        assert abs(_fsum(ev[q][0] ** 2 * ev[q][1] ** 2 * f[q] for q in range(27)) - rho * P[0] * P[1]) < tol


def test_rest_equilibrium_is_weights(oracle, st):
    # at RT = 1/3 and u = 0 the product equilibrium is w_q * rho
    _, _, w = st
    lib = oracle.load()
    for q in range(27):
        v = lib.orc_set_equilibrium_value(1.0, oracle.d3((0, 0, 0)), 1.0 / 3.0, q)
        assert v == pytest.approx(w[q], rel=4e-15, abs=0)


def test_non_extended_equals_extended_with_p_u2_plus_rt(oracle):
    lib = oracle.load()
    u = (0.03, -0.02, 0.01)
    rt = 0.2
    for q in range(27):
        a = lib.orc_set_equilibrium_value(1.1, oracle.d3(u), rt, q)
        b = lib.orc_set_extended_equilibrium_value(1.1, oracle.d3(u), u[0] ** 2 + rt, u[1] ** 2 + rt,
                                                   u[2] ** 2 + rt, q)
        assert a == b


def test_grad_expansion_moments(oracle, st):
    # Utilities.H:65-163 with frame velocity 0: sum g = E, sum c g = j,
    # sum c c g = flux (needs the D3Q27 fourth-order isotropy)
    ev, _, _ = st
    lib = oracle.load()
    rng = np.random.default_rng(2)
    for _ in range(20):
        E = 0.5 + rng.random()
        j = 0.05 * (rng.random(3) - 0.5)
        fl = np.array([0.3, 0.35, 0.4, 0.01, -0.02, 0.015]) * (0.9 + 0.2 * rng.random(6))
        g = [lib.orc_grad_expansion(E, oracle.d3(j), (oracle.c_double * 6)(*fl), q) for q in range(27)]
        tol = 1e-15 * 27 * (E + 1)
        assert abs(_fsum(g) - E) < tol
        for a in range(3):
            assert abs(_fsum(ev[q][a] * g[q] for q in range(27)) - j[a]) < tol
        pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
        for n, (a, b) in enumerate(pairs):
            assert abs(_fsum(ev[q][a] * ev[q][b] * g[q] for q in range(27)) - fl[n]) < tol


def test_equilibrium_moments_at_rest(oracle):
    # Utilities.H:245-277 at u = 0: q = 0, R_aa = 2 p h, p = rho R T,
    # h = E/(2 rho) + p/rho with T = (E/rho)/(2 cv)
    lib = oracle.load()
    rho, R, cv = 1.2, 1.0, 1.5
    T = 0.2
    E = rho * 2 * cv * T
    qeq = (oracle.c_double * 3)()
    req = (oracle.c_double * 6)()
    lib.orc_get_equilibrium_moments(rho, oracle.d3((0, 0, 0)), E, cv, R, qeq, req)
    p = rho * R * T
    h = E / (2 * rho) + p / rho
    assert list(qeq) == [0.0, 0.0, 0.0]
    for a in range(3):
        assert req[a] == pytest.approx(2 * p * h, rel=1e-15)
    assert list(req)[3:] == [0.0, 0.0, 0.0]


def test_temperature_and_energy_round_trip(oracle):
    lib = oracle.load()
    cv = 1.0 / (5.0 / 3.0 - 1.0)
    for T, rho, u in [(1 / 3, 1.0, (0.05, 0.0, 0.0)), (0.2, 0.9, (0.01, -0.03, 0.02))]:
        E = lib.orc_get_energy_vec(T, rho, oracle.d3(u), cv)
        assert lib.orc_get_temperature_scalar(E, rho, *u, cv) == pytest.approx(T, rel=1e-14)


def test_vector_temperature_expansion(oracle):
    # The RealVect overload of get_temperature (Utilities.H:198-207) expands
    # AMREX_D_TERM(a, +b, +c) to `a +b +c` after a minus sign, so it equals
    # the scalar overload plus (v^2 + w^2)/cv.  Only u_x^2 enters with the
    # correct sign; this changes T' in the g equilibrium whenever v or w != 0.
    lib = oracle.load()
    cv = 1.5
    E, rho = 1.0, 1.0
    for u in [(0.05, 0.0, 0.0), (0.0, 0.04, 0.0), (0.01, 0.02, -0.03)]:
        ts = lib.orc_get_temperature_scalar(E, rho, *u, cv)
        tv = lib.orc_get_temperature_vec(E, rho, oracle.d3(u), cv)
        assert tv - ts == pytest.approx((u[1] ** 2 + u[2] ** 2) / cv, rel=1e-9, abs=1e-17)


def test_masked_average_known_answers(oracle):
    # Utilities.H:315-350: mean of the fine values that are not the -1 sentinel
    lib = oracle.load()
    rng = np.random.default_rng(3)
    for nmask in range(9):
        v = rng.random(8) + 0.5
        idx = rng.permutation(8)[:nmask]
        v[idx] = -1.0
        r = lib.orc_masked_avgdown8((oracle.c_double * 8)(*v))
        if nmask == 8:
            assert math.isnan(r)  # coarse value left unchanged
        else:
            keep = np.setdiff1d(np.arange(8), idx)
            assert r == pytest.approx(v[keep].sum() / len(keep), rel=1e-15)
    # values within SMALL_NUM = eps*1e10 of -1 are masked too
    v = [-1.0 + 1e-7, 2.0, 2.0, 2.0, 2.0, 2.0, 2.0, 2.0]
    assert lib.orc_masked_avgdown8((oracle.c_double * 8)(*v)) == 2.0
