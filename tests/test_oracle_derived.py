"""Pins of the oracle's vorticity (LBM::compute_derived, LBM.cpp:909-955).

The velocity of populations set to f_eq(rho, u) is u (the product-form
equilibrium's first moment is rho u), and the reference's gradient() is a
central difference inside the domain and one-sided next to a domain face
or a solid cell, so for a velocity field linear in x, y, z the vorticity
is exact everywhere: solid-body rotation about axis a gives 2 Omega e_a,
a shear u_x = s y gives omega_z = -s.  Rotations about each axis catch a
swapped component or sign; a solid block inside the shear flow checks the
one-sided rule next to solids and the zero left in solid cells.
"""
import numpy as np
import pytest


def _level(oracle, n, velocity, solid=None):
    """Oracle level over an n = (nx, ny, nz) box with f = f_eq(1, u(x)) where
    u = velocity(X, Y, Z) at cell centres (ghosts included)."""
    from oracle.pyoracle import OracleLevel

    ng = 3
    O = OracleLevel(n, (0, 0, 0), ng=ng)
    O.set_phys(0.01)
    if solid is None:
        O.flags_all_regular()
    else:
        O.flags_mask(solid)
    nz, ny, nx = O.fab_shapes[0]
    Z, Y, X = np.meshgrid(*[np.arange(m, dtype=float) - ng + 0.5 for m in (nz, ny, nx)], indexing="ij")
    u, v, w = velocity(X, Y, Z)
    rho = np.ones_like(X)
    T = np.full_like(X, 0.2)
    O.ic_fields(rho.ravel(), u.ravel(), v.ravel(), w.ravel(), T.ravel())
    O.init_finish()
    return O


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_solid_body_rotation_gives_twice_the_angular_velocity(oracle, axis):
    n = (12, 10, 8)
    om = 2e-3
    c = [m / 2 for m in n]

    def rot(X, Y, Z):
        r = [X - c[0], Y - c[1], Z - c[2]]
        Om = [0.0, 0.0, 0.0]
        Om[axis] = om
        # u = Omega x r
        return (Om[1] * r[2] - Om[2] * r[1], Om[2] * r[0] - Om[0] * r[2], Om[0] * r[1] - Om[1] * r[0])

    O = _level(oracle, n, rot)
    w = O.vorticity()
    for a in range(3):
        expect = 2 * om if a == axis else 0.0
        np.testing.assert_allclose(w[a], expect, rtol=0, atol=1e-13)
    np.testing.assert_allclose(w[3], 2 * om, rtol=1e-12)


def test_shear_with_a_solid_block_uses_one_sided_differences(oracle):
    n = (12, 12, 8)
    s = 3e-3
    solid = np.zeros((n[2], n[1], n[0]), dtype=np.uint8)
    solid[3:5, 4:7, 5:8] = 1  # [z, y, x]

    def shear(X, Y, Z):
        return s * Y, np.zeros_like(X), np.zeros_like(X)

    O = _level(oracle, n, shear, solid)
    w = O.vorticity()
    fluid = solid == 0
    np.testing.assert_allclose(w[2][fluid], -s, rtol=1e-11)
    np.testing.assert_allclose(w[0][fluid], 0.0, atol=1e-14)
    np.testing.assert_allclose(w[1][fluid], 0.0, atol=1e-14)
    assert np.all(w[:, solid == 1] == 0.0)


def test_nonlinear_field_matches_the_difference_formula(oracle):
    """u_y = sin(2 pi x / nx): interior omega_z is the central difference
    (sin(k(x+1)) - sin(k(x-1))) / 2, the faces the one-sided ones."""
    n = (16, 6, 6)
    k = 2 * np.pi / n[0]

    def wave(X, Y, Z):
        return np.zeros_like(X), 1e-2 * np.sin(k * X), np.zeros_like(X)

    O = _level(oracle, n, wave)
    w = O.vorticity()
    x = np.arange(n[0]) + 0.5
    f = 1e-2 * np.sin(k * x)
    d = np.empty_like(f)
    d[1:-1] = 0.5 * f[2:] - 0.5 * f[:-2]
    d[0] = f[1] - f[0]
    d[-1] = f[-1] - f[-2]
    np.testing.assert_allclose(w[2], np.broadcast_to(d, w[2].shape), rtol=0, atol=1e-15)
