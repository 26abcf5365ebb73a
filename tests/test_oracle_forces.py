"""Pins of the oracle's momentum-exchange force (LBM::compute_eb_forces,
LBM.cpp:994-1044).

Halfway bounce-back reverses every population that streams from a fluid
cell into a solid one, so the next stream changes the fluid momentum by
exactly minus the summed 2 c_q f_q over those links: the force the
reference reports.  For a cuboid body every solid cell hit by a link has a
fluid face neighbour (is_fluid comp1 = 1), so the two sets of links agree
and F = -dP(stream) up to the rounding of the momentum sums.
"""
import numpy as np

from marbles_amd import configs


def _periodic_body(n, solid, u=0.05):
    cfg = configs.channel_sphere((n, n, n), radius=4.0, center=(n / 2,) * 3, u=u)
    cfg["periodic"] = (1, 1, 1)
    cfg["bc_lo"] = (0, 0, 0)
    cfg["bc_hi"] = (0, 0, 0)
    cfg["velocity_bc"] = None
    cfg["eb"] = {"type": "mask", "solid": solid}
    return cfg


def test_force_equals_momentum_lost_by_the_next_stream(oracle):
    n = 24
    solid = np.zeros((n, n, n), dtype=np.uint8)
    solid[9:15, 10:14, 8:16] = 1  # [z, y, x]
    O = oracle.build_oracle_level(_periodic_body(n, solid))
    O.evolve(3)
    F = O.eb_force()
    P0 = O.conservation()[1:4]
    for lat in (0, 1):
        O.fillpatch(lat)
        O.stream(lat)
    dP = O.conservation()[1:4] - P0
    scale = np.abs(O.conservation()[0])  # momentum sums are O(mass * u)
    assert np.abs(F[0]) > 1e-3
    np.testing.assert_allclose(F, -dP, rtol=0, atol=1e-14 * scale)


def test_force_on_a_sphere_has_no_cross_flow_component(oracle):
    n = 24
    cfg = _periodic_body(n, None)
    cfg["eb"] = {"type": "sphere", "center": (12.0, 12.0, 12.0), "radius": 5.0, "has_fluid_inside": 0}
    O = oracle.build_oracle_level(cfg)
    O.evolve(2)
    F = O.eb_force()
    # flow along x past a body symmetric under y -> -y and z -> -z
    assert F[0] > 0
    assert abs(F[1]) <= 1e-13 * F[0] and abs(F[2]) <= 1e-13 * F[0], F


def test_force_at_rest_vanishes(oracle):
    n = 16
    cfg = _periodic_body(n, None, u=0.0)
    cfg["eb"] = {"type": "sphere", "center": (8.0, 8.0, 8.0), "radius": 4.0, "has_fluid_inside": 0}
    O = oracle.build_oracle_level(cfg)
    O.evolve(1)
    assert np.abs(O.eb_force()).max() <= 1e-14
