"""Pins of the oracle's slip walls (BCFill's SLIPWALL{X,Y,Z}NORMAL cases,
BC.H:456-462, with the reflection tables of Stencil.H:136-167).

- After a ghost fill and a stream, a wall cell receives, for every direction
  that enters the domain through the wall, its own pre-stream population of
  the mirrored direction (specular reflection at the cell itself): brute
  force on random data, each axis.
- A uniform flow tangential to slip walls is a fixed point of the full step
  (no-slip walls or a wrong reflection table break it).
"""
import numpy as np
import pytest

from marbles_amd import configs


def _slip_box(oracle, axis, n=(6, 5, 4)):
    """Periodic box except slip walls normal to `axis`."""
    per = [1, 1, 1]
    per[axis] = 0
    lo = [0, 0, 0]
    lo[axis] = 6 + axis
    L = oracle.OracleLevel(n, per)
    L.set_phys(0.01)
    L.set_bc(lo, lo)
    L.flags_all_regular()
    return L


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_slip_wall_reflects_at_the_wall_cell(oracle, axis):
    """The ghost at cell - c_q holds the mirrored population of cell itself
    (data(iv + c_q, reflect(q))), and the stream carries it back to cell:
    after[cell, q] = before[cell, reflect(q)] for the 9 directions entering
    through each wall."""
    ev = np.array(oracle.stencil()[0])
    L = _slip_box(oracle, axis)
    rng = np.random.default_rng(axis)
    fab = L.fab("f", 0)
    before = rng.uniform(0.5, 1.5, fab.shape)
    L.set_fab("f", 0, before)
    L.fillpatch(0)
    L.stream(0)
    after = L.fab("f", 0)
    ng = L.ng
    valid = (slice(None),) + tuple(slice(ng, s - ng) for s in fab.shape[1:])
    b, a = before[valid], after[valid]
    ax = 2 - axis  # (z, y, x) axis of the wall normal in a component's array
    nd = b.shape[1 + ax]
    checked = 0
    for side, c_in in ((0, 1), (nd - 1, -1)):
        for q in range(27):
            if ev[q][axis] != c_in:
                continue
            m = ev[q].copy()
            m[axis] = -m[axis]
            qm = int(np.flatnonzero((ev == m).all(axis=1))[0])
            wall = [slice(None)] * 3
            wall[ax] = side
            assert np.array_equal(a[q][tuple(wall)], b[qm][tuple(wall)]), (axis, side, q)
            checked += 1
    assert checked == 18


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_uniform_tangential_flow_between_slip_walls_is_a_fixed_point(oracle, axis):
    n = (8, 8, 8)
    per = [1, 1, 1]
    per[axis] = 0
    lo = [0, 0, 0]
    lo[axis] = 6 + axis
    T0, u = 0.2, 0.04
    mach = [0.0, 0.0, 0.0]
    mach[(axis + 1) % 3] = u / np.sqrt(5.0 / 3.0 * T0)
    L = oracle.OracleLevel(n, per)
    L.set_phys(0.02)
    L.set_bc(lo, lo)
    L.flags_all_regular()
    L.ic_constant(1.0, mach, T0)
    L.init_finish()
    f0, g0 = L.valid("f"), L.valid("g")
    L.evolve(10)
    assert np.abs(L.valid("f") - f0).max() <= 1e-14
    assert np.abs(L.valid("g") - g0).max() <= 1e-14
    # the same flow between no-slip walls is not a fixed point
    M = oracle.OracleLevel(n, per)
    M.set_phys(0.02)
    M.set_bc([1 if d == axis else 0 for d in range(3)], [1 if d == axis else 0 for d in range(3)])
    M.flags_all_regular()
    M.ic_constant(1.0, mach, T0)
    M.init_finish()
    M.evolve(10)
    assert np.abs(M.valid("f") - f0).max() > 1e-6


def test_slip_channel_deck_runs_finite(oracle):
    """The single_rotated_box boundary set (x slip walls, y inlet, y
    outflow, z periodic) with a sphere: finite, and the inflow crosses the
    domain."""
    cfg = configs.slip_channel((16, 24, 8))
    L = oracle.build_oracle_level(cfg)
    L.evolve(30)
    f = L.valid("f")
    assert np.isfinite(f).all()
