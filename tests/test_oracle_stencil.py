"""Pins of the oracle's D3Q27 stencil against the reference tables
(Source/Stencil.H:49-167) and the invariants check_stencil() enforces
(Source/Stencil.cpp:5-62)."""
from fractions import Fraction

import numpy as np

# Source/Stencil.H:49-79, in order
REF_EVS = [
    (0, 0, 0),
    (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
    (1, 1, 0), (-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1),
    (1, 0, -1), (-1, 0, 1), (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1),
    (1, 1, 1), (-1, -1, -1), (1, -1, 1), (-1, 1, -1), (1, -1, -1), (-1, 1, 1),
    (1, 1, -1), (-1, -1, 1),
]
# Source/Stencil.H:125-134 (enum values of the opposite directions)
REF_BOUNCE = [0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17,
              20, 19, 22, 21, 24, 23, 26, 25]
# Source/Stencil.H:136-145 (bounce_Xdir)
REF_BOUNCE_X = [0, 2, 1, 3, 4, 5, 6, 10, 9, 8, 7, 14, 13, 12, 11, 15, 16, 17, 18,
                24, 23, 26, 25, 20, 19, 22, 21]
# Source/Stencil.H:147-156 and 158-167
REF_BOUNCE_Y = [0, 1, 2, 4, 3, 5, 6, 9, 10, 7, 8, 11, 12, 13, 14, 18, 17, 16, 15,
                21, 22, 19, 20, 25, 26, 23, 24]
REF_BOUNCE_Z = [0, 1, 2, 3, 4, 6, 5, 7, 8, 9, 10, 13, 14, 11, 12, 17, 18, 15, 16,
                25, 26, 23, 24, 21, 22, 19, 20]


def test_velocity_table_and_opposites(oracle):
    ev, bounce, w = oracle.stencil()
    assert [tuple(r) for r in ev] == REF_EVS
    assert list(bounce) == REF_BOUNCE
    for q in range(27):
        assert tuple(ev[bounce[q]]) == tuple(-ev[q])  # Stencil.cpp:20-24


def test_slip_reflection_tables_match_reference(oracle):
    assert oracle.bounce_dir(0) == REF_BOUNCE_X
    assert oracle.bounce_dir(1) == REF_BOUNCE_Y
    assert oracle.bounce_dir(2) == REF_BOUNCE_Z


def test_weights(oracle):
    ev, _, w = oracle.stencil()
    # weight classes (Stencil.cpp:44-59) and WO == 8/27 exactly in double
    s = np.abs(ev).sum(axis=1)
    assert np.all(w[s == 1] == 2.0 / 27.0)
    assert np.all(w[s == 2] == 1.0 / 54.0)
    assert np.all(w[s == 3] == 1.0 / 216.0)
    assert w[0] == 8.0 / 27.0
    # isotropy in exact arithmetic
    W = [Fraction(8, 27) if a == 0 else Fraction(2, 27) if a == 1 else Fraction(1, 54) if a == 2
         else Fraction(1, 216) for a in s]
    assert sum(W) == 1
    th = Fraction(1, 3)
    for a in range(3):
        for b in range(3):
            m2 = sum(W[q] * ev[q][a] * ev[q][b] for q in range(27))
            assert m2 == (th if a == b else 0)
    # fourth-order isotropy used by the Grad expansion's second moment
    assert sum(W[q] * ev[q][0] ** 4 for q in range(27)) == 3 * th * th
    assert sum(W[q] * ev[q][0] ** 2 * ev[q][1] ** 2 for q in range(27)) == th * th


def test_engine_host_stencil_matches_reference_table():
    from marbles_amd.stencil import EV

    assert EV.tolist() == [list(c) for c in REF_EVS]
