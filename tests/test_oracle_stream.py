"""Pins of the oracle's stream (Source/LBM.cpp:558-604) and gradient
(Source/Utilities.H:279-312) by brute force and known answers."""
import numpy as np
import pytest


def _slices(c, n):
    # (source slice, target slice) along one axis for a shift by c
    if c == 1:
        return slice(0, n - 1), slice(1, n)
    if c == -1:
        return slice(1, n), slice(0, n - 1)
    return slice(0, n), slice(0, n)


def np_push_stream(f, isf0, ev, bounce):
    """Literal push with halfway bounce-back into a -1 pre-filled temporary."""
    fs = -np.ones_like(f)
    nz, ny, nx = isf0.shape
    for q in range(27):
        cx, cy, cz = ev[q]
        sx, tx = _slices(cx, nx)
        sy, ty = _slices(cy, ny)
        sz, tz = _slices(cz, nz)
        src_f = isf0[sz, sy, sx] == 1
        tgt_f = isf0[tz, ty, tx] != 0
        v = f[q][sz, sy, sx]
        view = fs[q][tz, ty, tx]
        view[src_f & tgt_f] = v[src_f & tgt_f]
        back = fs[bounce[q]][sz, sy, sx]
        back[src_f & ~tgt_f] = v[src_f & ~tgt_f]
    return fs


def np_pull_valid(f, isf0, ev, bounce, g):
    """The engine's pull: for a fluid cell y, fluid(y-c) ? f_q(y-c) : f_qbar(y)."""
    nz, ny, nx = isf0.shape
    out = -np.ones((27, nz - 2 * g, ny - 2 * g, nx - 2 * g))
    yz, yy, yx = np.meshgrid(np.arange(g, nz - g), np.arange(g, ny - g), np.arange(g, nx - g), indexing="ij")
    fl = isf0[yz, yy, yx] == 1
    for q in range(27):
        cx, cy, cz = ev[q]
        sz, sy, sx = yz - cz, yy - cy, yx - cx
        src_fluid = isf0[sz, sy, sx] == 1
        v = np.where(src_fluid, f[q][sz, sy, sx], f[bounce[q]][yz, yy, yx])
        out[q] = np.where(fl, v, -1.0)
    return out


def _random_level(oracle, periodic, n=(9, 7, 6), solid_frac=0.2, seed=0):
    L = oracle.OracleLevel(n, periodic, ng=3)
    rng = np.random.default_rng(seed)
    shape = L.fab_shapes[0]
    isf = np.zeros((2,) + shape, dtype=np.int32)
    isf[0] = (rng.random(shape) > solid_frac).astype(np.int32)
    L.set_fab("is_fluid", 0, isf)
    f = rng.random((27,) + shape)
    L.set_fab("f", 0, f)
    return L, f, isf[0]


def test_push_stream_matches_bruteforce(oracle):
    ev, bounce, _ = oracle.stencil()
    for seed in range(3):
        L, f, isf0 = _random_level(oracle, (0, 0, 0), seed=seed)
        L.stream(0)
        assert np.array_equal(L.fab("f", 0), np_push_stream(f, isf0, ev, bounce))


def test_pull_equals_push_on_valid_cells(oracle):
    ev, bounce, _ = oracle.stencil()
    for seed in range(3):
        L, f, isf0 = _random_level(oracle, (0, 0, 0), seed=10 + seed)
        L.stream(0)
        got = L.fab("f", 0)[:, 3:-3, 3:-3, 3:-3]
        assert np.array_equal(got, np_pull_valid(f, isf0, ev, bounce, 3))


def test_periodic_streaming_is_a_translation(oracle):
    # all fluid, periodic 8^3: 8 fill+stream steps move every population by 8*c = 0
    L = oracle.OracleLevel((8, 8, 8), (1, 1, 1), ng=3)
    L.flags_all_regular()
    rng = np.random.default_rng(5)
    f0 = rng.random((27,) + L.fab_shapes[0])
    L.set_fab("f", 0, f0)
    L.fill_boundary("f")
    v0 = L.valid("f")
    for _ in range(8):
        L.fillpatch(0)
        L.stream(0)
    assert np.array_equal(L.valid("f"), v0)
    # one step is a pure periodic shift by c_q
    L.fillpatch(0)
    L.stream(0)
    ev, _, _ = oracle.stencil()
    v1 = L.valid("f")
    for q in range(27):
        assert np.array_equal(v1[q], np.roll(v0[q], shift=(ev[q][2], ev[q][1], ev[q][0]), axis=(0, 1, 2)))


def test_single_solid_cell_bounce_back(oracle):
    ev, bounce, _ = oracle.stencil()
    n = 7
    solid = np.zeros((n, n, n), dtype=np.uint8)
    solid[3, 3, 3] = 1
    L = oracle.OracleLevel((n, n, n), (1, 1, 1), ng=3)
    L.flags_mask(solid)
    rng = np.random.default_rng(6)
    L.set_fab("f", 0, rng.random((27,) + L.fab_shapes[0]))
    L.fill_boundary("f")
    v0 = L.valid("f")
    L.fillpatch(0)
    L.stream(0)
    v1 = L.valid("f")
    assert np.all(v1[:, 3, 3, 3] == -1.0)  # sentinel in the solid
    for q in range(1, 27):
        cx, cy, cz = ev[q]
        z, y, x = 3 - cz, 3 - cy, 3 - cx  # neighbour whose q-population points into the solid
        assert v1[bounce[q], z, y, x] == v0[q, z, y, x]
    # eb_boundary ring: the solid cell has fluid face neighbours
    isf = L.valid("is_fluid")
    assert isf[1].sum() == 1 and isf[1, 3, 3, 3] == 1


def test_gradient_known_answers(oracle):
    # Q_corr,x = rho u ((1 - 3 R T) - u^2); put a known Q field through
    # compute_q_corrections and check every branch of gradient()
    n = 6
    L = oracle.OracleLevel((n, 4, 4), (1, 1, 1), prob_hi=(3.0, 4.0, 4.0), ng=3)  # inv_dx = 2 in x
    solid = np.zeros((4, 4, n), dtype=np.uint8)
    solid[1, 1, 3] = 1
    L.flags_mask(solid)
    md = np.zeros((19,) + L.shapes1[0])
    qx = np.arange(n + 2, dtype=float) ** 2  # along x incl. the ghost layer
    md[6] = qx[None, None, :]
    L.set_fab("md", 0, md)
    L.compute_q_corrections()
    dq = L.fab("dq", 0)[0]
    # md index = x + 1 (one ghost); the domain box test makes x = 0 and
    # x = n-1 one-sided even though x is periodic (LBM.cpp:963,979-986)
    Q = lambda x: (x + 1.0) ** 2
    inv = 2.0
    assert dq[0, 0, 0] == (Q(1) - Q(0)) * inv
    assert dq[0, 0, n - 1] == (Q(n - 1) - Q(n - 2)) * inv
    assert dq[0, 0, 2] == (0.5 * Q(3) - 0.5 * Q(1)) * inv
    # next to the solid cell at x = 3 (z=1, y=1)
    assert dq[1, 1, 2] == (Q(2) - Q(1)) * inv
    assert dq[1, 1, 4] == (Q(5) - Q(4)) * inv
    assert dq[1, 1, 3] == 0.0  # solid cells are skipped (left zero)
