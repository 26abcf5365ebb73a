"""Pins of the oracle's regrid fill (orc_amr_regrid_fill, LBM::RemakeLevel's
FillPatchTwoLevels): a new fine level with the old boxes is the old level's
valid data exactly; over newly refined cells the conservative interpolation
makes the mean of each coarse cell's 8 children equal to the coarse value
(to rounding), and where the coarse data is linear the children reproduce
it exactly; solid cells are 0."""
import numpy as np

from amr_cases import oracle_hierarchy, small_amr, fine_ncell


def _new_fine(oracle, cfg, boxes, solid=None):
    F2 = oracle.OracleLevel(fine_ncell(cfg), cfg["periodic"], cfg["prob_lo"], cfg["prob_hi"], boxes=boxes, ng=3)
    F2.set_phys(cfg["nu"], cfg.get("alpha"), dt=0.5)
    F2.set_bc(cfg["bc_lo"], cfg["bc_hi"], cfg.get("velocity_bc"))
    if solid is None:
        F2.flags_all_regular()
    else:
        F2.flags_mask(solid)
    return F2


def test_same_boxes_copy_the_old_level(oracle):
    cfg = small_amr(n=16, nbox=24, seed=5, fraction=0.3)
    C, F = oracle_hierarchy(cfg)
    oracle.amr_evolve(C, F, 2)
    F2 = _new_fine(oracle, cfg, F.boxes)
    oracle.amr_regrid_fill(C, F, F2)
    for name in ("f", "g"):
        assert np.array_equal(F2.valid(name), F.valid(name))


def test_new_region_is_the_conservative_interpolation(oracle):
    cfg = small_amr(n=16, nbox=24, seed=5, fraction=0.3)
    C, F = oracle_hierarchy(cfg)
    oracle.amr_evolve(C, F, 2)
    nf = fine_ncell(cfg)
    F2 = _new_fine(oracle, cfg, [((0, 0, 0), tuple(n - 1 for n in nf))])  # the whole domain
    oracle.amr_regrid_fill(C, F, F2)
    old = np.zeros(tuple(reversed(nf)), dtype=bool)
    for lo, hi in F.boxes:
        old[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = True
    cold = old.reshape(nf[2] // 2, 2, nf[1] // 2, 2, nf[0] // 2, 2).any(axis=(1, 3, 5))
    for name in ("f", "g"):
        fine, crse = F2.valid(name), C.valid(name)
        assert np.array_equal(fine[:, old], F.valid(name)[:, old])
        mean = fine.reshape(27, nf[2] // 2, 2, nf[1] // 2, 2, nf[0] // 2, 2).mean(axis=(2, 4, 6))
        np.testing.assert_allclose(mean[:, ~cold], crse[:, ~cold], rtol=1e-14, atol=0)


def test_solid_cells_are_zero(oracle):
    cfg = small_amr(n=16, nbox=24, seed=5, fraction=0.3)
    C, F = oracle_hierarchy(cfg)
    nf = fine_ncell(cfg)
    solid = np.zeros(tuple(reversed(nf)), dtype=np.uint8)
    solid[10:14, 10:14, 10:14] = 1
    F2 = _new_fine(oracle, cfg, [((8, 8, 8), (15, 15, 15))], solid=solid)
    oracle.amr_regrid_fill(C, F, F2)
    f = F2.valid("f")
    assert np.all(f[:, 10:14, 10:14, 10:14] == 0.0)
    assert np.all(f[:, 8:10, 8:16, 8:16] > 0.0)


def test_linear_coarse_data_is_reproduced(oracle):
    """Coarse populations linear in x (away from the periodic wrap): the new
    fine cells lie on the line, child centres at X + 1/2 -+ 1/4."""
    cfg = small_amr(n=16, nbox=24, seed=5, fraction=0.3)
    C, F = oracle_hierarchy(cfg)
    for b, (lo, hi) in enumerate(C.boxes):
        shp = C.fab_shapes[b]
        X = np.arange(shp[2]) + lo[0] - 3 + 0.5
        arr = np.empty((27,) + shp)
        for q in range(27):
            arr[q] = 1.0 + 0.01 * q + 0.002 * X[None, None, :]
        C.set_fab("f", b, arr)
    nf = fine_ncell(cfg)
    F2 = _new_fine(oracle, cfg, [((0, 0, 0), tuple(n - 1 for n in nf))])
    oracle.amr_regrid_fill(C, F, F2)
    fine = F2.valid("f")
    old = np.zeros(tuple(reversed(nf)), dtype=bool)
    for lo, hi in F.boxes:
        old[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = True
    x = (np.arange(nf[0]) + 0.5) / 2.0
    want = np.array([1.0 + 0.01 * q + 0.002 * x for q in range(27)])
    interior = np.zeros(nf[0], dtype=bool)
    interior[2:-2] = True
    sel = (~old) & interior[None, None, :]
    got = fine[:, sel]
    ref = np.broadcast_to(want[:, None, None, :], fine.shape)[:, sel]
    assert sel.any()
    np.testing.assert_allclose(got, ref, rtol=1e-14, atol=0)
