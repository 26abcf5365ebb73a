"""GPU parity of the two-level transfers and schedule against the oracle,
through the C ABI (bit-exact expected: FP64, same operation order)."""
import numpy as np
import pytest

from amr_cases import oracle_hierarchy, small_amr, with_sphere_across_fine_edge

pytestmark = pytest.mark.gpu


def _nonperiodic(cfg):
    """Outflow at both x faces.  The fine level is a full-span slab on the
    x-lo face plus the random fine boxes that keep 2 coarse cells from both
    x faces: in the reference schedule a fine box whose coarse/fine edge
    meets a non-periodic face feeds -1 sentinels into valid cells at the
    second substep (physbc copies a streamed, unrefreshed ghost), so such
    geometry diverges in the reference itself."""
    cfg = dict(cfg)
    cfg["periodic"] = (0, 1, 1)
    cfg["bc_lo"] = (5, 0, 0)
    cfg["bc_hi"] = (5, 0, 0)
    nf = [2 * n for n in cfg["ncell"]]
    keep = [(lo, hi) for lo, hi in cfg["fine_boxes"] if lo[0] >= 12 and hi[0] <= nf[0] - 13]
    cfg["fine_boxes"] = [((0, 0, 0), (7, nf[1] - 1, nf[2] - 1))] + keep
    return cfg


def _inlet_pressure(cfg):
    """The outflow_x geometry with a constant velocity inlet (u = 0.04) at
    x-lo, which the fine slab touches, and a pressure outlet at x-hi on the
    coarse level."""
    cfg = _nonperiodic(cfg)
    cfg["bc_lo"] = (2, 0, 0)
    cfg["bc_hi"] = (3, 0, 0)
    cfg["velocity_bc"] = {"kind": "constant", "dir": 0, "Mach_ref": 0.04 / (5.0 / 3.0 * 0.2) ** 0.5,
                          "initial_density": 1.0, "initial_temperature": 0.2}
    return cfg


CASES = {
    "periodic": lambda: small_amr(),
    "outflow_x": lambda: _nonperiodic(small_amr(n=32, nbox=80, seed=9, fraction=0.25)),
    "uneven_32": lambda: small_amr(n=32, nbox=80, seed=4, fraction=0.2),
    "inlet_pressure": lambda: _inlet_pressure(small_amr(n=32, nbox=80, seed=9, fraction=0.25)),
    "eb_sphere": lambda: with_sphere_across_fine_edge(small_amr(n=32, nbox=80, seed=4, fraction=0.2)),
}


def _copy_state(O, E):
    for name in ("f", "g"):
        E.upload(name, O.flat(name))


def _same(E, O, names=("f", "g")):
    for name in names:
        a, b = E.download(name), O.flat(name)
        if not np.array_equal(a, b):
            bad = np.flatnonzero(a != b)
            return f"{name}: {bad.size} values differ, first at {bad[0]}: {a[bad[0]]!r} vs {b[bad[0]]!r}"
    return None


@pytest.mark.parametrize("case", list(CASES))
def test_hierarchy_initial_state_matches_oracle(oracle, engine, case):
    cfg = CASES[case]()
    C, F = oracle_hierarchy(cfg)
    H = engine.Hierarchy(cfg)
    assert _same(H.crse, C) is None
    assert _same(H.fine, F) is None
    H.close()


@pytest.mark.parametrize("case", list(CASES))
def test_explode_and_coalesce_bitexact(oracle, engine, case):
    cfg = CASES[case]()
    C, F = oracle_hierarchy(cfg)
    oracle.amr_evolve(C, F, 2)   # a state with sentinels in the fine ghosts
    H = engine.Hierarchy(cfg)
    _copy_state(C, H.crse)
    _copy_state(F, H.fine)
    for lat in (0, 1):
        oracle.amr_explode(C, F, lat)
        H.explode(lat)
    err = _same(H.fine, F)
    assert err is None, "explode: " + err
    oracle.amr_coalesce(F, C, 0)
    oracle.amr_coalesce(F, C, 1)
    H.coalesce()
    err = _same(H.crse, C)
    assert err is None, "coalesce: " + err
    H.close()


@pytest.mark.parametrize("case", list(CASES))
def test_amr_advance_matches_oracle(oracle, engine, case):
    cfg = CASES[case]()
    C, F = oracle_hierarchy(cfg)
    H = engine.Hierarchy(cfg)
    n = 6
    oracle.amr_evolve(C, F, n)
    H.advance(n)
    exact = True
    for lvl, O in ((H.crse, C), (H.fine, F)):
        for name in ("f", "g"):
            a, b = lvl.valid(name), O.valid(name)
            exact &= bool(np.array_equal(a, b))
            err = float(np.abs(a - b).max() / np.abs(b).max())
            assert err <= 1e-12, (case, name, err)
    print(f"{case}: {n} coarse steps, bit-exact={exact}")
    H.close()


@pytest.mark.parametrize("case", ["outflow_x", "uneven_32", "inlet_pressure"])
def test_batched_fab_step_on_uneven_boxes_matches_oracle(oracle, engine, case):
    """Single level over the coarse level's uneven boxes (edges 4 and 8):
    the batched fab step (one launch per pass over every box) against the
    oracle's literal stages, including ring-1 ghosts next to outflow faces."""
    from oracle.pyoracle import OracleLevel
    from marbles_amd import configs

    cfg = CASES[case]()
    S = OracleLevel(cfg["ncell"], cfg["periodic"], cfg["prob_lo"], cfg["prob_hi"], boxes=cfg["coarse_boxes"], ng=3)
    S.set_phys(cfg["nu"], cfg["alpha"])
    S.set_bc(cfg["bc_lo"], cfg["bc_hi"], cfg.get("velocity_bc"))
    S.flags_all_regular()
    S.ic_fields(*configs.level_fields(cfg, S.ncell, S.boxes, 3))
    S.init_finish()
    E = engine.Level(cfg["ncell"], cfg["periodic"], cfg["prob_lo"], cfg["prob_hi"], boxes=cfg["coarse_boxes"],
                     ng=3, layout="fab")
    E.set_phys(cfg["nu"], cfg["alpha"])
    E.set_bc(cfg["bc_lo"], cfg["bc_hi"], cfg.get("velocity_bc"))
    E.flags_all_regular()
    E.init_fields(*configs.level_fields(cfg, E.ncell, E.boxes, 3))
    S.evolve(5)
    E.advance(5)
    for name in ("f", "g"):
        a, b = E.valid(name), S.valid(name)
        assert np.array_equal(a, b), (case, name, float(np.abs(a - b).max()))


def test_graph_replay_matches_eager_steps(oracle, engine):
    """hipGraph replay of the coarse step (two graphs, one per coarse buffer
    parity) against eager launches, replay after replay, with eager work and
    synchronisations between the replays."""
    cfg = CASES["uneven_32"]()
    He = engine.Hierarchy(cfg)
    Hg = engine.Hierarchy(cfg)
    for rep in range(5):
        He.advance(1)
        Hg.advance(1, graph=True)
        for lvl_e, lvl_g in ((He.crse, Hg.crse), (He.fine, Hg.fine)):
            for name in ("f", "g"):
                assert np.array_equal(lvl_e.valid(name), lvl_g.valid(name)), (rep, name)
        # eager work on the graph's levels between replays
        assert np.isfinite(Hg.crse.conservation()).all()
    He.close()
    Hg.close()


def test_fab_autotune_keeps_the_state_and_results(oracle, engine):
    cfg = CASES["uneven_32"]()
    C, F = oracle_hierarchy(cfg)
    H = engine.Hierarchy(cfg)
    before = H.fine.download("f")
    times = H.fine.autotune(reps=2)
    times.update({"crse:" + k: v for k, v in H.crse.autotune(reps=2).items()})
    assert all(v > 0 for v in times.values()), times
    assert np.array_equal(H.fine.download("f"), before)
    oracle.amr_evolve(C, F, 3)
    H.advance(3)
    for lvl, O in ((H.crse, C), (H.fine, F)):
        for name in ("f", "g"):
            assert np.array_equal(lvl.valid(name), O.valid(name)), name
    H.close()
