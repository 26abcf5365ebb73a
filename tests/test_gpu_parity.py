"""GPU parity of the HIP engine against the CPU oracle (restatement of the
reference step), through the C ABI.

Bar: bit-exact for streaming, sentinel and flag work; populations within
1e-12 relative after N steps (the FP64 path is built without FMA
contraction and reproduces the reference operation order, so it is
expected to be bit-exact as well; the tests assert the 1e-12 bar and
report exactness).
"""
import numpy as np
import pytest

from marbles_amd import configs

pytestmark = pytest.mark.gpu

REL_TOL = 1e-12


def _rel_err(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale)


def _small_channel(max_box=None):
    return configs.channel_sphere((32, 16, 16), radius=4.0, center=(10.0, 8.0, 8.0), max_box=max_box)


def _fab_pair(oracle, engine, cfg):
    from oracle.pyoracle import build_oracle_level

    O = build_oracle_level(cfg)
    boxes = O.boxes
    E = engine.Level(cfg["ncell"], cfg["periodic"], cfg["prob_lo"], cfg["prob_hi"], boxes=boxes, ng=3,
                     layout="fab")
    E.set_phys(cfg["nu"], cfg.get("alpha"))
    E.set_bc(cfg["bc_lo"], cfg["bc_hi"], cfg.get("velocity_bc"))
    E.flags_upload(O.flat("is_fluid"))
    E.upload("f", O.flat("f"))
    E.upload("g", O.flat("g"))
    return O, E


def _compact_from(oracle_level, engine, cfg):
    E = engine.Level(cfg["ncell"], cfg["periodic"], cfg["prob_lo"], cfg["prob_hi"], layout="compact")
    E.set_phys(cfg["nu"], cfg.get("alpha"))
    E.set_bc(cfg["bc_lo"], cfg["bc_hi"], cfg.get("velocity_bc"))
    E.flags_upload(oracle_level.valid("is_fluid"))
    E.upload("f", oracle_level.valid("f"))
    E.upload("g", oracle_level.valid("g"))
    return E


@pytest.mark.parametrize("max_box", [None, 8])
def test_fab_stream_bitexact(oracle, engine, max_box):
    cfg = _small_channel(max_box)
    O, E = _fab_pair(oracle, engine, cfg)
    for lat in (0, 1):
        O.fillpatch(lat)
        E.fillpatch(lat)
    for name in ("f", "g"):
        assert np.array_equal(E.download(name), O.flat(name)), f"ghost fill of {name} differs"
    for lat in (0, 1):
        O.stream(lat)
        E.stream(lat)
    for name in ("f", "g"):
        assert np.array_equal(E.download(name), O.flat(name)), f"stream of {name} differs"


@pytest.mark.parametrize("max_box", [None, 8])
def test_fab_stages_one_step_bitexact(oracle, engine, max_box):
    """The stage-by-stage fab path (fill ghosts, stream f, g, collide) leaves
    every fab cell, ghosts included, as the oracle does."""
    cfg = _small_channel(max_box)
    O, E = _fab_pair(oracle, engine, cfg)
    O.evolve(1)
    for lat in (0, 1):
        E.fillpatch(lat)
    for lat in (0, 1):
        E.stream(lat)
    E.collide()
    for name in ("f", "g"):
        assert np.array_equal(E.download(name), O.flat(name)), f"one step of {name} differs"


@pytest.mark.parametrize("max_box", [None, 8])
def test_fab_batched_steps(oracle, engine, max_box):
    """The batched fab step (mrb_stream_collide on a fab level) against the
    oracle's stages: valid cells bit for bit.  Ghost cells beyond
    non-periodic faces are not streamed by the batched step; the boundary
    functor rewrites them before they are read."""
    cfg = _small_channel(max_box)
    O, E = _fab_pair(oracle, engine, cfg)
    O.evolve(1)
    E.advance(1)
    for name in ("f", "g"):
        assert np.array_equal(E.valid(name), O.valid(name)), f"one step of {name} differs"
    O.evolve(9)
    E.advance(9)
    for name in ("f", "g"):
        err = _rel_err(E.valid(name), O.valid(name))
        assert err <= REL_TOL, (name, err)


def test_compact_stream_bitexact(oracle, engine):
    from oracle.pyoracle import build_oracle_level

    cfg = _small_channel()
    O = build_oracle_level(cfg)
    O.evolve(3)
    E = _compact_from(O, engine, cfg)
    for lat in (0, 1):
        O.fillpatch(lat)
        O.stream(lat)
    E.stream_compact()
    for name in ("f", "g"):
        assert np.array_equal(E.valid(name), O.valid(name)), f"compact stream of {name} differs"


FUSED_CASES = {
    # the literal Taylor-Green workload (tau = 0.6, T = 1/3) stays finite for
    # ~50 steps before the reference model's energy-lattice instability
    "tg_tau0.6_40steps": (lambda: configs.taylor_green(32), 40),
    "tg_tau1.0_150steps": (lambda: configs.taylor_green(32, tau=1.0), 150),
    "channel_T0.2_120steps": (
        lambda: configs.channel_sphere((48, 24, 24), radius=5.0, center=(16.0, 12.0, 12.0)), 120),
}


@pytest.mark.parametrize("case", list(FUSED_CASES))
def test_compact_fused_steps(oracle, engine, case):
    from oracle.pyoracle import build_oracle_level

    make, nsteps = FUSED_CASES[case]
    cfg = make()
    O = build_oracle_level(cfg)
    E = _compact_from(O, engine, cfg)
    O.evolve(nsteps)
    E.advance(nsteps)
    exact = True
    for name in ("f", "g"):
        a, b = E.valid(name), O.valid(name)
        err = _rel_err(a, b)
        exact &= bool(np.array_equal(a, b))
        assert err <= REL_TOL, (case, name, err)
    print(f"{case}: {nsteps} fused steps, bit-exact={exact}")


@pytest.mark.parametrize("case", ["taylor_green", "channel"])
def test_engine_initial_state_matches_oracle(oracle, engine, case):
    from oracle.pyoracle import build_oracle_level

    cfg = configs.taylor_green(16) if case == "taylor_green" else _small_channel()
    O = build_oracle_level(cfg)
    E = configs.build_level(cfg, layout="compact")
    assert np.array_equal(E.valid("is_fluid"), O.valid("is_fluid"))
    for name in ("f", "g"):
        assert np.array_equal(E.valid(name), O.valid(name)), name
    Ef = configs.build_level(cfg, layout="fab", ng=3)
    assert np.array_equal(Ef.download("is_fluid"), O.flat("is_fluid"))
    for name in ("f", "g"):
        assert np.array_equal(Ef.download(name), O.flat(name)), name


def test_conservation_reduction(oracle, engine):
    from oracle.pyoracle import build_oracle_level

    cfg = configs.taylor_green(16)
    O = build_oracle_level(cfg)
    E = _compact_from(O, engine, cfg)
    O.evolve(5)
    E.advance(5)
    a, b = E.conservation(), O.conservation()
    assert np.allclose(a, b, rtol=1e-12, atol=1e-10), (a, b)


@pytest.mark.parametrize("case", ["tg", "channel", "odd"])
def test_every_hot_kernel_variant_matches_oracle(oracle, engine, case):
    """Every compact-layout kernel variant (two-pass and the single-pass
    z-sweep tile shapes) gives the oracle's populations."""
    from oracle.pyoracle import build_oracle_level

    if case == "tg":
        cfg, n = configs.taylor_green(32), 30
    elif case == "channel":
        cfg, n = configs.channel_sphere((48, 24, 24), radius=5.0, center=(16.0, 12.0, 12.0)), 40
    else:  # sizes that leave partial tiles in x, y and z
        cfg, n = configs.channel_sphere((37, 29, 23), radius=4.0, center=(12.0, 14.0, 11.0)), 20
    O = build_oracle_level(cfg)
    O.evolve(n)
    ref = {name: O.valid(name) for name in ("f", "g")}
    checked = []
    for v, vname in enumerate(engine.Level.hot_kernel_variants()):
        if "_h32" in vname or "_f32" in vname:
            continue  # FP32-arithmetic collides of the reduced storages (their own tests)
        O0 = build_oracle_level(cfg)
        E = _compact_from(O0, engine, cfg)
        E.set_hot_kernel(v)
        E.advance(n)
        for name in ("f", "g"):
            err = _rel_err(E.valid(name), ref[name])
            assert err <= REL_TOL, (case, vname, name, err)
        E.close()
        checked.append(vname)
    assert "k_sweep" in checked and "k_sweep_x64" in checked and "k_onepass" in checked


def test_autotune_does_not_advance_the_state(oracle, engine):
    from oracle.pyoracle import build_oracle_level

    cfg = configs.channel_sphere((40, 24, 24), radius=5.0, center=(14.0, 12.0, 12.0))
    O = build_oracle_level(cfg)
    E = _compact_from(O, engine, cfg)
    f0 = E.valid("f")
    times = E.autotune(reps=2)
    assert len(times) == len(engine.Level.hot_kernel_variants())
    assert np.array_equal(E.valid("f"), f0)
    E.advance(5)
    O.evolve(5)
    assert _rel_err(E.valid("f"), O.valid("f")) <= REL_TOL


def _velocity(E):
    ev, _, _ = __import__("oracle.pyoracle", fromlist=["stencil"]).stencil()
    f = E.valid("f")
    rho = f.sum(axis=0)
    return np.stack([np.tensordot(ev[:, d].astype(float), f, axes=(0, 0)) / rho for d in range(3)]), rho


def _tracking_case(case):
    if case == "tg_tau1.0":
        return configs.taylor_green(32, tau=1.0), 100
    if case == "channel_profile":
        return configs.channel_profile((40, 20, 24), kind="channel"), 100
    if case == "pressure_walled":
        return configs.pressure_case("walled", (24, 16, 32)), 100
    return configs.channel_sphere((48, 24, 24), radius=5.0, center=(16.0, 12.0, 12.0)), 100


@pytest.mark.parametrize("case", ["tg_tau1.0", "channel_T0.2", "channel_profile", "pressure_walled"])
def test_f32_shifted_storage_tracks_f64(oracle, engine, case):
    """FP32 storage of f - f_eq(rho0, 0, T0) with FP64 arithmetic: velocity
    and density within 1e-6 (relative to max |u|, rho) of the FP64 run, also
    with a Channel-profile inlet and a walled pressure outlet."""
    cfg, n = _tracking_case(case)
    runs = {}
    for storage in ("f64", "f32s"):
        E = configs.build_level(cfg, layout="compact", storage=storage)
        if storage == "f32s":
            E.set_hot_kernel("k_qpass_y2+k_collide_y2")  # FP64 arithmetic
        E.advance(n)
        runs[storage] = _velocity(E)
        E.close()
    (u64, r64), (u32, r32) = runs["f64"], runs["f32s"]
    fluid = r64 > 0
    du = np.abs(u32 - u64)[:, fluid].max() / np.abs(u64[:, fluid]).max()
    dr = np.abs(r32 - r64)[fluid].max() / r64[fluid].max()
    print(f"{case}: {n} steps, max rel velocity error {du:.2e}, density {dr:.2e}")
    assert du <= 1e-6 and dr <= 1e-6, (du, dr)


@pytest.mark.parametrize("case", ["tg_tau1.0", "channel_T0.2", "channel_profile"])
def test_f32_shifted_with_fp32_arithmetic_tracks_f64(engine, case):
    """Shifted FP32 storage with the FP32-arithmetic collide in deviation
    form (k_collide_f32, 64x2 and 256x1 blocks, bit-identical to each
    other): after 100 steps velocity and density within 1e-6 of the FP64
    run (measured 0.99e-7, 4.2e-7 and 1.9e-7 in velocity with the momentum
    residual of the FP32 equilibria returned to the axis pairs; 1.2e-6 and
    2.0e-6 without it), and mass conserved to FP64 rounding of the stored
    values."""
    cfg, n = _tracking_case(case)
    runs, cons, fields = {}, {}, {}
    for key, storage, kernel in (("f64", "f64", None), ("f32a", "f32s", "k_qpass_y2+k_collide_f32"),
                                 ("f32b", "f32s", "k_qpass_x256+k_collide_f32_x256")):
        E = configs.build_level(cfg, layout="compact", storage=storage)
        if kernel:
            E.set_hot_kernel(kernel)
        c0 = E.conservation()
        E.advance(n)
        runs[key] = _velocity(E)
        cons[key] = (c0, E.conservation())
        fields[key] = (E.valid("f"), E.valid("g"))
        E.close()
    for a, b in zip(fields["f32a"], fields["f32b"]):
        assert np.array_equal(a, b), case
    (u64, r64), (u32, r32) = runs["f64"], runs["f32a"]
    fluid = r64 > 0
    du = np.abs(u32 - u64)[:, fluid].max() / np.abs(u64[:, fluid]).max()
    dr = np.abs(r32 - r64)[fluid].max() / r64[fluid].max()
    print(f"{case}: FP32 arithmetic, max rel velocity error {du:.2e}, density {dr:.2e}")
    assert du <= 1e-6 and dr <= 1e-6, (du, dr)
    if case == "tg_tau1.0":  # periodic: mass is conserved
        m0, m1 = cons["f32a"][0][0], cons["f32a"][1][0]
        assert abs(m1 - m0) <= 1e-9 * m0, (m0, m1)


def test_f32_arithmetic_variant_survives_a_pressure_face(engine):
    """A shifted FP32 level on the FP32-arithmetic collide that gets a
    PRESSURE face steps on the FP64-arithmetic collide of the same block
    shape, and returns to its own when the faces change back."""
    cfg = configs.pressure_case("walled", (24, 16, 32))
    E = configs.build_level(cfg, layout="compact", storage="f32s")
    E.set_bc((1, 0, 1), (1, 0, 1))
    E.set_hot_kernel("k_qpass_x256+k_collide_f32_x256")
    E.set_bc(cfg["bc_lo"], cfg["bc_hi"], cfg.get("velocity_bc"))
    assert E.hot_kernel_name() == "k_qpass_x256+k_collide_x256"
    E.advance(5)
    E.set_bc((1, 0, 1), (1, 0, 1))
    assert E.hot_kernel_name() == "k_qpass_x256+k_collide_f32_x256"
    E.advance(5)
    assert np.isfinite(E.valid("f")).all()
    E.close()


@pytest.mark.parametrize("case", ["nz1", "nz2", "nz3", "ny1", "row256", "mostly_solid"])
def test_sweep_edge_shapes_bitwise(engine, case):
    """The z-sweeps at the shapes their prologue, epilogue and lane
    handling treat apart: one, two and three planes (the planes past the
    last one are replaced by the last), one row, rows exactly 256 cells
    wide (no lane past the row's end), and a sphere filling most of the
    domain (waves of solid cells only, which pull and collide like fluid
    ones and store the solid marker).  FP64 and shifted FP32 storage, both
    arithmetics, against the two-pass step bit for bit over 6 steps."""
    if case == "nz1":
        cfg = configs.channel_sphere((24, 12, 1), radius=3.0, center=(8.0, 6.0, 0.0))
    elif case == "nz2":
        cfg = configs.channel_sphere((24, 16, 2), radius=3.0, center=(8.0, 8.0, 1.0))
    elif case == "nz3":
        cfg = configs.channel_sphere((24, 16, 3), radius=3.0, center=(8.0, 8.0, 1.0))
    elif case == "ny1":
        cfg = configs.channel_sphere((24, 1, 16), radius=3.0, center=(8.0, 0.0, 8.0))
    elif case == "row256":
        cfg = configs.channel_sphere((256, 6, 8), radius=3.0, center=(40.0, 3.0, 4.0))
    else:
        cfg = configs.channel_sphere((32, 16, 16), radius=11.0, center=(16.0, 8.0, 8.0))
    runs = (("f64", ["k_qpass+k_collide", "k_sweep", "k_sweep_x64"]),
            ("f32s", ["k_qpass+k_collide", "k_sweep"]),
            ("f32s", ["k_qpass_y2+k_collide_f32", "k_sweep_f32"]))
    for storage, names in runs:
        out = {}
        for name in names:
            E = configs.build_level(cfg, layout="compact", storage=storage)
            E.set_hot_kernel(name)
            assert E.hot_kernel_name() == name
            E.advance(6)
            out[name] = (E.valid("f"), E.valid("g"))
            E.close()
        ref = out[names[0]]
        assert np.isfinite(ref[0]).all() and np.isfinite(ref[1]).all(), (case, storage)
        for name in names[1:]:
            for a, b in zip(out[name], ref):
                assert np.array_equal(a, b), (case, storage, name)


@pytest.mark.parametrize("case", ["tg_tau1.0", "channel", "odd", "channel_profile"])
def test_fp32_arithmetic_sweep_bitwise(engine, case):
    """The z-sweep with the FP32-arithmetic collide (k_sweep_f32: stored
    FP32 differences held in LDS, neighbours' Q_corr formed in FP64 from the
    populations as pass 1 forms them) equals the two-pass FP32-arithmetic
    step bit for bit, over 20 steps: periodic, an EB sphere, sizes that
    leave partial planes, and a Channel-profile inlet."""
    if case == "tg_tau1.0":
        cfg = configs.taylor_green(32, tau=1.0)
    elif case == "channel":
        cfg = configs.channel_sphere((48, 24, 24), radius=5.0, center=(16.0, 12.0, 12.0))
    elif case == "odd":
        cfg = configs.channel_sphere((37, 29, 23), radius=4.0, center=(12.0, 14.0, 11.0))
    else:
        cfg = configs.channel_profile((40, 20, 24), kind="channel")
    out = {}
    for name in ("k_qpass_y2+k_collide_f32", "k_sweep_f32"):
        E = configs.build_level(cfg, layout="compact", storage="f32s")
        E.set_hot_kernel(name)
        assert E.hot_kernel_name() == name
        E.advance(20)
        out[name] = (E.valid("f"), E.valid("g"))
        E.close()
    (fa, ga), (fb, gb) = out.values()
    assert np.isfinite(fa).all() and np.isfinite(ga).all()
    assert np.array_equal(fa, fb) and np.array_equal(ga, gb), case


def test_fp32_arithmetic_sweep_yields_to_a_pressure_face(engine):
    """k_sweep_f32 has no PRESSURE faces: a level on it that gets one steps
    on the FP64-arithmetic sweep of the same storage, and returns to the
    FP32 one when the faces change back."""
    cfg = configs.pressure_case("walled", (24, 16, 32))
    E = configs.build_level(cfg, layout="compact", storage="f32s")
    E.set_bc((1, 0, 1), (1, 0, 1))
    E.set_hot_kernel("k_sweep_f32")
    E.set_bc(cfg["bc_lo"], cfg["bc_hi"], cfg.get("velocity_bc"))
    assert E.hot_kernel_name() == "k_sweep"
    E.advance(5)
    E.set_bc((1, 0, 1), (1, 0, 1))
    assert E.hot_kernel_name() == "k_sweep_f32"
    E.advance(5)
    assert np.isfinite(E.valid("f")).all()
    E.close()


@pytest.mark.parametrize("case", ["tg_tau1.0", "channel", "channel_profile", "pressure_walled"])
def test_f16_scaled_storage_tracks_f64(oracle, engine, case):
    """FP16 storage of (f - f_eq(rho0, 0, T0)) * 2^e, e per (z-slab box,
    lattice), the level's default step (FP32 arithmetic in the collide, FP64
    with PRESSURE faces): velocity and density against the FP64 run after
    100 steps (FP16 keeps 11 significant bits, so the bars are set by that,
    not by 1e-6; measured 3.8e-4 / 8.2e-4 velocity); no stored value
    saturates; conservation sums in FP64 agree with the FP64 run's; the
    shift cannot change once populations are stored."""
    cfg, n = _tracking_case(case)
    runs, cons = {}, {}
    for storage in ("f64", "f16s"):
        E = configs.build_level(cfg, layout="compact", storage=storage)
        if storage == "f16s":
            # the FP32-arithmetic collide, or FP64 arithmetic with PRESSURE faces
            pressure = case.startswith("pressure")
            assert E.hot_kernel_name() == ("k_qpass+k_collide" if pressure else "k_qpass+k_collide_h32")
        E.advance(n)
        runs[storage] = _velocity(E)
        cons[storage] = E.conservation()
        if storage == "f16s":
            assert E.storage_saturations() == 0
            from marbles_amd._lib import MrbError

            with pytest.raises(MrbError):
                E.set_shift_state(1.0, 0.25)
        E.close()
    (u64, r64), (u16, r16) = runs["f64"], runs["f16s"]
    fluid = r64 > 0
    du = np.abs(u16 - u64)[:, fluid].max() / np.abs(u64[:, fluid]).max()
    dr = np.abs(r16 - r64)[fluid].max() / r64[fluid].max()
    dm = abs(cons["f16s"][0] - cons["f64"][0]) / abs(cons["f64"][0])
    print(f"{case}: {n} steps, FP16 max rel velocity error {du:.2e}, density {dr:.2e}, mass {dm:.2e}")
    assert du <= 5e-3 and dr <= 1e-3 and dm <= 1e-4, (du, dr, dm)


@pytest.mark.parametrize("case", ["tg_tau1.0", "channel_T0.2", "channel_profile"])
def test_f16_storage_with_fp32_arithmetic_tracks_f64(oracle, engine, case):
    """Scaled FP16 storage with FP32 arithmetic in the collide pass
    (k_collide_h32): after 100 steps its velocity and density errors against
    the FP64 run stay within the FP16 bars and within 1.5x (plus 1e-5) of
    the errors of the same storage with FP64 arithmetic, since the FP16
    rounding of the stored deviations dominates the FP32 rounding."""
    cfg, n = _tracking_case(case)
    runs = {}
    for key, storage, kernel in (("f64", "f64", None), ("f16s", "f16s", None),
                                 ("h32", "f16s", "k_qpass+k_collide_h32")):
        E = configs.build_level(cfg, layout="compact", storage=storage)
        if kernel:
            E.set_hot_kernel(kernel)
        E.advance(n)
        runs[key] = _velocity(E)
        if storage == "f16s":
            assert E.storage_saturations() == 0
        E.close()
    u64, r64 = runs["f64"]
    fluid = r64 > 0
    err = {}
    for key in ("f16s", "h32"):
        u, r = runs[key]
        err[key] = (np.abs(u - u64)[:, fluid].max() / np.abs(u64[:, fluid]).max(),
                    np.abs(r - r64)[fluid].max() / r64[fluid].max())
    print(f"{case}: FP64 arithmetic {err['f16s']}, FP32 arithmetic {err['h32']}")
    du, dr = err["h32"]
    assert du <= 5e-3 and dr <= 1e-3, err
    assert du <= 1.5 * err["f16s"][0] + 1e-5 and dr <= 1.5 * err["f16s"][1] + 1e-5, err


@pytest.mark.parametrize("case", ["channel", "odd", "channel_profile"])
def test_fp32_arithmetic_collide_block_shapes_bitwise(engine, case):
    """The FP32-arithmetic FP16 collide on 256x1-cell blocks equals the
    64x4 one bit for bit (same operations per cell)."""
    if case == "channel":
        cfg = configs.channel_sphere((48, 24, 24), radius=5.0, center=(16.0, 12.0, 12.0))
    elif case == "odd":
        cfg = configs.channel_sphere((37, 29, 23), radius=4.0, center=(12.0, 14.0, 11.0))
    else:
        cfg = configs.channel_profile((40, 20, 24), kind="channel")
    out = {}
    for name in ("k_qpass+k_collide_h32", "k_qpass_x256+k_collide_h32_x256"):
        E = configs.build_level(cfg, layout="compact", storage="f16s")
        E.set_hot_kernel(name)
        E.advance(20)
        out[name] = (E.valid("f"), E.valid("g"))
        E.close()
    for a, b in zip(*out.values()):
        assert np.array_equal(a, b), case


def test_f16_scaled_round_trip(oracle, engine):
    """Upload then download through the FP16 storage: every fluid value
    within 2^-11 (FP16 unit roundoff) of its box's largest |f - shift|;
    solid cells read back as the stream's -1 sentinel."""
    cfg = configs.channel_sphere((32, 16, 24), radius=4.0, center=(10.0, 8.0, 12.0))
    E64 = configs.build_level(cfg, layout="compact", storage="f64")
    E16 = configs.build_level(cfg, layout="compact", storage="f16s")
    isf = E64.valid("is_fluid")[0]
    for lat in ("f", "g"):
        a = E64.valid(lat)
        E16.upload(lat, a)
        b = E16.valid(lat)
        shift = np.array(E16.shift(lat))[:, None, None, None]
        fluid = isf == 1
        for z0 in range(0, a.shape[1], 8):
            sl = slice(z0, z0 + 8)
            fz = fluid[sl]
            amax = np.abs(a[:, sl] - shift)[:, fz].max()
            err = np.abs(b[:, sl] - a[:, sl])[:, fz].max()
            assert err <= amax * 2.0 ** -11 * (1 + 1e-12), (lat, z0, err, amax)
        assert np.all(b[:, ~fluid] == -1.0)
    E64.close()
    E16.close()


@pytest.mark.parametrize("layout", ["compact", "fab", "compact_f32s"])
def test_eb_force_matches_oracle(oracle, engine, layout):
    """Momentum-exchange force on the channel's sphere after 20 steps (FP64
    sums in a different order: 1e-12 relative; FP32-shifted storage: the
    populations differ by ~1e-7 relative, so 1e-5)."""
    from oracle.pyoracle import build_oracle_level

    cfg = configs.channel_sphere((48, 24, 24), radius=5.0, center=(16.0, 12.0, 12.0))
    O = build_oracle_level(cfg)
    O.evolve(20)
    ref = O.eb_force()
    if layout == "fab":
        O0 = build_oracle_level(cfg)
        E = engine.Level(cfg["ncell"], cfg["periodic"], cfg["prob_lo"], cfg["prob_hi"], boxes=O0.boxes, ng=3,
                         layout="fab")
        E.set_phys(cfg["nu"], cfg.get("alpha"))
        E.set_bc(cfg["bc_lo"], cfg["bc_hi"], cfg.get("velocity_bc"))
        E.flags_upload(O0.flat("is_fluid"))
        E.upload("f", O0.flat("f"))
        E.upload("g", O0.flat("g"))
    else:
        E = configs.build_level(cfg, layout="compact", storage="f32s" if layout == "compact_f32s" else "f64")
    E.advance(20)
    got = E.eb_force()
    tol = 1e-5 if layout == "compact_f32s" else 1e-12
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max(), (layout, got, ref)
    E.close()


@pytest.mark.parametrize("kind", ["channel", "parabolic"])
@pytest.mark.parametrize("layout", ["compact", "fab", "fab_batched"])
def test_velocity_profile_inlet_matches_oracle(oracle, engine, kind, layout):
    """Walled channel driven by a Channel or Parabolic inlet profile (ghost
    values from the ghost cell's index): compact two-pass step, fab stages
    and the batched fab step against the oracle after 30 steps."""
    from oracle.pyoracle import build_oracle_level

    cfg = configs.channel_profile((40, 20, 24), kind=kind, max_box=None if layout == "compact" else 12)
    O = build_oracle_level(cfg)
    if layout == "compact":
        E = _compact_from(O, engine, cfg)
    else:
        O, E = _fab_pair(oracle, engine, cfg)
    n = 30
    if layout == "fab":
        for _ in range(n):
            for lat in (0, 1):
                E.fillpatch(lat)
                E.stream(lat)
            E.collide()
    else:
        E.advance(n)
    O.evolve(n)
    for name in ("f", "g"):
        a, b = E.valid(name), O.valid(name)
        err = _rel_err(a, b)
        assert err <= REL_TOL, (kind, layout, name, err)
    u = O.valid("f")
    assert np.abs(np.tensordot(np.array([c[0] for c in oracle.stencil()[0]], float), u, axes=(0, 0))).max() > 1e-3
    E.close()


PRESSURE_CASES = {"outlet": 30, "outlet_body": 2, "walled": 30, "both": 30, "channel": 30}


@pytest.mark.parametrize("case", list(PRESSURE_CASES))
@pytest.mark.parametrize("layout", ["compact", "fab", "fab_batched"])
def test_pressure_boundary_matches_oracle(oracle, engine, case, layout):
    """PRESSURE faces (pressure_bc_f / pressure_bc_g): an outlet behind a
    velocity inlet, a walled outlet (edge ghosts beyond a wall and the face),
    pressure at both ends, a Channel-profile inlet with a pressure outlet;
    and two steps of an outlet whose sums reach solid cells (their contents
    enter the sums: 0 in the initial state, -1 after a stream).  Compact
    two-pass step, fab stages and the batched fab step against the oracle."""
    from oracle.pyoracle import build_oracle_level

    cfg = configs.pressure_case(case, max_box=None if layout == "compact" else 12)
    O = build_oracle_level(cfg)
    if layout == "compact":
        E = _compact_from(O, engine, cfg)
    else:
        O, E = _fab_pair(oracle, engine, cfg)
    n = PRESSURE_CASES[case]
    if layout == "fab":
        for _ in range(n):
            for lat in (0, 1):
                E.fillpatch(lat)
                E.stream(lat)
            E.collide()
    else:
        E.advance(n)
    O.evolve(n)
    for name in ("f", "g"):
        a, b = E.valid(name), O.valid(name)
        assert np.isfinite(b).all()
        err = _rel_err(a, b)
        assert err <= REL_TOL, (case, layout, name, err)
    E.close()


@pytest.mark.parametrize("layout", ["compact", "fab", "fab_batched"])
def test_slip_walls_match_oracle(oracle, engine, layout):
    """Slip walls (SLIPWALLXNORMAL on both x faces, BC.H:456-462) with the
    single_rotated_box deck's other faces (velocity inlet along y at y-lo,
    outflow at y-hi, periodic z) and a sphere: compact two-pass step, fab
    stages and the batched fab step against the oracle after 30 steps."""
    from oracle.pyoracle import build_oracle_level

    cfg = configs.slip_channel((24, 40, 16), max_box=None if layout == "compact" else 12)
    O = build_oracle_level(cfg)
    if layout == "compact":
        E = _compact_from(O, engine, cfg)
    else:
        O, E = _fab_pair(oracle, engine, cfg)
    n = 30
    if layout == "fab":
        for _ in range(n):
            for lat in (0, 1):
                E.fillpatch(lat)
                E.stream(lat)
            E.collide()
    else:
        E.advance(n)
    O.evolve(n)
    for name in ("f", "g"):
        a, b = E.valid(name), O.valid(name)
        assert np.isfinite(b).all()
        err = _rel_err(a, b)
        assert err <= REL_TOL, (layout, name, err)
    E.close()


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_slip_wall_every_axis_compact(oracle, engine, axis):
    """Slip walls normal to each axis (the other two periodic) around a
    sphere, started from a flow along the next axis: compact step against
    the oracle after 20 steps."""
    from oracle.pyoracle import build_oracle_level

    n = (20, 18, 16)
    per = [1, 1, 1]
    per[axis] = 0
    bc = [0, 0, 0]
    bc[axis] = 6 + axis
    mach = [0.0, 0.0, 0.0]
    mach[(axis + 1) % 3] = 0.04 / np.sqrt(5.0 / 3.0 * 0.2)
    cfg = {"ncell": n, "periodic": tuple(per), "prob_lo": (0.0, 0.0, 0.0), "prob_hi": tuple(float(v) for v in n),
           "bc_lo": tuple(bc), "bc_hi": tuple(bc), "velocity_bc": None, "nu": 0.02, "T0": 0.2,
           "ic": {"type": "constant", "density": 1.0, "mach": tuple(mach), "temperature": 0.2,
                  "adiabatic_exponent": 5.0 / 3.0},
           "eb": {"type": "sphere", "center": (10.0, 9.0, 8.0), "radius": 3.5, "has_fluid_inside": 0},
           "max_box": None}
    O = build_oracle_level(cfg)
    E = _compact_from(O, engine, cfg)
    O.evolve(20)
    E.advance(20)
    for name in ("f", "g"):
        err = _rel_err(E.valid(name), O.valid(name))
        assert err <= REL_TOL, (axis, name, err)
    E.close()


def test_pressure_bc_rejects_fill_order_dependent_faces(engine):
    """A PRESSURE face whose tangential axis is neither periodic nor no-slip
    on both sides, or whose own axis has one cell, is refused."""
    from marbles_amd._lib import MrbError

    E = engine.Level((8, 6, 4), (0, 0, 1), (0.0, 0.0, 0.0), (8.0, 6.0, 4.0), layout="compact")
    E.set_bc((2, 1, 0), (3, 1, 0))
    for lo, hi in (((2, 5, 0), (3, 1, 0)), ((2, 2, 0), (3, 1, 0)), ((2, 6, 0), (3, 6, 0)), ((2, 3, 0), (5, 1, 0))):
        with pytest.raises(MrbError):
            E.set_bc(lo, hi)
    E.close()
    F = engine.Level((1, 6, 4), (0, 1, 1), (0.0, 0.0, 0.0), (1.0, 6.0, 4.0), layout="compact")
    with pytest.raises(MrbError):
        F.set_bc((3, 0, 0), (1, 0, 0))
    F.close()


PLUGIN_CASES = {
    "viscosity_test": (lambda: configs.shear_wave(32, tau=0.8), 60),
    "thermaldiffusivity_test": (lambda: configs.plugin_case(
        "thermaldiffusivity_test", (4, 32, 4), {"wave_length": 32.0, "initial_temperature": 0.2},
        nu=0.02, alpha=0.03), 60),
    "sod": (lambda: configs.sod_tube(64), 40),
}


@pytest.mark.parametrize("case", list(PLUGIN_CASES))
@pytest.mark.parametrize("layout", ["compact", "fab"])
def test_ic_plugin_workloads_match_oracle(oracle, engine, case, layout):
    """The reference's ViscosityTest, ThermalDiffusivityTest and SodTest
    initial states (evaluated per fab cell on the host, equilibria on the
    device) and the steps that follow, against the oracle."""
    from oracle.pyoracle import build_oracle_level

    make, n = PLUGIN_CASES[case]
    cfg = make()
    if layout == "fab":
        cfg["max_box"] = 16
    O = build_oracle_level(cfg)
    E = configs.build_level(cfg, layout=layout)
    for name in ("f", "g"):
        assert _rel_err(E.valid(name), O.valid(name)) <= REL_TOL, (case, layout, "initial", name)
    O.evolve(n)
    E.advance(n)
    for name in ("f", "g"):
        err = _rel_err(E.valid(name), O.valid(name))
        assert err <= REL_TOL, (case, layout, name, err)
    E.close()


@pytest.mark.parametrize("storage", ["f64", "f32s", "f16s"])
@pytest.mark.parametrize("case", ["channel", "odd", "channel_profile", "pressure_walled", "outlet", "slip"])
def test_step_variants_equal_two_pass_bitwise(engine, case, storage):
    """The other block shapes of the two passes (64x8, 64x2, 128x2, 256x1),
    k_onepass, which recomputes the neighbours' Q_corr from their pulls
    instead of reading pass 1's array, and the single-pass z-sweeps k_sweep
    (FP64 and FP32 storage) and k_sweep_x64 (FP64), which pass Q_corr
    between rows, do the same operations in the
    same order: their populations equal the 64x4 two-pass step's bit for bit
    (partial tiles, the sphere, inlet profiles, pressure ghosts, slip
    walls)."""
    if case == "channel":
        cfg, n = configs.channel_sphere((48, 24, 24), radius=5.0, center=(16.0, 12.0, 12.0)), 20
    elif case == "odd":
        cfg, n = configs.channel_sphere((37, 29, 23), radius=4.0, center=(12.0, 14.0, 11.0)), 20
    elif case == "channel_profile":
        cfg, n = configs.channel_profile((40, 20, 24), kind="channel"), 20
    elif case == "pressure_walled":
        cfg, n = configs.pressure_case("walled", (24, 16, 32)), 20
    elif case == "slip":
        cfg, n = configs.slip_channel((24, 40, 16)), 20
    else:
        cfg, n = configs.pressure_case("outlet", (32, 16, 16)), 20
    names = ["k_qpass+k_collide", "k_qpass_y8+k_collide_y8", "k_qpass_y2+k_collide_y2",
             "k_qpass_x128+k_collide_x128", "k_qpass_x256+k_collide_x256"]
    if storage != "f16s":
        names.append("k_onepass")
    if storage == "f64":
        names += ["k_sweep", "k_sweep_x64"]
    elif storage == "f32s":
        names.append("k_sweep")
    out = {}
    for name in names:
        E = configs.build_level(cfg, layout="compact", storage=storage)
        E.set_hot_kernel(name)
        E.advance(n)
        out[name] = (E.valid("f"), E.valid("g"))
        E.close()
    for name in names[1:]:
        for a, b in zip(out[name], out[names[0]]):
            assert np.array_equal(a, b), (case, storage, name)
