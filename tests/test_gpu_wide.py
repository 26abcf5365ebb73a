"""GPU parity at the benchmarked shapes.

The compact-layout kernels run rows of 64, 128 or 256 cells per block; a
domain wider than 256 cells puts several blocks on each x-row, the sphere
across the x = 256 block edge and both end lanes of every row of k_onepass
inside the domain.  The workloads of the benchmark are checked at the sizes
they are measured on (the 256^3 channel), or as close as the CPU oracle
allows in seconds (the 64^3 Taylor-Green vortex, the spectral field).

Bar: populations within 1e-12 relative of the oracle (FP64 storage); the
reduced-precision storages against the FP64 run within their bars.
"""
import numpy as np
import pytest

from marbles_amd import configs
from marbles_amd._lib import MRB_E_UNSUPPORTED, MrbError

pytestmark = pytest.mark.gpu

REL_TOL = 1e-12


def _rel_err(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale)


def _wide_channel():
    # 272 cells in x: one full 256-wide row block plus a partial one, four
    # 64-wide blocks plus a partial one; the sphere straddles x = 256
    return configs.channel_sphere((272, 24, 16), radius=5.0, center=(256.0, 12.0, 8.0))


def _compact_from(O, engine, cfg, storage="f64"):
    E = engine.Level(cfg["ncell"], cfg["periodic"], cfg["prob_lo"], cfg["prob_hi"], layout="compact",
                     storage=storage)
    E.set_phys(cfg["nu"], cfg.get("alpha"))
    E.set_bc(cfg["bc_lo"], cfg["bc_hi"], cfg.get("velocity_bc"))
    if storage != "f64":
        E.set_shift_state(*configs.reference_state(cfg))
    E.flags_upload(O.valid("is_fluid"))
    E.upload("f", O.valid("f"))
    E.upload("g", O.valid("g"))
    return E


def test_wide_rows_every_variant_matches_oracle(oracle, engine):
    """Every compact-layout step variant of FP64 storage on a channel 272
    cells wide, 20 steps, against the oracle."""
    from oracle.pyoracle import build_oracle_level

    cfg, n = _wide_channel(), 20
    O0 = build_oracle_level(cfg)
    O = build_oracle_level(cfg)
    O.evolve(n)
    ref = {name: O.valid(name) for name in ("f", "g")}
    assert np.isfinite(ref["f"]).all()
    checked = []
    for v, vname in enumerate(engine.Level.hot_kernel_variants()):
        if "_h32" in vname or "_f32" in vname:
            continue  # FP32-arithmetic collides of the reduced storages
        E = _compact_from(O0, engine, cfg)
        E.set_hot_kernel(v)
        try:
            E.advance(n)
        except MrbError as e:
            assert e.code == MRB_E_UNSUPPORTED and vname == "k_sweep", (vname, e)  # rows over 256 cells
            E.close()
            continue
        for name in ("f", "g"):
            err = _rel_err(E.valid(name), ref[name])
            assert err <= REL_TOL, (vname, name, err)
        E.close()
        checked.append(vname)
    assert "k_qpass_x256+k_collide_x256" in checked and "k_onepass" in checked and "k_sweep_x64" in checked
    print("wide rows: " + ", ".join(checked))


@pytest.mark.parametrize("storage", ["f32s", "f16s"])
def test_wide_rows_reduced_storages(oracle, engine, storage):
    """Shifted FP32 and scaled FP16 storage on the 272-wide channel: every
    block shape gives the same populations bit for bit, and the velocity
    stays within the storage's bar of the FP64 run after 60 steps."""
    cfg, n = _wide_channel(), 60
    names = [v for v in engine.Level.hot_kernel_variants() if v.startswith("k_qpass") or v == "k_onepass"]
    if storage == "f16s":
        names = [v for v in names if v != "k_onepass" and "_w3" not in v and "_f32" not in v]
    else:
        names = [v for v in names if "_h32" not in v]
    out = {}
    for name in names:
        E = configs.build_level(cfg, layout="compact", storage=storage)
        E.set_hot_kernel(name)
        E.advance(n)
        out[name] = (E.valid("f"), E.valid("g"))
        if storage == "f16s":
            assert E.storage_saturations() == 0
        E.close()
    groups = {}
    for name in names:
        groups.setdefault("fp32" if ("_h32" in name or "_f32" in name) else "f64", []).append(name)
    for group in groups.values():
        for name in group[1:]:
            for a, b in zip(out[name], out[group[0]]):
                assert np.array_equal(a, b), (storage, name, group[0])
    E = configs.build_level(cfg, layout="compact", storage="f64")
    E.advance(n)
    f64 = E.valid("f")
    E.close()
    ev = np.array(__import__("oracle.pyoracle", fromlist=["stencil"]).stencil()[0], dtype=float)

    def vel(f):
        rho = f.sum(axis=0)
        fl = rho > 0
        return np.stack([np.tensordot(ev[:, d], f, axes=(0, 0))[fl] / rho[fl] for d in range(3)])

    u64 = vel(f64)
    for name in names:
        # FP32 storage: 1e-6 with FP64 or FP32 arithmetic
        bar = 1e-6 if storage == "f32s" else 5e-3
        du = float(np.abs(vel(out[name][0]) - u64).max() / np.abs(u64).max())
        assert du <= bar, (storage, name, du)


def test_taylor_green_64_matches_oracle(oracle, engine):
    """The Taylor-Green workload at its full 64^3 size (tau = 0.6, T = 1/3),
    40 fused steps on the level's default kernel, against the oracle; the
    reference model's energy lattice amplifies round-off at this point and
    the run leaves double range between steps 60 and 70, so 40 steps is
    the longest comparison that stays finite."""
    from oracle.pyoracle import build_oracle_level

    cfg, n = configs.taylor_green(64), 40
    O = build_oracle_level(cfg)
    E = _compact_from(O, engine, cfg)
    assert E.hot_kernel_name() == "k_qpass_x256+k_collide_x256"
    O.evolve(n)
    E.advance(n)
    for name in ("f", "g"):
        a, b = E.valid(name), O.valid(name)
        assert np.isfinite(b).all()
        err = _rel_err(a, b)
        assert err <= REL_TOL, (name, err)
        print(f"taylor_green_64 {name}: max rel {err:.2e}, bit-exact={np.array_equal(a, b)}")
    E.close()
    O.close()


@pytest.mark.timeout(900)
def test_channel_256_matches_oracle(oracle, engine):
    """The benchmarked workload itself: the 256^3 channel with the r = 24
    sphere, 3 steps of the level's default kernel (k_qpass_x256 +
    k_collide_x256) and of the single-pass sweep the bench's autotuner
    selects at this size (k_sweep), each against the oracle on 64^3 boxes."""
    from oracle.pyoracle import build_oracle_level

    cfg, n = configs.channel_sphere((256, 256, 256), max_box=64), 3
    O = build_oracle_level(cfg)
    O.evolve(n)
    for kernel in (None, "k_sweep"):
        E = configs.build_level(cfg, layout="compact")
        if kernel is None:
            assert E.hot_kernel_name() == "k_qpass_x256+k_collide_x256"
            assert np.array_equal(E.valid("is_fluid"), O.valid("is_fluid"))
        else:
            E.set_hot_kernel(kernel)
        E.advance(n)
        for name in ("f", "g"):
            a = E.valid(name)
            b = O.valid(name)
            err = _rel_err(a, b)
            exact = bool(np.array_equal(a, b))
            del a, b
            assert err <= REL_TOL, (E.hot_kernel_name(), name, err)
            print(f"channel_256 {E.hot_kernel_name()} {name}: max rel {err:.2e}, bit-exact={exact}")
        E.close()
    O.close()


@pytest.mark.timeout(600)
def test_sweep_long_run_equals_two_pass_256(engine):
    """1000 steps of the 256^3 channel (all 256 rows resident, neighbours'
    Q_y through the tagged ring, one block per CU) on the single-pass sweep
    and on the two-pass step: bit for bit equal at the end, so no ring
    entry was read stale or overwritten early in 256 x 256 x 1000
    row-plane exchanges."""
    cfg, n = configs.channel_sphere((256, 256, 256)), 1000
    out = {}
    for kernel in ("k_qpass_x256+k_collide_x256", "k_sweep"):
        E = configs.build_level(cfg, layout="compact")
        E.set_hot_kernel(kernel)
        E.advance(n)
        out[kernel] = E.conservation()
        for name in ("f", "g"):
            out[(kernel, name)] = E.valid(name)
        E.close()
    for name in ("f", "g"):
        a, b = out[("k_sweep", name)], out[("k_qpass_x256+k_collide_x256", name)]
        assert np.isfinite(b).all()
        assert np.array_equal(a, b), name
        del a, b
    print("sweep 1000 steps: conservation", out["k_sweep"])


def _spectral(n=64, kmax=8):
    # tau = 0.55: at 0.52 the thermal model diverges on this field (the
    # precision study's point); T0 = 0.2
    return configs.spectral_turbulence(n, kmax=kmax, amplitude=0.04, tau=0.55)


def test_spectral_field_matches_oracle(oracle, engine):
    """The precision workload's spectral field scaled down (64^3, k <= 8):
    initial state and 10 FP64 steps against the oracle."""
    from oracle.pyoracle import build_oracle_level

    cfg, n = _spectral(), 10
    O = build_oracle_level(cfg)
    E = configs.build_level(cfg, layout="compact")
    for name in ("f", "g"):
        assert _rel_err(E.valid(name), O.valid(name)) <= REL_TOL, ("initial", name)
    O.evolve(n)
    E.advance(n)
    for name in ("f", "g"):
        err = _rel_err(E.valid(name), O.valid(name))
        assert err <= REL_TOL, (name, err)
    E.close()
    O.close()


def test_spectral_field_reduced_storages_track_f64(engine):
    """Shifted FP32 storage within 1e-6 and scaled FP16 within its bars of
    the FP64 run after 200 steps on the scaled spectral field (relative to
    max |u| and max rho), kinetic energy included."""
    cfg, n = _spectral(), 200
    ev = np.array(__import__("oracle.pyoracle", fromlist=["stencil"]).stencil()[0], dtype=float)
    res = {}
    for storage in ("f64", "f32s", "f16s"):
        E = configs.build_level(cfg, layout="compact", storage=storage)
        E.advance(n)
        f = E.valid("f").reshape(27, -1)
        rho = f.sum(axis=0)
        u = np.stack([ev[:, d] @ f for d in range(3)]) / rho
        res[storage] = (rho, u, float(0.5 * (rho * (u * u).sum(0)).sum()))
        if storage == "f16s":
            assert E.storage_saturations() == 0
        E.close()
    r0, u0, k0 = res["f64"]
    bars = {"f32s": (1e-6, 1e-6, 1e-6), "f16s": (5e-3, 1e-3, 1e-3)}
    for storage, (bu, br, bk) in bars.items():
        r, u, k = res[storage]
        du = float(np.abs(u - u0).max() / np.abs(u0).max())
        dr = float(np.abs(r - r0).max() / r0.max())
        dk = abs(k - k0) / k0
        print(f"spectral_64 {storage}: velocity {du:.2e}, density {dr:.2e}, kinetic energy {dk:.2e}")
        assert du <= bu and dr <= br and dk <= bk, (storage, du, dr, dk)
