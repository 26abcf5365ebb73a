"""Scaled FP16 storage: scale choice next to quiescent boxes, variant
selection with PRESSURE faces, and the autotuner's effect on the
saturation count."""
import numpy as np
import pytest

from marbles_amd import configs

pytestmark = pytest.mark.gpu


def _quiescent_box_with_bump(engine, storage):
    """32^3 periodic box at rest (rho = 1, T0 = 0.2) except a velocity bump
    confined to the z-planes 16..23 (one scale box): every other scale box
    starts exactly at the shift state."""
    n, T0 = 32, 0.2
    E = engine.Level((n, n, n), (1, 1, 1), (0.0, 0.0, 0.0), (float(n),) * 3, layout="compact", storage=storage)
    E.set_phys((0.6 - 0.5) * T0)
    E.set_bc((0, 0, 0), (0, 0, 0))
    if storage != "f64":
        E.set_shift_state(1.0, T0)
    E.flags_all_regular()
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    r2 = (x - 16.0) ** 2 + (y - 16.0) ** 2 + (z - 19.5) ** 2
    inside = (z >= 16) & (z < 24)
    u = np.where(inside, 0.03 * np.exp(-r2 / 8.0), 0.0).ravel()
    zero = np.zeros(n ** 3)
    E.init_fields(np.ones(n ** 3), u, zero, zero.copy(), np.full(n ** 3, T0))
    return E


def test_quiescent_neighbour_boxes_do_not_saturate(engine):
    """A disturbance pulled into a scale box that was exactly at the shift
    state is stored at the disturbed neighbour's scale: no value saturates
    and the velocity tracks the FP64 run within the FP16 bar."""
    ev = np.array(__import__("oracle.pyoracle", fromlist=["stencil"]).stencil()[0], dtype=float)
    out = {}
    for storage in ("f64", "f16s"):
        E = _quiescent_box_with_bump(engine, storage)
        E.advance(40)
        f = E.valid("f")
        rho = f.sum(axis=0)
        out[storage] = np.stack([np.tensordot(ev[:, d], f, axes=(0, 0)) / rho for d in range(3)])
        if storage == "f16s":
            assert E.storage_saturations() == 0
        E.close()
    u64, u16 = out["f64"], out["f16s"]
    du = float(np.abs(u16 - u64).max() / np.abs(u64).max())
    # the bump has spread over every scale box by step 40
    assert np.abs(u64[:, :4]).max() > 1e-9
    assert du <= 5e-3, du


def test_fp32_arithmetic_variant_survives_a_pressure_face(engine):
    """An FP16 level set to the 256x1-block FP32-arithmetic collide that
    then gets a PRESSURE face steps on the FP64-arithmetic 256x1 variant,
    and returns to its own when the faces change back."""
    cfg = configs.pressure_case("walled", (24, 16, 32))
    E = configs.build_level(cfg, layout="compact", storage="f16s")
    E.set_bc((1, 0, 1), (1, 0, 1))  # walls only
    E.set_hot_kernel("k_qpass_x256+k_collide_h32_x256")
    E.set_bc(cfg["bc_lo"], cfg["bc_hi"], cfg.get("velocity_bc"))
    assert E.hot_kernel_name() == "k_qpass_x256+k_collide_x256"
    E.advance(5)
    E.set_bc((1, 0, 1), (1, 0, 1))
    assert E.hot_kernel_name() == "k_qpass_x256+k_collide_h32_x256"
    E.advance(5)
    assert np.isfinite(E.valid("f")).all()
    E.close()


def test_autotune_keeps_arithmetic_and_saturation_count(engine):
    """The autotuner on an FP16 level tries only the variants with the
    current collide arithmetic and leaves the saturation count unchanged."""
    cfg = configs.channel_sphere((48, 24, 24), radius=5.0, center=(16.0, 12.0, 12.0))
    E = configs.build_level(cfg, layout="compact", storage="f16s")
    E.advance(3)
    before = E.storage_saturations()
    times = E.autotune(reps=2)
    assert E.storage_saturations() == before
    tried = [k for k, v in times.items() if v >= 0]
    assert tried and all("_h32" in k for k in tried), tried
    assert "_h32" in E.hot_kernel_name()
    E.close()
