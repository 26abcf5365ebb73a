"""GPU checks of the derived-field and tagging entry points against the
oracle (vorticity, LBM::compute_derived) and a numpy restatement of the
tagging criteria, through the C ABI."""
import numpy as np
import pytest
import torch

from marbles_amd import configs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["channel", "taylorgreen"])
def test_vorticity_matches_oracle(oracle, engine, case):
    """Vorticity of the state after 20 steps: the channel (EB sphere, so
    one-sided differences next to solid cells and at the x faces) and a
    periodic Taylor-Green box; 1e-12 of the largest magnitude."""
    from oracle.pyoracle import build_oracle_level

    if case == "channel":
        cfg = configs.channel_sphere((48, 24, 24), radius=5.0, center=(16.0, 12.0, 12.0))
    else:
        cfg = configs.taylor_green(24, tau=1.0)
    O = build_oracle_level(cfg)
    O.evolve(20)
    ref = O.vorticity()
    E = configs.build_level(cfg, layout="compact")
    E.advance(20)
    n = int(np.prod(cfg["ncell"]))
    out = torch.empty(4 * n, dtype=torch.float64, device="cuda")
    E.vorticity(out)
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(ref.shape)
    E.close()
    err = np.abs(got - ref).max() / np.abs(ref[3]).max()
    print(f"{case}: max |vorticity - oracle| / max |omega| = {err:.2e} (bit-exact: {np.array_equal(got, ref)})")
    assert np.abs(ref[3]).max() > 0
    assert err <= 1e-12


def _tags_reference(f, periodic, criterion, value):
    if criterion == "greater":
        return f > value
    if criterion == "less":
        return f < value
    d = np.zeros_like(f)
    for ax, per in zip((2, 1, 0), periodic):  # f is [z, y, x]; periodic is (x, y, z)
        for s in (-1, 1):
            nb = np.roll(f, -s, axis=ax)
            diff = np.abs(nb - f)
            if not per:
                idx = [slice(None)] * 3
                idx[ax] = -1 if s == 1 else 0
                diff[tuple(idx)] = 0.0
            d = np.maximum(d, diff)
    return d > value


@pytest.mark.parametrize("criterion", ["greater", "less", "adjacent_difference"])
def test_tag_cells_matches_restatement(engine, criterion):
    cfg = configs.channel_sphere((40, 24, 16), radius=4.0, center=(12.0, 12.0, 8.0))
    E = configs.build_level(cfg, layout="compact")
    n = int(np.prod(cfg["ncell"]))
    rng = np.random.default_rng(11)
    f = rng.standard_normal((cfg["ncell"][2], cfg["ncell"][1], cfg["ncell"][0]))
    value = {"greater": 0.7, "less": -0.4, "adjacent_difference": 2.5}[criterion]
    fd = torch.tensor(f.ravel(), device="cuda")
    tags = torch.zeros(n, dtype=torch.uint8, device="cuda")
    E.tag_cells(fd, criterion, value, tags)
    torch.cuda.synchronize()
    got = tags.cpu().numpy().reshape(f.shape).astype(bool)
    E.close()
    ref = _tags_reference(f, cfg["periodic"], criterion, value)
    assert 0 < ref.sum() < ref.size
    assert np.array_equal(got, ref)
