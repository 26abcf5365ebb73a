"""Tag clustering for regrid (mrb_cluster_tags: tags grown by n_error_buf,
coarsened to blocking-factor blocks, Berger-Rigoutsos to grid_eff, chopped
to max_grid_size); host code, no GPU.  AMReX's own clustering is not part
of the reference, so these are the method's invariants, not AMReX's boxes:
every grown tag is covered, boxes are disjoint, aligned to the blocking
factor, inside the domain and no longer than max_grid_size, and separated
clusters are not merged."""
import numpy as np
import pytest


def _cluster(tags, **kw):
    from marbles_amd.amr import cluster_tags

    return cluster_tags(tags, **kw)


def _check(tags, boxes, n_error_buf, bf, mgs):
    nz, ny, nx = tags.shape
    fine = np.zeros((2 * nz, 2 * ny, 2 * nx), dtype=np.int32)
    for lo, hi in boxes:
        for d in range(3):
            assert lo[d] % bf == 0 and (hi[d] + 1) % bf == 0
            assert 0 <= lo[d] <= hi[d] < 2 * tags.shape[2 - d]
            assert hi[d] - lo[d] + 1 <= mgs
        fine[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] += 1
    assert fine.max() <= 1, "boxes overlap"
    grown = np.zeros_like(tags, dtype=bool)
    for k, j, i in zip(*np.nonzero(tags)):
        grown[max(0, k - n_error_buf):k + n_error_buf + 1, max(0, j - n_error_buf):j + n_error_buf + 1,
              max(0, i - n_error_buf):i + n_error_buf + 1] = True
    covered = fine.reshape(nz, 2, ny, 2, nx, 2).min(axis=(1, 3, 5)) > 0
    assert np.all(covered[grown]), "a grown tag is not covered"
    return fine.sum()


def test_two_blobs_are_split_at_the_hole(engine):
    tags = np.zeros((16, 16, 32), dtype=np.uint8)
    tags[4:8, 4:8, 2:6] = 1
    tags[8:12, 8:12, 24:28] = 1
    boxes = _cluster(tags, n_error_buf=1, blocking_factor=4, max_grid_size=64, grid_eff=0.7)
    vol = _check(tags, boxes, 1, 4, 64)
    assert len(boxes) >= 2
    # not one box spanning both blobs
    assert vol < 0.5 * (2 * 16) * (2 * 16) * (2 * 32)
    xs = sorted(lo[0] for lo, _ in boxes)
    assert xs[0] < 16 and xs[-1] >= 40


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_tags_invariants(engine, seed):
    rng = np.random.default_rng(seed)
    tags = (rng.random((16, 24, 20)) < 0.03).astype(np.uint8)
    for buf, bf, mgs, eff in ((0, 2, 8, 0.7), (1, 4, 16, 0.5), (2, 8, 16, 0.9)):
        boxes = _cluster(tags, n_error_buf=buf, blocking_factor=bf, max_grid_size=mgs, grid_eff=eff)
        _check(tags, boxes, buf, bf, mgs)


def test_empty_and_full(engine):
    assert _cluster(np.zeros((8, 8, 8), dtype=np.uint8)) == []
    boxes = _cluster(np.ones((8, 8, 8), dtype=np.uint8), blocking_factor=4, max_grid_size=8)
    assert _check(np.ones((8, 8, 8), dtype=np.uint8), boxes, 1, 4, 8) == 16 ** 3
    assert len(boxes) == 8


def test_bad_parameters_raise(engine):
    from marbles_amd._lib import MrbError

    with pytest.raises(MrbError):
        _cluster(np.ones((8, 8, 8), dtype=np.uint8), blocking_factor=3)
    with pytest.raises(MrbError):
        _cluster(np.ones((8, 8, 8), dtype=np.uint8), blocking_factor=4, max_grid_size=6)
