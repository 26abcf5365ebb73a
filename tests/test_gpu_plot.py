"""Plotfile fields and files (LBM::write_plot_file, LBM.cpp:1629-1682)
through the C ABI: the engine's plot variables equal the oracle's
(macrodata of the last collide's post-stream state, f, g, vorticity and
dQCorr of those macrodata, is_fluid), and the plotfile on disk follows
AMReX's layout (Header read line by line, Cell MultiFab by an independent
VisMF reader)."""
import os

import numpy as np
import pytest
from vismf import read_vismf

from marbles_amd import configs

pytestmark = pytest.mark.gpu

# LBM.cpp:14-82 (macrodata, f, g, derived, is_fluid names)
NAMES = (["rho", "vel_x", "vel_y", "vel_z", "vel_mag", "two_rho_e", "QCorrX", "QCorrY", "QCorrZ", "pxx", "pyy",
          "pzz", "pxy", "pxz", "pyz", "qx", "qy", "qz", "temperature"]
         + [f"f_{q:02d}" for q in range(27)] + [f"g_{q:02d}" for q in range(27)]
         + ["vort_x", "vort_y", "vort_z", "vort_mag", "dQCorrX", "dQCorrY", "dQCorrZ", "is_fluid", "eb_boundary"])


def _pair(oracle, engine, cfg):
    from oracle.pyoracle import build_oracle_level

    O = build_oracle_level(cfg)
    E = engine.Level(cfg["ncell"], cfg["periodic"], cfg["prob_lo"], cfg["prob_hi"], layout="compact")
    E.set_phys(cfg["nu"], cfg.get("alpha"))
    E.set_bc(cfg["bc_lo"], cfg["bc_hi"], cfg.get("velocity_bc"))
    E.flags_upload(O.valid("is_fluid"))
    E.upload("f", O.valid("f"))
    E.upload("g", O.valid("g"))
    return O, E


def _compare(got, ref):
    for k in range(ref.shape[0]):
        scale = max(np.abs(ref[k]).max(), 1e-300)
        err = np.abs(got[k] - ref[k]).max() / scale
        assert err <= 1e-12, (NAMES[k] if ref.shape[0] == len(NAMES) else k, err)


@pytest.mark.parametrize("case", ["channel", "pressure"])
def test_plot_fields_match_oracle(oracle, engine, case):
    """Before any step (macrodata of the initial state) and after 15 steps
    (the pull of the previous buffer), with an EB sphere, the velocity
    inlet and outflow, or a pressure outlet behind a walled inlet."""
    if case == "channel":
        cfg = configs.channel_sphere((40, 24, 24), radius=5.0, center=(14.0, 12.0, 12.0))
    else:
        cfg = configs.pressure_case("walled", (24, 16, 20))
    O, E = _pair(oracle, engine, cfg)
    _compare(E.plot_fields(streaming=True, derived=True), oracle.plot_vars(O, streaming=True, derived=True))
    E.advance(15)
    O.evolve(15)
    got = E.plot_fields(streaming=True, derived=True)
    ref = oracle.plot_vars(O, streaming=True, derived=True)
    assert np.abs(ref[19 + 54:19 + 54 + 4]).max() > 0  # a vorticity to compare
    _compare(got, ref)
    E.close()


def test_plotfile_layout(oracle, engine, tmp_path):
    cfg = configs.channel_sphere((32, 16, 12), radius=3.0, center=(10.0, 8.0, 6.0))
    O, E = _pair(oracle, engine, cfg)
    E.advance(4)
    path = str(tmp_path / "plt00004")
    E.write_plotfile(path, time=4.0, step=4, streaming=True, derived=True)
    head = open(os.path.join(path, "Header")).read().split("\n")
    nv = len(NAMES)
    assert head[0] == "HyperCLaw-V1.1" and int(head[1]) == nv and head[2:2 + nv] == NAMES
    rest = head[2 + nv:]
    assert rest[0] == "3" and float(rest[1]) == 4.0 and rest[2] == "0"
    assert [float(x) for x in rest[3].split()] == [0.0, 0.0, 0.0]
    assert [float(x) for x in rest[4].split()] == [32.0, 16.0, 12.0]
    assert rest[5] == "" and rest[6].strip() == "((0,0,0) (31,15,11) (0,0,0))" and rest[7].strip() == "4"
    assert [float(x) for x in rest[8].split()] == [1.0, 1.0, 1.0] and rest[9] == "0" and rest[10] == "0"
    assert rest[11].split() == ["0", "1", "4"] and rest[12] == "4"
    assert [[float(x) for x in r.split()] for r in rest[13:16]] == [[0.0, 32.0], [0.0, 16.0], [0.0, 12.0]]
    assert rest[16] == "Level_0/Cell"
    ncomp, ngrow, boxes, fabs = read_vismf(os.path.join(path, "Level_0", "Cell"))
    assert (ncomp, ngrow, boxes) == (nv, 0, [(0, 0, 0, 31, 15, 11)])
    np.testing.assert_array_equal(fabs[0], E.plot_fields(streaming=True, derived=True))
    E.close()


def _fab_pair(oracle, engine, cfg):
    from oracle.pyoracle import build_oracle_level

    O = build_oracle_level(cfg)
    E = engine.Level(cfg["ncell"], cfg["periodic"], cfg["prob_lo"], cfg["prob_hi"], boxes=O.boxes, ng=3,
                     layout="fab")
    E.set_phys(cfg["nu"], cfg.get("alpha"))
    E.set_bc(cfg["bc_lo"], cfg["bc_hi"], cfg.get("velocity_bc"))
    E.flags_upload(O.flat("is_fluid"))
    E.upload("f", O.flat("f"))
    E.upload("g", O.flat("g"))
    return O, E


@pytest.mark.parametrize("path", ["stages", "batched"])
def test_fab_plot_fields_match_oracle(oracle, engine, path):
    """Multi-box fab level (boxes of 12): the initial state's macrodata, then
    after 9 steps through the stage calls (the collide's own macrodata) or
    the batched step (the pull of the state before the step), with vorticity
    across box edges."""
    cfg = configs.channel_sphere((36, 24, 24), radius=5.0, center=(13.0, 12.0, 12.0), max_box=12)
    O, E = _fab_pair(oracle, engine, cfg)
    assert len(O.boxes) > 1
    _compare(E.plot_fields(streaming=True, derived=True), oracle.plot_vars(O, streaming=True, derived=True))
    n = 9
    if path == "stages":
        for _ in range(n):
            for lat in (0, 1):
                E.fillpatch(lat)
                E.stream(lat)
            E.collide()
    else:
        E.advance(n)
    O.evolve(n)
    ref = oracle.plot_vars(O, streaming=True, derived=True)
    assert np.abs(ref[19 + 54:19 + 54 + 4]).max() > 0
    _compare(E.plot_fields(streaming=True, derived=True), ref)
    E.close()


def test_fab_plotfile_boxes(oracle, engine, tmp_path):
    cfg = configs.channel_sphere((24, 16, 16), radius=3.0, center=(8.0, 8.0, 8.0), max_box=8)
    O, E = _fab_pair(oracle, engine, cfg)
    E.advance(3)
    path = str(tmp_path / "plt00003")
    E.write_plotfile(path, time=3.0, step=3, derived=True)
    head = open(os.path.join(path, "Header")).read().split("\n")
    nv = 19 + 7 + 2
    rest = head[2 + nv:]
    nbox = len(E.boxes)
    assert rest[11].split() == ["0", str(nbox), "3"]
    for b, (lo, hi) in enumerate(E.boxes):
        rows = [[float(x) for x in r.split()] for r in rest[13 + 3 * b:16 + 3 * b]]
        assert rows == [[float(lo[d]), float(hi[d] + 1)] for d in range(3)]
    assert rest[13 + 3 * nbox] == "Level_0/Cell"
    ncomp, ngrow, boxes, fabs = read_vismf(os.path.join(path, "Level_0", "Cell"))
    assert (ncomp, ngrow) == (nv, 0) and boxes == [tuple(lo) + tuple(hi) for lo, hi in E.boxes]
    P = E.plot_fields(derived=True)
    for (lo, hi), fab in zip(E.boxes, fabs):
        np.testing.assert_array_equal(fab, P[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1])
    E.close()


@pytest.mark.parametrize("case", ["uneven_32", "inlet_pressure"])
def test_two_level_plot_fields_and_plotfile(oracle, engine, tmp_path, case):
    """Both levels of a hierarchy against the oracle's: the initial state,
    then after 4 coarse steps (coarse macrodata with the coalesced fine
    values of its last pass, fine macrodata with the ring-1 ghosts over
    uncovered domain for the gradients at the coarse/fine edge); and the
    two-level plotfile (Header, Level_0/Cell, Level_1/Cell)."""
    from amr_cases import oracle_hierarchy
    from test_gpu_amr import CASES

    cfg = CASES[case]()
    C, F = oracle_hierarchy(cfg)
    H = engine.Hierarchy(cfg)
    for lvl, O in ((H.crse, C), (H.fine, F)):
        _compare(lvl.plot_fields(derived=True), oracle.plot_vars(O, derived=True))
    n = 4
    oracle.amr_evolve(C, F, n)
    H.advance(n)
    for lvl, O in ((H.crse, C), (H.fine, F)):
        ref = oracle.plot_vars(O, streaming=True, derived=True)
        _compare(lvl.plot_fields(streaming=True, derived=True), ref)
    path = str(tmp_path / "plt00004")
    H.write_plotfile(path, time=4.0, steps=(4, 8), derived=True)
    head = open(os.path.join(path, "Header")).read().split("\n")
    nv = 19 + 7 + 2
    rest = head[2 + nv:]
    nc, nf = cfg["ncell"], tuple(2 * x for x in cfg["ncell"])
    assert rest[2] == "1" and rest[5].split() == ["2"]
    assert rest[6].strip() == (f"((0,0,0) ({nc[0] - 1},{nc[1] - 1},{nc[2] - 1}) (0,0,0)) "
                               f"((0,0,0) ({nf[0] - 1},{nf[1] - 1},{nf[2] - 1}) (0,0,0))")
    assert rest[7].split() == ["4", "8"]
    assert [float(x) for x in rest[8].split()] == [1.0, 1.0, 1.0]
    assert [float(x) for x in rest[9].split()] == [0.5, 0.5, 0.5]
    assert rest[10] == "0" and rest[11] == "0"
    nb0, nb1 = len(H.crse.boxes), len(H.fine.boxes)
    assert rest[12].split() == ["0", str(nb0), "4"] and rest[13] == "4"
    j = 14 + 3 * nb0
    assert rest[j] == "Level_0/Cell" and rest[j + 1].split() == ["1", str(nb1), "4"] and rest[j + 2] == "8"
    assert rest[j + 3 + 3 * nb1] == "Level_1/Cell"
    P = H.fine.plot_fields(derived=True)
    ncomp, ngrow, boxes, fabs = read_vismf(os.path.join(path, "Level_1", "Cell"))
    assert (ncomp, ngrow, len(fabs)) == (nv, 0, nb1)
    for (lo, hi), fab in zip(H.fine.boxes, fabs):
        np.testing.assert_array_equal(fab, P[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1])
    H.close()


@pytest.mark.parametrize("storage,tol", [("f32s", 1e-6), ("f16s", 5e-3)])
def test_reduced_storage_plot_fields_track_f64(engine, storage, tol):
    """Plot macrodata (the pull of the pre-step buffer through the reduced
    storage) of a channel with an inlet, against the FP64 run: density,
    velocity and temperature within the storage's tolerance of their range;
    is_fluid identical."""
    cfg = configs.channel_sphere((40, 24, 24), radius=5.0, center=(14.0, 12.0, 12.0))
    P = {}
    for st in ("f64", storage):
        E = configs.build_level(cfg, layout="compact", storage=st)
        E.advance(30)
        P[st] = E.plot_fields(derived=True)
        E.close()
    a, b = P[storage], P["f64"]
    fluid = b[26] == 1
    for k in (0, 1, 2, 3, 18):
        err = np.abs(a[k] - b[k])[fluid].max() / max(np.abs(b[k][fluid]).max(), 1e-300)
        assert err <= tol, (storage, k, err)
    assert np.array_equal(a[26], b[26])
