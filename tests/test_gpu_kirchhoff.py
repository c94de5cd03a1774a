"""Kirchhoff path on the GPU against the oracle, bit for bit:

- fast-sweeping eikonal traveltimes (src/fd_cmex/eikonal2d_mex.cpp:27-267),
  one source and a batch of sources, no passes (iMax = 0), a non-integer dx
  (the reference reads dx as an int, eikonal2d_mex.cpp:35);
- migration of one shot (src/migrate.m:1-26, src/shot2RecTime.m:1-6) for
  receivers that are a subset of xShotAndRecGrid;
- the resident handle: its traveltime tables are the eikonal's with sources on
  row 1 (the app's [1 xs], applications/kirchhoff_2d_acoustic/
  mainKirchhoffCpmlFor2dAw.m:246), and its stack is the app's
  Stacked = Stacked + M over shots in order (:278-291), per-shot images too;
- an index past the record, and a receiver missing from xShotAndRecGrid, are
  errors (MATLAB's indexing and ismember errors)."""
import numpy as np
import pytest

import oracle
import sssi_amd
from sssi_amd import synth

pytestmark = pytest.mark.gpu


def _model(nz=40, nx=60, seed=3):
    V = synth.layered_2d(nz, nx, 4, 1500.0, 3000.0, seed=seed)
    return synth.smooth_copy(V, 2)


def _exact(a, b):
    assert a.shape == b.shape
    assert np.array_equal(a, b), f"max |diff| {np.max(np.abs(a - b))}"


@pytest.mark.parametrize("dx, sz, sx, iMax", [(10.0, 1, 30, 40), (10.0, 17, 1, 5),
                                              (7.9, 40, 60, 12), (10.0, 5, 5, 0)])
def test_eikonal_matches_oracle_bit_for_bit(dx, sz, sx, iMax):
    V = _model()
    print(f"eikonal dx={dx} src=({sz},{sx}) iMax={iMax}", flush=True)
    _exact(sssi_amd.eikonal2d(V, dx, sz, sx, iMax), oracle.eikonal2d(V, dx, sz, sx, iMax))


@pytest.mark.parametrize("layout", ["plain", "4x32", "4x64", "8x48", "8x64"])
def test_eikonal_pass_layouts_match_oracle(layout, monkeypatch):
    """Every pass kernel (one pass per launch, or H passes per launch on
    overlapping tiles) gives the oracle's tables bit for bit: tiles in both
    directions, pass counts that are and are not multiples of H, sources on
    edges, corners and tile seams."""
    monkeypatch.setenv("SSSI_EIK", layout)
    V = _model(131, 157, seed=13)
    sz = [1, 131, 48, 49, 97, 1]
    sx = [1, 157, 48, 49, 2, 90]
    for iMax in (9, 10):
        print(f"layout {layout} iMax={iMax}", flush=True)
        T = sssi_amd.eikonal2d_batch(V, 10.0, sz, sx, iMax)
        for k in range(len(sz)):
            _exact(T[:, :, k], oracle.eikonal2d(V, 10.0, sz[k], sx[k], iMax))


def test_eikonal_batch_is_each_source_alone():
    V = _model(33, 47, seed=5)
    sz = [1, 1, 33, 12, 20]
    sx = [1, 47, 47, 23, 2]
    T = sssi_amd.eikonal2d_batch(V, 10.0, sz, sx, 20)
    assert T.shape == (33, 47, 5)
    for k in range(5):
        _exact(T[:, :, k], oracle.eikonal2d(V, 10.0, sz[k], sx[k], 20))


def _geometry(nx):
    xShot = np.arange(1, nx + 1, 3)
    xRec = np.arange(2, nx + 1, 2)
    return xShot, xRec, np.union1d(xShot, xRec)


def _tables(V, dx, xsr, iMax):
    return np.stack([oracle.eikonal2d(V, dx, 1, int(x), iMax) for x in xsr], axis=2)


def test_migrate_one_shot_matches_oracle():
    nz, nx, nt, dt = 30, 40, 900, 1e-3
    V = _model(nz, nx, seed=7)
    _, xRec, xsr = _geometry(nx)
    tt = _tables(V, 10.0, xsr, 30)
    shot = np.random.default_rng(1).standard_normal((nt, nx))
    for ixs in (1, len(xsr) // 2, len(xsr)):
        print(f"migrate ixs={ixs}", flush=True)
        _exact(sssi_amd.migrate(tt, xRec, xsr, shot, dt, nz, ixs),
               oracle.kirchhoff_migrate(tt, xRec, xsr, shot, dt, nz, ixs))


def test_handle_tables_and_stack_match_app_loop():
    nz, nx, nt, dt, dx, iMax = 36, 48, 1000, 1e-3, 10.0, 30
    V = _model(nz, nx, seed=9)
    xShot, xRec, xsr = _geometry(nx)
    k = sssi_amd.Kirchhoff2D(V, dx, xsr, iMax, dt)
    tt = k.travel_time()
    _exact(tt, _tables(V, dx, xsr, iMax))
    rng = np.random.default_rng(2)
    shots = rng.standard_normal((nt, nx, len(xShot)))
    ixs = np.searchsorted(xsr, xShot) + 1  # table of each shot position (1-based)
    stack, images = k.migrate_shots(xRec, shots, ixs, images=True)
    ref = np.zeros((nz, nx))
    for s in range(len(xShot)):
        M = oracle.kirchhoff_migrate(tt, xRec, xsr, shots[:, :, s], dt, nz, int(ixs[s]))
        _exact(images[:, :, s], M)
        ref = ref + M
    _exact(stack, ref)
    # the stack alone is the same call without images
    _exact(k.migrate_shots(xRec, shots, ixs), ref)
    # resident records, a shot count that is not a multiple of the kernel's shot batch
    k.load_shots(shots[:, :, :7])
    ref7 = np.zeros((nz, nx))
    for s in range(7):
        ref7 = ref7 + images[:, :, s]
    st7, im7 = k.migrate_loaded(xRec, ixs[:7], images=True)
    _exact(im7, images[:, :, :7])
    _exact(st7, ref7)
    _exact(k.migrate_loaded(xRec[::3], ixs[:7]),
           sum(oracle.kirchhoff_migrate(tt, xRec[::3], xsr, shots[:, :, s], dt, nz, int(ixs[s]))
               for s in range(7)))
    k.close()


def test_index_past_record_and_unknown_receiver_are_errors():
    nz, nx, dt = 20, 24, 1e-3
    V = _model(nz, nx, seed=11)
    xsr = np.arange(1, nx + 1)
    tt = _tables(V, 10.0, xsr, 20)
    short = np.ones((50, nx))  # traveltimes reach past 50 samples
    with pytest.raises(sssi_amd.SssiError) as e:
        sssi_amd.migrate(tt, xsr, xsr, short, dt, nz, 1)
    assert e.value.code == -1
    with pytest.raises(ValueError):
        oracle.kirchhoff_migrate(tt, xsr, xsr, short, dt, nz, 1)
    with pytest.raises(sssi_amd.SssiError) as e:
        sssi_amd.migrate(tt, [nx + 5], xsr, np.ones((900, nx)), dt, nz, 1)
    assert e.value.code == -1
