"""Pins the FP64 oracle of the Kirchhoff path (CPU only): the reference's
fast-sweeping eikonal solver (src/fd_cmex/eikonal2d_mex.cpp:27-267) and its
Kirchhoff migration (src/migrate.m, src/shot2RecTime.m).

Eikonal -- checks that do not restate the update:
  - exact traveltime k*s*dx along the source's row and column (homogeneous);
  - the edge rule (a row or column at the grid edge takes its one inner
    neighbour) is a mirror: a corner source on a quarter grid equals, bit for
    bit, the centre source on the grid mirrored about it, while the far edges
    are still out of reach;
  - mirror symmetry of a model symmetric about the source column, bit for bit;
  - the iteration decreases monotonically to a fixed point;
  - first-order convergence to the exact traveltime r/v under refinement;
  - the spacing is truncated to an integer (eikonal2d_mex.cpp:35).
Migration:
  - a single-spike record images exactly onto the pixels whose source and
    receiver traveltime samples (rounded, each +1) add up to the spike sample;
  - linearity in the record; brute force on tiny inputs; samples outside the
    record are an error, as MATLAB's indexing is.
"""
import math

import numpy as np
import pytest

import oracle


def test_eikonal_axis_is_exact():
    v = np.full((41, 57), 2000.0)
    sz, sx, dx = 21, 29, 10.0
    T = oracle.eikonal2d(v, dx, sz, sx, 30)
    h = dx / 2000.0
    k = np.arange(1, 20)
    np.testing.assert_allclose(T[sz - 1, sx - 1 + k], k * h, rtol=1e-13)
    np.testing.assert_allclose(T[sz - 1, sx - 1 - k], k * h, rtol=1e-13)
    np.testing.assert_allclose(T[sz - 1 + k, sx - 1], k * h, rtol=1e-13)
    assert T[sz - 1, sx - 1] == 0.0


def test_eikonal_edge_rule_is_a_mirror():
    v = np.random.default_rng(0).uniform(1500, 3000, (30, 40))
    full = np.block([[v[::-1, ::-1][:-1, :-1], v[::-1, :][:-1, :]],
                     [v[:, ::-1][:, :-1], v]])
    for iters in (1, 3, 6):  # 4 * iters passes stay short of the far edges
        Tq = oracle.eikonal2d(v, 10.0, 1, 1, iters)
        Tf = oracle.eikonal2d(full, 10.0, 30, 40, iters)
        assert np.array_equal(Tq, Tf[29:, 39:]), iters


def test_eikonal_mirror_symmetry():
    rng = np.random.default_rng(1)
    half = rng.uniform(1500, 4000, (35, 20))
    v = np.concatenate([half[:, ::-1], half[:, :1], half], axis=1)   # 41 columns, centre 21
    T = oracle.eikonal2d(v, 10.0, 1, 21, 25)
    assert np.array_equal(T[:, :20], T[:, 21:][:, ::-1])


def test_eikonal_decreases_to_a_fixed_point():
    v = np.random.default_rng(2).uniform(1500, 4500, (30, 45))
    prev = None
    for iters in (2, 4, 8, 16, 24):
        T = oracle.eikonal2d(v, 10.0, 5, 12, iters)
        if prev is not None:
            assert np.all(T <= prev)
        prev = T
    assert np.array_equal(oracle.eikonal2d(v, 10.0, 5, 12, 25), prev)


def test_eikonal_converges_to_the_exact_traveltime():
    """Homogeneous medium, 400 m x 600 m: the error at fixed physical points
    falls with the spacing (first order)."""
    errs = []
    for dx in (20, 10, 5):
        nz, nx = 400 // dx + 1, 600 // dx + 1
        v = np.full((nz, nx), 2500.0)
        T = oracle.eikonal2d(v, dx, 1, nx // 2 + 1, (nz + nx) // 2)
        pts = [(300, 100), (200, 250), (400, 300)]  # (z, x) metres from the source
        e = max(abs(T[z // dx, nx // 2 + xo // dx] - math.hypot(z, xo) / 2500.0) /
                (math.hypot(z, xo) / 2500.0) for z, xo in pts)
        errs.append(e)
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 0.03


def test_eikonal_truncates_the_spacing():
    v = np.random.default_rng(3).uniform(1500, 3000, (20, 25))
    assert np.array_equal(oracle.eikonal2d(v, 10.9, 3, 4, 10), oracle.eikonal2d(v, 10.0, 3, 4, 10))


def _tables(nz, nx, grid, v=2000.0, dx=10.0):
    """Homogeneous traveltime tables for surface positions `grid` (1-based)."""
    zz = np.arange(nz)[:, None] * dx
    xx = np.arange(nx)[None, :] * dx
    return np.stack([np.hypot(zz, xx - (g - 1) * dx) / v for g in grid], axis=2)


def test_migrate_single_spike_images_onto_its_isochron():
    nz, nx, nt, dt = 12, 20, 150, 2e-3
    grid = list(range(1, nx + 1))
    tt = _tables(nz, nx, grid)
    xrec = [3, 9, 15]
    ixs, it0, r0 = 7, 17, 1        # spike at sample 17 (1-based) of receiver xrec[1]
    shot = np.zeros((nt, nx))
    shot[it0 - 1, xrec[r0] - 1] = 1.0
    m = oracle.kirchhoff_migrate(tt, xrec, grid, shot, dt, nz, ixs)
    ks = np.round(tt[:, :, ixs - 1] / dt) + 1
    kr = np.round(tt[:, :, xrec[r0] - 1] / dt) + 1
    assert np.array_equal(m, (ks + kr == it0).astype(float))
    assert m.sum() > 0


def test_migrate_linearity_and_brute_force():
    rng = np.random.default_rng(4)
    nz, nx, nt, dt = 6, 9, 40, 3e-3
    grid = [1, 3, 4, 6, 8, 9]
    tt = rng.uniform(0, 0.05, (nz, nx, len(grid)))
    xrec = [3, 8, 9, 4]
    a, b = rng.standard_normal((nt, nx)), rng.standard_normal((nt, nx))
    ma = oracle.kirchhoff_migrate(tt, xrec, grid, a, dt, nz, 2)
    mb = oracle.kirchhoff_migrate(tt, xrec, grid, b, dt, nz, 2)
    mab = oracle.kirchhoff_migrate(tt, xrec, grid, 2 * a - 0.5 * b, dt, nz, 2)
    np.testing.assert_allclose(mab, 2 * ma - 0.5 * mb, rtol=1e-12, atol=1e-12)
    ref = np.zeros((nz, nx))
    for xr in xrec:
        r = grid.index(xr)
        for i in range(nz):
            for j in range(nx):
                it = int(np.floor(tt[i, j, 1] / dt + 0.5)) + 1 + int(np.floor(tt[i, j, r] / dt + 0.5)) + 1
                ref[i, j] += a[it - 1, xr - 1]
    assert np.array_equal(ma, ref)


def test_migrate_outside_the_record_is_an_error():
    nz, nx, dt = 5, 6, 1e-3
    grid = list(range(1, nx + 1))
    tt = _tables(nz, nx, grid)
    with pytest.raises(ValueError):
        oracle.kirchhoff_migrate(tt, [2, 4], grid, np.ones((10, nx)), dt, nz, 1)
