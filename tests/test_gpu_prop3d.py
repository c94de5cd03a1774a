"""GPU parity of the 3-D CPML propagator (through the C ABI) against the FP64 oracle.

Tolerance: relative L2 error <= 1e-5 of the FP32 GPU seismogram and snapshots
against the FP64 oracle on the same seeded inputs.
"""
import os

import numpy as np
import pytest

import oracle
import sssi_amd
from sssi_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


@pytest.fixture
def kernel_env():
    saved = {k: os.environ.get(k) for k in ("SSSI_TY3", "SSSI_ZC3", "SSSI_RING3", "SSSI_PD3")}
    yield
    for k, v in saved.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


@pytest.mark.parametrize("N,ty,zc", [(2, 8, 64), (4, 8, 8), (4, 4, 16), (5, 16, 8), (1, 8, 12),
                                     (8, 4, 64)])
def test_fwd_cpml_3d_matches_oracle(N, ty, zc, kernel_env):
    os.environ["SSSI_TY3"] = str(ty)
    os.environ["SSSI_ZC3"] = str(zc)
    nz, nx, ny, nB, nt = 36, 80, 30, 6, 120
    v = synth.cake_3d_dipping(nz, nx, ny, 4, 1500.0, 3000.0, seed=7 + N)
    dx = dy = dz = 10.0
    dt = 0.4 * dx / 3000.0 / np.sqrt(3) / sum(abs(c) for c in synth.dCoef(N))
    w = synth.ricker(25, nt, dt)
    iz, ix, iy = [0, 20], [40, 10], [15, 22]
    ww = np.stack([w, -0.7 * w], axis=1)
    d_ref, s_ref = oracle.fwd_cpml_3d(v, nt, N, nB, dz, dx, dy, dt, src_iz=iz, src_ix=ix,
                                      src_iy=iy, wavelet=ww, want_snapshot=True, snap_stride=9)
    d_gpu, s_gpu = sssi_amd.fwd_cpml_3d_points(v, nt, N, nB, dz, dx, dy, dt, iz, ix, iy, ww,
                                               want_snapshot=True, snap_stride=9)
    assert rel_l2(d_gpu, d_ref) < TOL, rel_l2(d_gpu, d_ref)
    assert rel_l2(s_gpu, s_ref) < TOL, rel_l2(s_gpu, s_ref)


def test_fwdTimeCpmlFor3dAw_dense_signature():
    nz, nx, ny, nB, nt, N = 20, 24, 22, 4, 40, 3
    v = synth.cake_3d_dipping(nz, nx, ny, 3, 1500.0, 2500.0, seed=3)
    dt = 1e-3
    src = np.zeros((nz, nx, ny, nt))
    src[0, 12, 11, :] = synth.ricker(30, nt, dt)
    d_ref, s_ref = oracle.fwd_cpml_3d(v, nt, N, nB, 10.0, 10.0, 10.0, dt, source=src,
                                      want_snapshot=True)
    d_gpu, s_gpu = sssi_amd.fwdTimeCpmlFor3dAw(v, src, N, nB, 10.0, 10.0, 10.0, dt)
    assert d_gpu.shape == (nx, ny, nt) and s_gpu.shape == (nz, nx, ny, nt)
    assert rel_l2(d_gpu, d_ref) < TOL
    assert rel_l2(s_gpu, s_ref) < TOL


def test_dense_plane_and_volume_sources_beyond_4096_points():
    """A dense source with a whole plane (5040 cells, the first step) plus a
    buried block, and repeated cells in a point list: each cell's sources are
    summed in the caller's order, as the oracle sums them."""
    nz, nx, ny, nB, nt, N = 24, 72, 70, 5, 30, 2
    v = synth.cake_3d_dipping(nz, nx, ny, 3, 1500.0, 2500.0, seed=8)
    dt = 1e-3
    w = synth.ricker(30, nt, dt)
    src = np.zeros((nz, nx, ny, nt))
    src[2, :, :, :] = w[None, None, :] * np.linspace(0.5, 1.5, nx)[:, None, None]
    src[10:13, 30:34, 40:45, :] = -0.3 * w[None, None, None, :]
    d_ref, s_ref = oracle.fwd_cpml_3d(v, nt, N, nB, 10.0, 10.0, 10.0, dt, source=src,
                                      want_snapshot=True, snap_stride=7)
    d_gpu, s_gpu = sssi_amd.fwd_cpml_3d_points(
        v, nt, N, nB, 10.0, 10.0, 10.0, dt, *np.nonzero(np.any(src != 0, axis=3)),
        np.ascontiguousarray(src[np.any(src != 0, axis=3)].T), want_snapshot=True, snap_stride=7)
    assert rel_l2(d_gpu, d_ref) < TOL, rel_l2(d_gpu, d_ref)
    assert rel_l2(s_gpu, s_ref) < TOL, rel_l2(s_gpu, s_ref)
    d_dense, _ = sssi_amd.fwdTimeCpmlFor3dAw(v, src, N, nB, 10.0, 10.0, 10.0, dt)
    assert np.array_equal(d_dense, d_gpu)
    # repeated cells, interleaved with others
    iz, ix, iy = [3, 0, 3, 7, 3], [40, 10, 40, 5, 40], [20, 22, 20, 9, 20]
    ww = np.stack([w, -w, 0.5 * w, w, 0.25 * w], axis=1)
    d_ref, _ = oracle.fwd_cpml_3d(v, nt, N, nB, 10.0, 10.0, 10.0, dt, src_iz=iz, src_ix=ix,
                                  src_iy=iy, wavelet=ww)
    d_gpu, _ = sssi_amd.fwd_cpml_3d_points(v, nt, N, nB, 10.0, 10.0, 10.0, dt, iz, ix, iy, ww)
    assert rel_l2(d_gpu, d_ref) < TOL, rel_l2(d_gpu, d_ref)


@pytest.mark.parametrize("nz,nx,ny,nB,N,ty,zc", [
    (24, 70, 20, 0, 3, 8, 16),   # no damping strips: every block takes the interior sweep
    (10, 40, 24, 8, 4, 8, 8),    # z split at depth 0: every chunk reaches the bottom strip
    (40, 72, 26, 6, 4, 8, 8),    # chunks of 7 and 1 planes at the split and the bottom, 4-plane ring
    (40, 72, 40, 6, 4, 16, 12),  # 16-row tiles with the 4-plane ring
])
def test_fwd_cpml_3d_chunking_edge_cases(nz, nx, ny, nB, N, ty, zc, kernel_env):
    os.environ["SSSI_TY3"] = str(ty)
    os.environ["SSSI_ZC3"] = str(zc)
    nt = 90
    v = synth.cake_3d_dipping(nz, nx, ny, 3, 1500.0, 3000.0, seed=11 + nz)
    dx = dy = dz = 10.0
    dt = 0.4 * dx / 3000.0 / np.sqrt(3) / sum(abs(c) for c in synth.dCoef(N))
    w = synth.ricker(25, nt, dt)
    iz, ix, iy = [0, nz // 2], [nx // 2, 5], [ny // 2, ny - 4]
    ww = np.stack([w, 0.5 * w], axis=1)
    d_ref, s_ref = oracle.fwd_cpml_3d(v, nt, N, nB, dz, dx, dy, dt, src_iz=iz, src_ix=ix,
                                      src_iy=iy, wavelet=ww, want_snapshot=True, snap_stride=11)
    d_gpu, s_gpu = sssi_amd.fwd_cpml_3d_points(v, nt, N, nB, dz, dx, dy, dt, iz, ix, iy, ww,
                                               want_snapshot=True, snap_stride=11)
    assert rel_l2(d_gpu, d_ref) < TOL, rel_l2(d_gpu, d_ref)
    assert rel_l2(s_gpu, s_ref) < TOL, rel_l2(s_gpu, s_ref)


@pytest.mark.parametrize("N,ty,zc,pd", [(4, 8, 8, 0), (4, 8, 12, 2), (3, 16, 12, 0), (5, 16, 8, 0),
                                        (2, 4, 16, 0)])
def test_plane_ring_equals_two_tile_buffer(N, ty, zc, pd, kernel_env):
    """The opt-in LDS plane ring (halo loaded with the z window's front and
    parked in LDS for L planes) computes the same values as the default
    2-tile buffer, bit for bit, over strip and interior blocks, chunk seams
    and the prefetch ring's depths."""
    os.environ["SSSI_TY3"] = str(ty)
    os.environ["SSSI_ZC3"] = str(zc)
    if pd:
        os.environ["SSSI_PD3"] = str(pd)
    nz, nx, ny, nB, nt = 40, 90, 44, 6, 80
    v = synth.cake_3d_dipping(nz, nx, ny, 4, 1500.0, 3000.0, seed=21 + N)
    dx = dy = dz = 10.0
    dt = 0.4 * dx / 3000.0 / np.sqrt(3) / sum(abs(c) for c in synth.dCoef(N))
    w = synth.ricker(25, nt, dt)
    iz, ix, iy = [0, 25], [45, 70], [20, 9]
    ww = np.stack([w, -0.6 * w], axis=1)
    out = {}
    for ring in ("0", "1"):
        os.environ["SSSI_RING3"] = ring
        print(f"N={N} ty={ty} zc={zc} pd={pd} ring={ring}", flush=True)
        out[ring] = sssi_amd.fwd_cpml_3d_points(v, nt, N, nB, dz, dx, dy, dt, iz, ix, iy, ww,
                                                want_snapshot=True, snap_stride=13)
    assert np.array_equal(out["0"][0], out["1"][0])
    assert np.array_equal(out["0"][1], out["1"][1])
    d_ref = oracle.fwd_cpml_3d(v, nt, N, nB, dz, dx, dy, dt, src_iz=iz, src_ix=ix, src_iy=iy,
                               wavelet=ww)[0]
    assert rel_l2(out["1"][0], d_ref) < TOL, rel_l2(out["1"][0], d_ref)


@pytest.mark.parametrize("band", ["0", "1"])
def test_strip_block_orders_equal_default(band, kernel_env):
    """SSSI_STRIPBAND3 = 0 (natural order) and 1 (one banded list) only change
    where the strip blocks run: the results equal the default order's (banded
    per kind of strip) bit for bit."""
    os.environ["SSSI_TY3"] = "8"
    os.environ["SSSI_ZC3"] = "12"
    nz, nx, ny, nB, N, nt = 40, 150, 60, 6, 4, 70
    v = synth.cake_3d_dipping(nz, nx, ny, 3, 1500.0, 3000.0, seed=31)
    dx = dy = dz = 10.0
    dt = 0.4 * dx / 3000.0 / np.sqrt(3) / sum(abs(c) for c in synth.dCoef(N))
    w = synth.ricker(25, nt, dt)
    iz, ix, iy = [0, 30], [75, 10], [30, 50]
    ww = np.stack([w, 0.4 * w], axis=1)
    saved = os.environ.get("SSSI_STRIPBAND3")
    try:
        os.environ.pop("SSSI_STRIPBAND3", None)
        base = sssi_amd.fwd_cpml_3d_points(v, nt, N, nB, dz, dx, dy, dt, iz, ix, iy, ww,
                                           want_snapshot=True, snap_stride=9)
        os.environ["SSSI_STRIPBAND3"] = band
        alt = sssi_amd.fwd_cpml_3d_points(v, nt, N, nB, dz, dx, dy, dt, iz, ix, iy, ww,
                                          want_snapshot=True, snap_stride=9)
    finally:
        if saved is None:
            os.environ.pop("SSSI_STRIPBAND3", None)
        else:
            os.environ["SSSI_STRIPBAND3"] = saved
    assert np.array_equal(base[0], alt[0]) and np.array_equal(base[1], alt[1])
