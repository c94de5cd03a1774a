"""GPU parity of the 2-D propagators (through the C ABI) against the FP64 oracle.

Tolerance: relative L2 error <= 1e-5 of the FP32 GPU result against the FP64
oracle on the same seeded inputs (seismograms, snapshots, reverse-time
models), as required of the FP32 device path.
"""
import numpy as np
import pytest

import oracle
import sssi_amd
from sssi_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel_l2(a, b):
    b = np.asarray(b)
    return float(np.linalg.norm(np.asarray(a) - b) / np.linalg.norm(b))


def _case(nz, nx, seed, vmin=1500.0, vmax=3500.0):
    v = synth.layered_2d(nz, nx, 5, vmin, vmax, seed)
    v = v + 200.0 * synth.gaussian_perturbation((nz, nx), 1.0, 3, seed + 1)
    return np.clip(v, vmin, vmax)


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6, 7, 8])
def test_fwd_cpml_2d_points_matches_oracle(N):
    nz, nx, nB, nt = 70, 150, 12, 300
    v = _case(nz, nx, 10 + N)
    dx = dz = 10.0
    dt = 0.5 * dx / 3500.0 / np.sqrt(2) / sum(abs(c) for c in synth.dCoef(N))
    w = synth.ricker(20, nt, dt)
    iz, ix = [0, 30], [nx // 2, 40]
    ww = np.stack([w, 0.5 * w], axis=1)
    d_ref, s_ref = oracle.fwd_cpml_2d(v, nt, N, nB, dz, dx, dt, src_iz=iz, src_ix=ix, wavelet=ww,
                                      want_snapshot=True, snap_stride=7)
    d_gpu, s_gpu = sssi_amd.fwd_cpml_2d_points(v, nt, N, nB, dz, dx, dt, iz, ix, ww,
                                               want_snapshot=True, snap_stride=7)
    assert rel_l2(d_gpu, d_ref) < TOL, rel_l2(d_gpu, d_ref)
    assert rel_l2(s_gpu, s_ref) < TOL, rel_l2(s_gpu, s_ref)


@pytest.mark.parametrize("N", [2, 5])
def test_fwdTimeCpmlFor2dAw_dense_signature(N):
    nz, nx, nB, nt = 40, 64, 10, 120
    v = _case(nz, nx, 3)
    dt = 1e-3
    src = np.zeros((nz, nx, nt))
    src[0, nx // 2, :] = synth.ricker(25, nt, dt)
    src[5, 10, :] = 0.3 * synth.ricker(30, nt, dt)
    d_ref, s_ref = oracle.fwd_cpml_2d(v, nt, N, nB, 10.0, 10.0, dt, source=src, want_snapshot=True)
    d_gpu, s_gpu = sssi_amd.fwdTimeCpmlFor2dAw(v, src, N, nB, 10.0, 10.0, dt)
    assert d_gpu.shape == (nx, nt) and s_gpu.shape == (nz, nx, nt)
    assert rel_l2(d_gpu, d_ref) < TOL
    assert rel_l2(s_gpu, s_ref) < TOL
    d1, s1 = sssi_amd.fwdTimeCpmlFor2dAw(v, src, N, nB, 10.0, 10.0, dt, nargout=1)
    assert s1 is None and np.array_equal(d1, d_gpu)


@pytest.mark.parametrize("N", [2, 4])
def test_rvs_cpml_2d_matches_oracle(N):
    nz, nx, nB, nt = 60, 90, 15, 250
    v = _case(nz, nx, 21)
    dt = 1e-3
    rng = np.random.default_rng(5)
    data = np.zeros((nx, nt))
    data[:, 50:200] = rng.standard_normal((nx, 150))
    m_ref, s_ref = oracle.rvs_cpml_2d(v, data, N, nB, 10.0, 10.0, dt, want_snapshot=True)
    m_gpu, s_gpu = sssi_amd.rvsTimeCpmlFor2dAw(v, data, N, nB, 10.0, 10.0, dt)
    l = 2 * N - 1
    assert m_gpu.shape == (nz + 2 * l, nx + 2 * l)
    assert rel_l2(m_gpu, m_ref) < TOL
    assert rel_l2(s_gpu, s_ref) < TOL


def test_sponge_fwd_rvs_match_oracle():
    nz, nx, nB, nt = 60, 100, 20, 400
    v = _case(nz, nx, 31)
    dt = 1e-3
    w = synth.ricker(20, nt, dt)
    d_ref, s_ref = oracle.fwd_sponge_2d(v, nt, nB, 10.0, 10.0, dt, src_iz=[0], src_ix=[50],
                                        wavelet=w, want_snapshot=True)
    d_gpu, s_gpu = sssi_amd.fwd_sponge_2d_points(v, nt, nB, 10.0, 10.0, dt, [0], [50], w,
                                                 want_snapshot=True)
    assert rel_l2(d_gpu, d_ref) < TOL
    assert rel_l2(s_gpu, s_ref) < TOL
    m_ref, r_ref = oracle.rvs_sponge_2d(v, d_ref, nB, 10.0, 10.0, dt, want_snapshot=True)
    m_gpu, r_gpu = sssi_amd.rvsTimeSpongeFor2dAw(v, d_ref, nB, 10.0, 10.0, dt)
    assert rel_l2(m_gpu, m_ref) < TOL
    assert rel_l2(r_gpu, r_ref) < TOL


@pytest.mark.parametrize("nz,nx,nB", [(5, 3, 1), (17, 65, 8), (33, 130, 0), (70, 64, 31)])
def test_sponge_odd_shapes_match_oracle(nz, nx, nB):
    """Grids smaller than one 64 x 16 block tile or not a multiple of it: rows
    and columns past the model compute nothing that is stored, and the zero
    padding they would overwrite stays the boundary.  Several sources, at the
    corners and on the bottom row, sum per cell in source order."""
    v = _case(nz, nx, 50 + nz)
    nt, dt = 150, 1e-3
    w = synth.ricker(20, nt, dt)
    iz = [0, nz - 1, nz - 1, nz // 2, 0]
    ix = [0, nx - 1, 0, nx // 2, 0]
    ww = np.stack([w * (1.0 + 0.25 * k) for k in range(len(iz))], axis=1)
    d_ref, s_ref = oracle.fwd_sponge_2d(v, nt, nB, 10.0, 10.0, dt, src_iz=iz, src_ix=ix,
                                        wavelet=ww, want_snapshot=True, snap_stride=7)
    d_gpu, s_gpu = sssi_amd.fwd_sponge_2d_points(v, nt, nB, 10.0, 10.0, dt, iz, ix, ww,
                                                 want_snapshot=True, snap_stride=7)
    assert rel_l2(d_gpu, d_ref) < TOL
    assert rel_l2(s_gpu, s_ref) < TOL
    m_ref, r_ref = oracle.rvs_sponge_2d(v, d_ref, nB, 10.0, 10.0, dt, want_snapshot=True,
                                        snap_stride=5)
    m_gpu, r_gpu = sssi_amd.rvs_sponge_2d_ex(v, d_ref, nB, 10.0, 10.0, dt, want_snapshot=True,
                                             snap_stride=5)
    assert rel_l2(m_gpu, m_ref) < TOL
    assert rel_l2(r_gpu, r_ref) < TOL


def test_handle_repeat_runs_are_identical():
    nz, nx, nt = 50, 128, 200
    v = _case(nz, nx, 41)
    P = sssi_amd.Propagator(v, 3, 10, 10.0, 10.0, 1e-3)
    P.set_points([0], [64], synth.ricker(20, nt, 1e-3))
    P.run(nt)
    d1 = P.data(nt)
    P.reset()
    P.run(nt)
    d2 = P.data(nt)
    assert np.array_equal(d1, d2)


def test_invalid_arguments_raise():
    v = np.full((20, 30), 2000.0)
    with pytest.raises(sssi_amd.SssiError):
        sssi_amd.fwd_cpml_2d_points(v, 10, 9, 5, 10.0, 10.0, 1e-3, [0], [5], np.ones(10))
    with pytest.raises(sssi_amd.SssiError):
        sssi_amd.fwd_cpml_2d_points(v, 10, 2, 16, 10.0, 10.0, 1e-3, [0], [5], np.ones(10))
    with pytest.raises(sssi_amd.SssiError):
        sssi_amd.fwd_cpml_2d_points(v, 10, 2, 5, 10.0, 10.0, 1e-3, [25], [5], np.ones(10))
