"""The interior march (a wavefront walking NG groups of rz rows with its z
window in registers) against the tile path (one 64 x rz tile per wavefront)
at the same rz: identical arithmetic, so identical results bit for bit
(between different rz, tiles that straddle the edge of a damping strip use
the two-stage form for more or fewer rows, so results differ in the last
bits; each is checked against the FP64 oracle elsewhere) -- forward
seismograms and snapshots with sources inside and outside the march region,
reverse propagation with surface injection, a model with no damping strips
(the march then starts at column 0), and batched BF16-checkpoint RTM images.
The march layout is chosen per grid from SSSI_NG (0 = tile path only;
default 3 at rz 4) and SSSI_RZ when the grid is created.
"""
import os

import numpy as np
import pytest

import oracle
import sssi_amd
from sssi_amd import synth

pytestmark = pytest.mark.gpu


def with_layout(ng, rz, fn):
    old = {k: os.environ.get(k) for k in ("SSSI_NG", "SSSI_RZ")}
    os.environ["SSSI_NG"] = str(ng)
    os.environ["SSSI_RZ"] = str(rz)
    try:
        return fn()
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def model(nz, nx, seed):
    v = synth.layered_2d(nz, nx, 6, 1500.0, 4000.0, seed)
    v = v * (1.0 + synth.gaussian_perturbation((nz, nx), 0.05, 4, seed + 1))
    return np.clip(v, 1500.0, 4000.0)


LAYOUTS = [(2, 4), (3, 4), (4, 4), (8, 4), (2, 8), (4, 8)]


@pytest.mark.parametrize("N", [2, 5, 8])
def test_forward_march_equals_tile_path(N):
    nz, nx, nB, nt = 230, 610, 20, 260
    v = model(nz, nx, 30 + N)
    dt = 4e-4
    w = synth.ricker(25, nt, dt)
    iz = [0, 0, 97, 150, 225, 3]
    ix = [305, 64, 333, 600, 100, 5]
    ww = np.stack([w * (1 + 0.1 * k) for k in range(len(iz))], axis=1)

    def run():
        return sssi_amd.fwd_cpml_2d_points(v, nt, N, nB, 5.0, 5.0, dt, iz, ix, ww,
                                           want_snapshot=True, snap_stride=13)
    base = {rz: with_layout(0, rz, run) for rz in (4, 8)}
    d0, s0 = base[4]
    for ng, rz in LAYOUTS:
        d1, s1 = with_layout(ng, rz, run)
        assert np.array_equal(base[rz][0], d1), (ng, rz)
        assert np.array_equal(base[rz][1], s1), (ng, rz)
    d_ref, _ = oracle.fwd_cpml_2d(v, nt, N, nB, 5.0, 5.0, dt, src_iz=iz, src_ix=ix, wavelet=ww)
    assert np.linalg.norm(d0 - d_ref) / np.linalg.norm(d_ref) < 1e-5


@pytest.mark.parametrize("mode", ["1", "2"])
def test_x_exchange_variants_equal_lds_exchange(mode):
    """SSSI_XDPP=1 (DPP wave shifts) and =2 (shifted global loads), at
    nDiffOrder 5, 4-row tiles, 3 groups: the interior march takes its x
    neighbours without the LDS round trip -- the same operands in the same
    order, so the same bits, forward and reverse."""
    nz, nx, nB, N, nt = 230, 610, 20, 5, 200
    v = model(nz, nx, 35)
    dt = 4e-4
    w = synth.ricker(25, nt, dt)
    iz, ix = [0, 97, 150], [305, 333, 600]
    ww = np.stack([w * (1 + 0.1 * k) for k in range(3)], axis=1)
    data = np.zeros((nx, nt), order="F")
    data[:, :] = w[None, :] * np.linspace(0.5, 1.5, nx)[:, None]

    def run():
        f = sssi_amd.fwd_cpml_2d_points(v, nt, N, nB, 5.0, 5.0, dt, iz, ix, ww,
                                        want_snapshot=True, snap_stride=13)
        r = sssi_amd.rvs_cpml_2d_ex(v, data, N, nB, 5.0, 5.0, dt, want_snapshot=True,
                                    snap_stride=17)
        return f + r
    lds = with_layout(3, 4, run)
    os.environ["SSSI_XDPP"] = mode
    try:
        alt = with_layout(3, 4, run)
    finally:
        os.environ.pop("SSSI_XDPP", None)
    for a, b in zip(lds, alt):
        assert np.array_equal(a, b)


def test_reverse_march_equals_tile_path():
    nz, nx, nB, N, nt = 200, 450, 15, 4, 240
    v = model(nz, nx, 41)
    dt = 5e-4
    data = np.zeros((nx, nt), order="F")
    data[:, :] = synth.ricker(20, nt, dt)[None, :] * np.linspace(0.5, 1.5, nx)[:, None]

    def run():
        return sssi_amd.rvs_cpml_2d_ex(v, data, N, nB, 5.0, 5.0, dt, want_snapshot=True,
                                       snap_stride=11)
    base = {rz: with_layout(0, rz, run) for rz in (4, 8)}
    for ng, rz in LAYOUTS:
        m1, s1 = with_layout(ng, rz, run)
        assert np.array_equal(base[rz][0], m1), (ng, rz)
        assert np.array_equal(base[rz][1], s1), (ng, rz)


def test_no_damping_strips():
    nz, nx, N, nt = 140, 320, 3, 150
    v = model(nz, nx, 51)
    dt = 5e-4
    w = synth.ricker(30, nt, dt)

    def run():
        return sssi_amd.fwd_cpml_2d_points(v, nt, N, 0, 5.0, 5.0, dt, [0, 70], [0, 160],
                                           np.stack([w, w], axis=1), want_snapshot=True,
                                           snap_stride=10)
    base = {rz: with_layout(0, rz, run) for rz in (4, 8)}
    d0, s0 = base[4]
    for ng, rz in LAYOUTS:
        d1, s1 = with_layout(ng, rz, run)
        assert np.array_equal(base[rz][0], d1) and np.array_equal(base[rz][1], s1), (ng, rz)
    d_ref, s_ref = oracle.fwd_cpml_2d(v, nt, N, 0, 5.0, 5.0, dt, src_iz=[0, 70], src_ix=[0, 160],
                                      wavelet=np.stack([w, w], axis=1), want_snapshot=True,
                                      snap_stride=10)
    assert np.linalg.norm(s0 - s_ref) / np.linalg.norm(s_ref) < 1e-5


@pytest.mark.parametrize("fmt", ["bf16", "fp32"])
def test_rtm_march_equals_tile_path(fmt):
    nz, nx, nB, N, nt = 150, 900, 20, 4, 300
    V = model(nz, nx, 61)
    VS = synth.smooth_copy(V, 6)
    dt = 5e-4
    w = synth.ricker(25, nt, dt)
    shots = [(0, 400, 200), (250, 500, 240), (600, 300, 150)]

    def run():
        r = sssi_amd.Rtm2D(V, VS, N, nB, 5.0, 5.0, dt, nt, img_stride=2, ck_format=fmt)
        for xb, xc, s in shots:
            r.add_shot(xb, xc, s, w)
        r.migrate(batched=True)
        return r.stack(), [r.shot(k)["image"] for k in range(len(shots))]
    base = {rz: with_layout(0, rz, run) for rz in (4, 8)}
    for ng, rz in LAYOUTS:
        st1, im1 = with_layout(ng, rz, run)
        assert np.array_equal(base[rz][0], st1), (ng, rz)
        for a, b in zip(base[rz][1], im1):
            assert np.array_equal(a, b), (ng, rz)


def test_strip_banding_per_kind_equals_default_order():
    """SSSI_STRIPBAND2=2 only changes the order in which the strip tiles are
    spread over the XCDs (banded per kind of strip in reach): every tile still
    runs once per step, so forward and reverse results are the same bits."""
    nz, nx, nB, N, nt = 230, 610, 20, 5, 150
    v = model(nz, nx, 36)
    dt = 4e-4
    w = synth.ricker(25, nt, dt)
    iz, ix = [0, 210], [300, 15]
    ww = np.stack([w, -0.5 * w], axis=1)
    data = np.zeros((nx, nt), order="F")
    data[:, :] = w[None, :] * np.linspace(0.5, 1.5, nx)[:, None]

    def run():
        f = sssi_amd.fwd_cpml_2d_points(v, nt, N, nB, 5.0, 5.0, dt, iz, ix, ww,
                                        want_snapshot=True, snap_stride=11)
        r = sssi_amd.rvs_cpml_2d_ex(v, data, N, nB, 5.0, 5.0, dt, want_snapshot=True,
                                    snap_stride=11)
        return f + r
    base = run()
    os.environ["SSSI_STRIPBAND2"] = "2"
    try:
        alt = run()
    finally:
        os.environ.pop("SSSI_STRIPBAND2", None)
    for a, b in zip(base, alt):
        assert np.array_equal(a, b)
