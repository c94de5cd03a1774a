"""Pins the FP64 oracle (CPU only).

Checks, per oracle function:
  dcoef               golden coefficient table (UserGuide.tex:339-341), exact
                      Vandermonde residual for N = 1..8
  fwd CPML 2-D        composite stencil from exact rational weights (one and
                      two steps from a point source, no PML); light cone;
                      linearity; mirror symmetry; CPML absorption; analytic
                      2-D Green's function; closed-form impulse responses in
                      every damping strip (N = 1); exact onset step and
                      row/column of each strip (N = 1, 2, 3, 5)
  rvs CPML 2-D        equals the forward propagator driven by the time-reversed
                      surface data, bit for bit
  3-D CPML            equals the 2-D propagator on a y-invariant problem inside
                      the y light cone (unit spacing, where the 3-D index-unit
                      damping equals the 2-D one); closed-form impulse
                      responses in the x, y and z strips
  sponge fwd          closed-form one-, two- and three-step impulse responses in
                      every region (side strips, bottom strip, corners)
  sponge rvs          equals the forward sponge driven by the time-reversed
                      surface data, bit for bit; model = damped previous level
  imaging             brute-force cross-correlation sums
The vectorised restatement of the MATLAB files (matlab_restatement.py) is a
cross-check by the same author, not a pin: every check above also fails under
a planted index mistake in the oracle without it.
"""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle
import matlab_restatement as M
from sssi_amd import synth

# UserGuide.tex:339-341 -- the reference's own table of staggered coefficients,
# committed as a fixture
def _userguide_coeffs():
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        "staggered_coefficients.json")
    with open(path) as f:
        table = json.load(f)["coefficients"]
    return {int(N): [Fraction(a, b) for a, b in row] for N, row in table.items()}


USERGUIDE_COEFFS = _userguide_coeffs()


def test_dcoef_userguide_table():
    for N, table in USERGUIDE_COEFFS.items():
        c = oracle.dcoef(N)
        for got, want in zip(c, table):
            assert got == float(want), (N, got, want)


@pytest.mark.parametrize("N", range(1, 9))
def test_dcoef_solves_vandermonde_system(N):
    # sum_j (2j-1)^(2i-1) c_j = delta_{i1}  (src/dCoef.m:34-48), checked exactly on
    # the rationals recovered from the doubles' nearest fractions
    c = oracle.dcoef(N)
    exact = synth.dcoef_exact(N)
    for i in range(1, N + 1):
        s = sum(Fraction((2 * j - 1) ** (2 * i - 1)) * exact[j - 1] for j in range(1, N + 1))
        assert s == (1 if i == 1 else 0)
    for got, want in zip(c, exact):
        assert got == float(want)
    # residual of the rounded coefficients is at rounding level
    A = np.array([[(2 * j - 1) ** (2 * i - 1) for j in range(1, N + 1)] for i in range(1, N + 1)],
                 dtype=float)
    r = A @ c - np.eye(N)[0]
    assert np.max(np.abs(r)) < 1e-13 * np.max(np.abs(A))


def _velocity(nz, nx, seed, vmin=1500.0, vmax=3000.0):
    rng = np.random.default_rng(seed)
    return rng.uniform(vmin, vmax, (nz, nx))


@pytest.mark.parametrize("N,nB", [(1, 4), (2, 5), (3, 0), (4, 3), (5, 4), (8, 2)])
def test_cpml_2d_bit_exact_vs_matlab_restatement(N, nB):
    nz, nx, nt = 16, 22, 25
    v = _velocity(nz, nx, N)
    rng = np.random.default_rng(100 + N)
    src = np.zeros((nz, nx, nt))
    src[0, 9, :] = rng.standard_normal(nt)
    src[7, 14, :] = rng.standard_normal(nt)
    c = oracle.dcoef(N)
    d1, s1 = oracle.fwd_cpml_2d(v, nt, N, nB, 10.0, 8.0, 1e-3, source=src, want_snapshot=True)
    d2, s2 = M.fwd_cpml_2d(v, src, N, nB, 10.0, 8.0, 1e-3, c)
    assert np.array_equal(d1, d2) and np.array_equal(s1, s2)
    m1, r1 = oracle.rvs_cpml_2d(v, d1, N, nB, 10.0, 8.0, 1e-3, want_snapshot=True)
    m2, r2 = M.rvs_cpml_2d(v, d1, N, nB, 10.0, 8.0, 1e-3, c)
    assert np.array_equal(m1, m2) and np.array_equal(r1, r2)


def test_point_sources_equal_dense_source():
    nz, nx, nt, N, nB = 20, 30, 40, 3, 5
    v = _velocity(nz, nx, 3)
    w = np.random.default_rng(1).standard_normal((nt, 2))
    src = np.zeros((nz, nx, nt))
    src[0, 10, :] += w[:, 0]
    src[4, 20, :] += w[:, 1]
    d1, s1 = oracle.fwd_cpml_2d(v, nt, N, nB, 10.0, 10.0, 1e-3, source=src, want_snapshot=True)
    d2, s2 = oracle.fwd_cpml_2d(v, nt, N, nB, 10.0, 10.0, 1e-3, src_iz=[0, 4], src_ix=[10, 20],
                                wavelet=w, want_snapshot=True)
    assert np.array_equal(d1, d2) and np.array_equal(s1, s2)


def _composite_weights(N):
    c = synth.dcoef_exact(N)
    L = 2 * N - 1
    w = [Fraction(0)] * (2 * L + 1)
    for k in range(N):
        for q in range(N):
            cc = c[k] * c[q]
            w[L + k + 1 + q] += cc
            w[L + k - q] -= cc
            w[L - k + q] -= cc
            w[L - 1 - k - q] += cc
    return w


def test_composite_weights_match_published_table():
    # N=2 composite D-D+ weights w0..w3 (exact)
    w = _composite_weights(2)
    assert w[3:] == [Fraction(-365, 144), Fraction(87, 64), Fraction(-3, 32), Fraction(1, 576)]
    for N in range(1, 6):
        w = _composite_weights(N)
        assert sum(w) == 0                       # Laplacian of a constant
        L = 2 * N - 1
        # second moment: sum m^2 w_m = 2 (exact second derivative of x^2)
        assert sum(Fraction(m * m) * w[L + m] for m in range(-L, L + 1)) == 2


@pytest.mark.parametrize("N", [1, 2, 3, 5])
def test_interior_is_composite_laplacian(N):
    """Without PML the two-stage CPML step is p_new = v^2dt^2 (Lz p + Lx p + s) + 2p - p_old
    with the composite weights; check the first three steps from a point source."""
    nz, nx, nt = 40, 44, 3
    dz, dx, dt = 7.0, 9.0, 1e-3
    v = _velocity(nz, nx, 11)
    s = np.array([1.0, -0.5, 0.25])
    iz, ix = 20, 21
    _, snap = oracle.fwd_cpml_2d(v, nt, N, 0, dz, dx, dt, src_iz=[iz], src_ix=[ix], wavelet=s,
                                 want_snapshot=True)
    L = 2 * N - 1
    w = [float(x) for x in _composite_weights(N)]

    def lap(p):
        out = np.zeros_like(p)
        pp = np.pad(p, L)
        for m in range(-L, L + 1):
            out += w[L + m] / dz ** 2 * pp[L + m:L + m + nz, L:L + nx]
            out += w[L + m] / dx ** 2 * pp[L:L + nz, L + m:L + m + nx]
        return out

    vdt2 = (v * dt) ** 2
    pold = np.zeros((nz, nx))
    p = np.zeros((nz, nx))
    for t in range(nt):
        src = np.zeros((nz, nx))
        src[iz, ix] = s[t]
        pnew = vdt2 * (lap(p) + src) + 2 * p - pold
        pold, p = p, pnew
        np.testing.assert_allclose(snap[:, :, t], p, rtol=1e-12, atol=1e-12 * np.abs(p).max())


def test_light_cone_and_zero_source():
    nz, nx, nt, N = 60, 80, 6, 3
    L = 2 * N - 1
    v = _velocity(nz, nx, 5)
    d0, s0 = oracle.fwd_cpml_2d(v, nt, N, 10, 10.0, 10.0, 1e-3, src_iz=[0], src_ix=[40],
                                wavelet=np.zeros(nt), want_snapshot=True)
    assert not np.any(s0) and not np.any(d0)
    _, s = oracle.fwd_cpml_2d(v, nt, N, 10, 10.0, 10.0, 1e-3, src_iz=[30], src_ix=[40],
                              wavelet=np.ones(nt), want_snapshot=True)
    for t in range(nt):
        reach = L * t
        field = s[:, :, t]
        nzr = np.argwhere(field != 0)
        assert np.all(np.abs(nzr[:, 0] - 30) <= reach)
        assert np.all(np.abs(nzr[:, 1] - 40) <= reach)
        assert field[30, 40] != 0


def test_linearity():
    nz, nx, nt, N, nB = 30, 40, 60, 2, 8
    v = _velocity(nz, nx, 9)
    rng = np.random.default_rng(0)
    w1, w2 = rng.standard_normal(nt), rng.standard_normal(nt)
    a, b = 1.7, -0.3
    d1, _ = oracle.fwd_cpml_2d(v, nt, N, nB, 10.0, 10.0, 1e-3, src_iz=[0], src_ix=[20], wavelet=w1)
    d2, _ = oracle.fwd_cpml_2d(v, nt, N, nB, 10.0, 10.0, 1e-3, src_iz=[0], src_ix=[20], wavelet=w2)
    d3, _ = oracle.fwd_cpml_2d(v, nt, N, nB, 10.0, 10.0, 1e-3, src_iz=[0], src_ix=[20],
                               wavelet=a * w1 + b * w2)
    np.testing.assert_allclose(d3, a * d1 + b * d2, rtol=0, atol=1e-12 * np.abs(d3).max())


def test_mirror_symmetry_before_strips():
    nz, nx, nt, N, nB = 50, 81, 9, 2, 10
    v = np.tile(np.linspace(1500, 3000, nz)[:, None], (1, nx))
    w = synth.ricker(40, nt, 1e-3)
    d, s = oracle.fwd_cpml_2d(v, nt, N, nB, 10.0, 10.0, 1e-3, src_iz=[0], src_ix=[40], wavelet=w,
                              want_snapshot=True)
    # the wavefront (L = 3 cells per step: columns 13..67 after 9 steps) has not
    # reached the x strips (0..9, 71..80): exact mirror symmetry
    assert np.array_equal(d, d[::-1, :])
    assert np.array_equal(s, s[:, ::-1, :])


def test_rvs_equals_fwd_with_reversed_surface_data():
    nz, nx, nt, N, nB = 24, 30, 35, 3, 6
    v = _velocity(nz, nx, 13)
    data = np.random.default_rng(2).standard_normal((nx, nt))
    model, snap = oracle.rvs_cpml_2d(v, data, N, nB, 10.0, 10.0, 1e-3, want_snapshot=True)
    src = np.zeros((nz, nx, nt))
    src[0, :, :] = data[:, ::-1]
    _, fsnap = oracle.fwd_cpml_2d(v, nt, N, nB, 10.0, 10.0, 1e-3, source=src, want_snapshot=True)
    assert np.array_equal(snap, fsnap[:, :, ::-1])
    L = 2 * N - 1
    # model = previous level = forward snapshot at step nt-2, zero-padded by l
    assert np.array_equal(model[L:L + nz, L:L + nx], fsnap[:, :, nt - 2])
    assert not np.any(model[:L]) and not np.any(model[:, :L])


def test_cpml_absorbs_reflections():
    """Data recorded with a 20-cell CPML equal (to 1 %) the data of a domain large
    enough that no reflection returns in the window; without the PML they do not."""
    N, nB, nt, dt, dx = 2, 20, 700, 1e-3, 10.0
    nz, nx = 60, 100
    v = np.full((nz, nx), 2000.0)
    w = synth.ricker(15, nt, dt)
    Vs = synth.extBoundary(v, nB, 2)
    d_pml, _ = oracle.fwd_cpml_2d(Vs, nt, N, nB, dx, dx, dt, src_iz=[0], src_ix=[nB + 50],
                                  wavelet=w)
    big = 600
    Vb = np.full((nz + big, nx + 2 * big), 2000.0)
    d_big, _ = oracle.fwd_cpml_2d(Vb, nt, N, 0, dx, dx, dt, src_iz=[0], src_ix=[big + 50],
                                  wavelet=w)
    ref = d_big[big:big + nx, :]
    got = d_pml[nB:nB + nx, :]
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err < 0.01, err
    d_rigid, _ = oracle.fwd_cpml_2d(v, nt, N, 0, dx, dx, dt, src_iz=[0], src_ix=[50], wavelet=w)
    assert np.linalg.norm(d_rigid - ref) / np.linalg.norm(ref) > 0.1


def test_green_function_2d():
    """Homogeneous medium, interior point source: p_tt = v^2 (lap p + s) with s a cell
    value, i.e. a point-source density v^2 dx dz s(t).  The trace at distance r matches
    v^2 dx dz (G2 * s)(t), G2 = H(t - r/v) / (2 pi v sqrt(v^2 t^2 - r^2)), with the
    source sample of step m acting over [m dt, (m+1) dt] and snapshot n at (n+1) dt."""
    N, dx, dt, v0 = 5, 5.0, 4e-4, 2000.0
    nt, rcell = 400, 20
    nz = nx = 148            # walls farther than the time window's reach
    c0 = nz // 2
    w = synth.ricker(30.0, nt, dt)
    v = np.full((nz, nx), v0)
    _, snap = oracle.fwd_cpml_2d(v, nt, N, 0, dx, dx, dt, src_iz=[c0], src_ix=[c0], wavelet=w,
                                 want_snapshot=True)
    trace = snap[c0, c0 + rcell, :]
    r = rcell * dx
    F = lambda t: math.acosh(v0 * t / r) / (2 * math.pi * v0 * v0)  # integral of G2
    ana = np.zeros(nt)
    for n in range(nt):
        acc = 0.0
        for m in range(n + 1):
            tau0, tau1 = (n - m) * dt, (n - m + 1) * dt
            a0 = max(tau0, r / v0)
            if tau1 > a0:
                acc += w[m] * (F(tau1) - F(a0))
        ana[n] = v0 * v0 * dx * dx * acc
    err = np.linalg.norm(trace - ana) / np.linalg.norm(ana)
    assert err < 0.06, err
    assert abs(np.abs(trace).max() / np.abs(ana).max() - 1) < 0.02


@pytest.mark.parametrize("N,nB", [(2, 3), (4, 2), (1, 0)])
def test_cpml_3d_bit_exact_vs_matlab_restatement(N, nB):
    nz, nx, ny, nt = 10, 12, 11, 12
    v = np.random.default_rng(N).uniform(1500, 3000, (nz, nx, ny))
    src = np.zeros((nz, nx, ny, nt))
    src[0, 6, 5, :] = np.random.default_rng(7).standard_normal(nt)
    c = oracle.dcoef(N)
    d1, s1 = oracle.fwd_cpml_3d(v, nt, N, nB, 10.0, 9.0, 8.0, 1e-3, source=src, want_snapshot=True)
    d2, s2 = M.fwd_cpml_3d(v, src, N, nB, 10.0, 9.0, 8.0, 1e-3, c)
    assert np.array_equal(d1, d2) and np.array_equal(s1, s2)


def test_cpml_3d_equals_2d_on_y_invariant_problem():
    """Unit spacing makes the 3-D index-unit damping equal the 2-D one; with a
    y-invariant model and a y-line source the centre plane is the 2-D field
    until the y strips' influence (L cells per step) arrives."""
    N, nB, nt = 2, 4, 8
    L = 2 * N - 1
    nz, nx = 14, 18
    ny = 2 * nB + 2 * L * nt + 3
    dt = 1e-4
    vz = np.linspace(1500, 2500, nz)[:, None] * np.ones((1, nx))
    v3 = np.repeat(vz[:, :, None], ny, axis=2)
    w = synth.ricker(200, nt, dt)
    iy = np.arange(ny)
    d3, s3 = oracle.fwd_cpml_3d(v3, nt, N, nB, 1.0, 1.0, 1.0, dt, src_iz=np.zeros(ny, int),
                                src_ix=np.full(ny, 9), src_iy=iy,
                                wavelet=np.repeat(w[:, None], ny, 1), want_snapshot=True)
    d2, s2 = oracle.fwd_cpml_2d(vz, nt, N, nB, 1.0, 1.0, dt, src_iz=[0], src_ix=[9], wavelet=w,
                                want_snapshot=True)
    yc = ny // 2
    # the 3-D file forms d0 .* (u/L).^2 (dampPml.m), the 2-D MEX d0*(u/L)*(u/L): compare
    # bit for bit with the 2-D restatement using the .m rounding order, and to rounding
    # level with the 2-D oracle
    src2 = np.zeros((nz, nx, nt))
    src2[0, 9, :] = w
    d2m, s2m = M.fwd_cpml_2d(vz, src2, N, nB, 1.0, 1.0, dt, oracle.dcoef(N), cmex_damping=False)
    assert np.array_equal(s3[:, :, yc, :], s2m)
    assert np.array_equal(d3[:, yc, :], d2m)
    np.testing.assert_allclose(s3[:, :, yc, :], s2, rtol=0, atol=1e-12 * np.abs(s2).max())


def test_cpml_profiles_closed_form():
    nz, nx, nB = 12, 20, 5
    v = _velocity(nz, nx, 17)
    dz, dx, dt = 7.0, 9.0, 1e-3
    xb, zb = oracle.cpml_profiles_2d(v, nB, dz, dx, dt)
    # interior and top: no damping
    assert np.all(xb[:, nB:nx - nB] == 1.0) and np.all(zb[:nz - nB, :] == 1.0)
    # outermost cells: u = L, d = d0 = -(3v/2L) ln 1e-6
    d0l = -(3 * v[:, 0]) / (2 * nB * dx) * math.log(1e-6)
    np.testing.assert_allclose(-np.log(xb[:, 0]) / dt, d0l, rtol=1e-12)
    d0b = -(3 * v[-1, :]) / (2 * nB * dz) * math.log(1e-6)
    np.testing.assert_allclose(-np.log(zb[-1, :]) / dt, d0b, rtol=1e-12)
    # quadratic growth with the distance into the layer (u = (nB - j) dx on the left)
    q = -np.log(xb[:, :nB]) / dt / d0l[:, None] * 1.0
    np.testing.assert_allclose(q, (((nB - np.arange(nB)) / nB) ** 2)[None, :] *
                               (v[:, :nB] / v[:, :1]), rtol=1e-12)
    # right layer mirrors the left one in u
    qr = -np.log(xb[:, nx - nB:]) / dt
    np.testing.assert_allclose(qr / (v[:, nx - nB:] / v[:, :1]), (-np.log(xb[:, :nB]) / dt /
                               (v[:, :nB] / v[:, :1]))[:, ::-1], rtol=1e-12)


def test_cpml_profiles_3d_index_units():
    nz, nx, ny, nB = 8, 10, 12, 3
    v = np.random.default_rng(3).uniform(1500, 3000, (nz, nx, ny))
    xb, yb, zb = oracle.cpml_profiles_3d(v, nB, 1e-3)
    d0 = -(3 * v[:, 0, 0]) / (2 * nB) * math.log(1e-6)
    np.testing.assert_allclose(-np.log(xb[:, 0, 5]) / 1e-3, d0, rtol=1e-12)
    # x profile replicated over y from the first y slice; y profile from the top row
    assert np.array_equal(xb[:, :, 0], xb[:, :, 7])
    assert np.array_equal(yb[0], yb[5])
    d0y = -(3 * v[0, :, 0]) / (2 * nB) * math.log(1e-6)
    np.testing.assert_allclose(-np.log(yb[4, :, 0]) / 1e-3, d0y, rtol=1e-12)


@pytest.mark.parametrize("nB", [0, 3, 6])
def test_sponge_bit_exact_vs_matlab_restatement(nB):
    nz, nx, nt = 15, 21, 30
    v = _velocity(nz, nx, nB)
    src = np.zeros((nz, nx, nt))
    src[0, 10, :] = np.random.default_rng(4).standard_normal(nt)
    d1, s1 = oracle.fwd_sponge_2d(v, nt, nB, 10.0, 9.0, 1e-3, source=src, want_snapshot=True)
    d2, s2 = M.fwd_sponge_2d(v, src, nB, 10.0, 9.0, 1e-3)
    assert np.array_equal(d1, d2) and np.array_equal(s1, s2)
    m1, r1 = oracle.rvs_sponge_2d(v, d1, nB, 10.0, 9.0, 1e-3, want_snapshot=True)
    m2, r2 = M.rvs_sponge_2d(v, d1, nB, 10.0, 9.0, 1e-3)
    assert np.array_equal(m1, m2) and np.array_equal(r1, r2)


def test_sponge_single_step_closed_form():
    """Two steps from a unit source in the interior (no damping reach): the
    5-point scheme gives p1 = -v^2dt^2 at the source and the textbook stencil
    at the second step; then the edge factor g(0) = exp(-(0.015 nB)^2)^10."""
    nz, nx, nB = 20, 24, 6
    v = _velocity(nz, nx, 8)
    dt, dz, dx = 1e-3, 10.0, 9.0
    _, s = oracle.fwd_sponge_2d(v, 2, nB, dz, dx, dt, src_iz=[5], src_ix=[12],
                                wavelet=np.array([1.0, 0.0]), want_snapshot=True)
    vd2 = (v * dt) ** 2
    p1 = np.zeros((nz, nx))
    p1[5, 12] = -vd2[5, 12]
    assert np.array_equal(s[:, :, 0], p1)
    ax = (v * dt / dx) ** 2
    az = (v * dt / dz) ** 2
    p2 = np.zeros((nz, nx))
    p2[5, 12] = (2 - 2 * ax[5, 12] - 2 * az[5, 12]) * p1[5, 12]
    p2[5, 11] = ax[5, 11] * p1[5, 12]
    p2[5, 13] = ax[5, 13] * p1[5, 12]
    p2[4, 12] = az[4, 12] * p1[5, 12]
    p2[6, 12] = az[6, 12] * p1[5, 12]
    np.testing.assert_allclose(s[:, :, 1], p2, rtol=1e-14, atol=1e-20)
    g0 = math.exp(-(0.015 * nB) ** 2) ** 10
    assert g0 < 1
    # a source inside the left sponge column 0 is damped by g(0) once in its step
    _, s2 = oracle.fwd_sponge_2d(v, 1, nB, dz, dx, dt, src_iz=[3], src_ix=[0],
                                 wavelet=np.array([1.0]), want_snapshot=True)
    np.testing.assert_allclose(s2[3, 0, 0], -vd2[3, 0] * g0, rtol=1e-15)


def test_xcorr_image_brute_force():
    rng = np.random.default_rng(6)
    f = rng.standard_normal((5, 7, 23))
    r = rng.standard_normal((5, 7, 23))
    for stride, eps in [(1, 1e-3), (4, 1e-6)]:
        num, den, img = oracle.xcorr_image(f, r, stride=stride, eps=eps)
        ts = range(0, 23, stride)
        Mb = np.zeros((5, 7))
        s2 = np.full((5, 7), eps)
        for t in ts:
            Mb = f[:, :, t] * r[:, :, t] + Mb
            s2 = f[:, :, t] ** 2 + s2
        assert np.array_equal(num, Mb) and np.array_equal(den, s2)
        assert np.array_equal(img, Mb / s2)


def test_rtm_shot_composes_fwd_rvs_and_imaging():
    nz, nx, nt, N, nB = 20, 30, 50, 2, 5
    V = synth.layered_2d(nz, nx, 3, 1500, 3000, seed=1)
    VS = synth.smooth_copy(V, 2)
    w = synth.ricker(25, nt, 1e-3)
    out = oracle.rtm_shot_cpml_2d(V, VS, nt, [0], [15], w, N, nB, 10.0, 10.0, 1e-3,
                                  img_stride=3, eps=1e-3)
    dT, _ = oracle.fwd_cpml_2d(V, nt, N, nB, 10.0, 10.0, 1e-3, src_iz=[0], src_ix=[15], wavelet=w)
    dS, fs = oracle.fwd_cpml_2d(VS, nt, N, nB, 10.0, 10.0, 1e-3, src_iz=[0], src_ix=[15],
                                wavelet=w, want_snapshot=True)
    _, rs = oracle.rvs_cpml_2d(V, dT - dS, N, nB, 10.0, 10.0, 1e-3, want_snapshot=True)
    num, den, img = oracle.xcorr_image(fs, rs, stride=3, eps=1e-3)
    assert np.array_equal(out["data_true"], dT)
    assert np.array_equal(out["num"], num) and np.array_equal(out["den"], den)
    assert np.array_equal(out["image"], img)


# --------------------------------------------------------------------------
# Damping-strip conventions, pinned independently of any restatement of the
# loops: closed-form responses to a single impulse and exact onset times.
#
# CPML (fwdTimeCpmlFor2dAw_mex.c:185-253): the half point h+1/2 between rows
# h and h+1 carries Phi damped with b(h) (the row above it), and the integer
# point i carries Psi damped with b(i).  With Phi = Psi = 0 before the step,
# A(h+1/2) = b(h) D+p(h+1/2) and P(i) = b(i) D-A(i).  For N = 1 and an impulse
# of value a at (i0, j0) the second step is therefore
#   P_z(i0-1) = a b(i0-1)^2 / dz^2,   P_z(i0+1) = a b(i0+1) b(i0) / dz^2,
#   P_z(i0)   = -a b(i0) (b(i0) + b(i0-1)) / dz^2       (x alike with b_x, dx)
# which tells b(h) at half point h+1/2 apart from b(h+1).

def _b_profiles(v, nB, dz, dx, dt):
    """b = exp(-d dt), d = -(3 v / 2L) ln(1e-6) (u/L)^2 (src/dampPml.m:16-18) on
    the strips of fwdTimeCpmlFor2dAw_mex.c:100-145: left u = (nB-j) dx,
    right u = (j-(nx-nB)+1) dx, bottom u = (i-(nz-nB)+1) dz."""
    nz, nx = v.shape
    bx = np.ones((nz, nx))
    bz = np.ones((nz, nx))
    Lx, Lz = nB * dx, nB * dz
    for j in range(nB):
        u = (nB - j) * dx
        bx[:, j] = np.exp(-(-(3 * v[:, j]) / (2 * Lx) * math.log(1e-6)) * (u / Lx) ** 2 * dt)
        jr = nx - nB + j
        u = (j + 1) * dx
        bx[:, jr] = np.exp(-(-(3 * v[:, jr]) / (2 * Lx) * math.log(1e-6)) * (u / Lx) ** 2 * dt)
    for i in range(nB):
        ir = nz - nB + i
        u = (i + 1) * dz
        bz[ir, :] = np.exp(-(-(3 * v[ir, :]) / (2 * Lz) * math.log(1e-6)) * (u / Lz) ** 2 * dt)
    return bx, bz


@pytest.mark.parametrize("i0,j0", [(19, 12), (21, 3), (17, 2), (20, 25), (22, 28)])
def test_cpml_strip_impulse_two_steps_closed_form(i0, j0):
    """N = 1, impulse at step 0 inside (or at the edge of) the bottom, left and
    right strips, including a corner; second step against the closed form."""
    nz, nx, nB, N = 24, 30, 6, 1
    dz, dx, dt = 10.0, 8.0, 1e-3
    v = _velocity(nz, nx, 31)
    _, s = oracle.fwd_cpml_2d(v, 2, N, nB, dz, dx, dt, src_iz=[i0], src_ix=[j0],
                              wavelet=np.array([1.0, 0.0]), want_snapshot=True)
    vd2 = (v * dt) ** 2
    bx, bz = _b_profiles(v, nB, dz, dx, dt)
    # the closed-form profile is the oracle's, to rounding
    oxb, ozb = oracle.cpml_profiles_2d(v, nB, dz, dx, dt)
    np.testing.assert_allclose(oxb, bx, rtol=1e-14)
    np.testing.assert_allclose(ozb, bz, rtol=1e-14)
    a = vd2[i0, j0]
    p1 = np.zeros((nz, nx))
    p1[i0, j0] = a
    assert np.array_equal(s[:, :, 0], p1)
    P = np.zeros((nz, nx))
    bzc = bz[:, j0]
    P[i0, j0] += -a * bzc[i0] * (bzc[i0] + (bzc[i0 - 1] if i0 > 0 else 1.0)) / dz ** 2
    if i0 > 0:
        P[i0 - 1, j0] += a * bzc[i0 - 1] ** 2 / dz ** 2
    if i0 + 1 < nz:
        P[i0 + 1, j0] += a * bzc[i0 + 1] * bzc[i0] / dz ** 2
    bxr = bx[i0, :]
    P[i0, j0] += -a * bxr[j0] * (bxr[j0] + (bxr[j0 - 1] if j0 > 0 else 1.0)) / dx ** 2
    if j0 > 0:
        P[i0, j0 - 1] += a * bxr[j0 - 1] ** 2 / dx ** 2
    if j0 + 1 < nx:
        P[i0, j0 + 1] += a * bxr[j0 + 1] * bxr[j0] / dx ** 2
    p2 = vd2 * P + 2 * p1
    np.testing.assert_allclose(s[:, :, 1], p2, rtol=1e-13, atol=1e-13 * abs(a) * 1e-6)
    # the strip really varies across the impulse's neighbours, so b(h) and b(h+1)
    # give different answers
    if i0 >= nz - nB:
        assert bzc[i0 - 1] != bzc[i0]


@pytest.mark.parametrize("N", [1, 2, 3, 5])
@pytest.mark.parametrize("side", ["bottom", "left", "right"])
def test_cpml_strip_onset_is_exact(N, side):
    """Before any damped memory variable sees the wave, the CPML run is the
    undamped run bit for bit; the first difference appears exactly at the step
    and at the rows / columns that the damped points' stencil reach predicts.

    Damped half points h+1/2 are those with b(h) != 1 (bottom h >= nz-nB, left
    h <= nB-1, right h >= nx-nB); Phi(h+1/2) reads p over h-N+1 .. h+N, and
    through A it changes D-A at i in h-N+1 .. h+N.  Damped integer points i
    carry Psi(i), which reads p over i-L .. i+L (the D-D+ reach, L = 2N-1) and
    changes P at i alone.  Snapshot t holds the field after step t, which
    reaches L*t cells from the source, so a path that reads p at distance D
    from the source first acts in snapshot ceil(D/L) + 1.  With b(h+1) in
    place of b(h) at half point h+1/2 the Phi path moves by one cell and this
    prediction fails."""
    L = 2 * N - 1
    nB, nt, gap = 6, 8, 4 * L
    far = L * nt + nB + N + 2  # farther than any reach in nt steps
    if side == "bottom":
        iz = 1
        nz = iz + gap + N - 1 + nB
        nx = 2 * far + 1
        jx = nx // 2
        zb0 = nz - nB
        paths = [(zb0 - L - iz, zb0),              # Psi(zb0): changes row zb0
                 (zb0 - N + 1 - iz, zb0 - N + 1)]  # Phi(zb0+1/2): rows >= zb0-N+1
    elif side == "left":
        jx = nB + N - 1 + gap
        nx = jx + far
        iz = L * nt + 1
        nz = iz + far
        paths = [(jx - (nB - 1 + L), nB - 1), (jx - (nB - 1 + N), nB - 1 + N)]
    else:
        iz = L * nt + 1
        nz = iz + far
        jx = far
        nx = jx + gap + nB + N - 1
        paths = [(nx - nB - L - jx, nx - nB), (nx - nB - N + 1 - jx, nx - nB - N + 1)]
    v = np.full((nz, nx), 2000.0)
    args = dict(src_iz=[iz], src_ix=[jx], wavelet=np.ones(nt), want_snapshot=True)
    _, sd = oracle.fwd_cpml_2d(v, nt, N, nB, 10.0, 10.0, 1e-3, **args)
    _, sf = oracle.fwd_cpml_2d(v, nt, N, 0, 10.0, 10.0, 1e-3, **args)
    first = [(-(-dist // L) + 1, edge) for dist, edge in paths]
    t1 = min(t for t, _ in first)
    assert 2 <= t1 < nt
    for t in range(t1):
        assert np.array_equal(sd[:, :, t], sf[:, :, t]), t
    diff = np.argwhere(sd[:, :, t1] != sf[:, :, t1])
    assert diff.size
    edges = [e for t, e in first if t == t1]
    if side == "bottom":
        assert diff[:, 0].min() == min(edges)
    elif side == "left":
        assert diff[:, 1].max() == max(edges)
    else:
        assert diff[:, 1].min() == min(edges)


def _sponge_g(nB):
    # fwdTimeSpongeFor2dAw.m:47-50: boundary(m) = exp(-(0.015 (nB - m + 1))^2)^10, m = 1..nB
    return np.array([math.exp(-(0.015 * (nB - m)) ** 2) ** 10 for m in range(nB)])


def _sponge_factor(nz, nx, nB):
    """Cerjan factor per cell (fwdTimeSpongeFor2dAw.m:83-92): side columns only in
    rows 0..nz-nB-1, the bottom rows' factor across every column (corners get
    the bottom factor alone)."""
    g = _sponge_g(nB)
    f = np.ones((nz, nx))
    for i in range(nz - nB):
        for j in range(nB):
            f[i, j] = g[j]
            f[i, nx - 1 - j] = g[j]
    for i in range(nz - nB, nz):
        f[i, :] = g[nz - 1 - i]
    return f


@pytest.mark.parametrize("cell", [(3, 0), (13, 0), (14, 2), (19, 0), (19, 23), (14, 21),
                                  (13, 23), (16, 11), (5, 11)])
def test_sponge_region_one_and_two_steps_closed_form(cell):
    """Impulse in a side strip, at the last row that has side damping, in the
    first bottom row, in corners, in the bottom strip and in the interior.
    Step 1: p1 = -f v^2dt^2 at the cell (f = the cell's factor).  Step 2: the
    neighbours get f(n) a(n) p1(c) and the cell f(c) b(c) p1(c): p1 is damped
    once more as the previous level, which is exactly what the update reads
    as p_old -- zero here -- so only the current level matters."""
    nz, nx, nB = 20, 24, 6
    i0, j0 = cell
    v = _velocity(nz, nx, 41)
    dt, dz, dx = 1e-3, 10.0, 9.0
    _, s = oracle.fwd_sponge_2d(v, 3, nB, dz, dx, dt, src_iz=[i0], src_ix=[j0],
                                wavelet=np.array([1.0, 0.0, 0.0]), want_snapshot=True)
    f = _sponge_factor(nz, nx, nB)
    vd2 = (v * dt) ** 2
    ax = (v * dt / dx) ** 2
    az = (v * dt / dz) ** 2
    p1 = np.zeros((nz, nx))
    p1[i0, j0] = -vd2[i0, j0] * f[i0, j0]
    np.testing.assert_allclose(s[:, :, 0], p1, rtol=1e-15, atol=0)
    p2 = np.zeros((nz, nx))
    p2[i0, j0] = f[i0, j0] * (2 - 2 * ax[i0, j0] - 2 * az[i0, j0]) * p1[i0, j0]
    for di, dj, a in [(0, 1, ax), (0, -1, ax), (1, 0, az), (-1, 0, az)]:
        i, j = i0 + di, j0 + dj
        if 0 <= i < nz and 0 <= j < nx:
            p2[i, j] = f[i, j] * a[i, j] * p1[i0, j0]
    np.testing.assert_allclose(s[:, :, 1], p2, rtol=1e-13, atol=1e-13 * abs(p1).max() * 1e-3)
    # step 3 at the cell: the previous level enters as p_old with its second
    # factor, p3(c) = f(c) [b(c) p2(c) - f(c) p1(c) + a-weighted neighbours]
    i, j = i0, j0
    nb = 0.0
    for di, dj, a in [(0, 1, ax), (0, -1, ax), (1, 0, az), (-1, 0, az)]:
        ii, jj = i + di, j + dj
        if 0 <= ii < nz and 0 <= jj < nx:
            nb += a[i, j] * p2[ii, jj]
    p3c = f[i, j] * ((2 - 2 * ax[i, j] - 2 * az[i, j]) * p2[i, j] - f[i, j] * p1[i, j] + nb)
    np.testing.assert_allclose(s[i, j, 2], p3c, rtol=1e-12)


@pytest.mark.parametrize("nB,nt", [(0, 9), (5, 31), (7, 2), (4, 1)])
def test_rvs_sponge_equals_fwd_with_reversed_surface_data(nB, nt):
    """rvsTimeSpongeFor2dAw.m:69-111 is the forward sponge scheme run for
    it = nt..1 with source(1,:) = data(:,it)' (same -v^2dt^2 sign): snapshot it
    is the forward snapshot nt-it+1 of the dense source s(0,:,t) = data(:,nt-1-t),
    and model is the previous level, which after the loop carries the
    damping factor once more than the snapshot it was stored as."""
    nz, nx = 17, 23
    v = _velocity(nz, nx, 50 + nB)
    data = np.random.default_rng(nt).standard_normal((nx, nt))
    model, snap = oracle.rvs_sponge_2d(v, data, nB, 10.0, 9.0, 1e-3, want_snapshot=True)
    src = np.zeros((nz, nx, nt))
    src[0, :, :] = data[:, ::-1]
    _, fsnap = oracle.fwd_sponge_2d(v, nt, nB, 10.0, 9.0, 1e-3, source=src, want_snapshot=True)
    assert np.array_equal(snap, fsnap[:, :, ::-1])
    want = np.zeros((nz + 2, nx + 2))
    if nt >= 2:
        want[1:-1, 1:-1] = _sponge_factor(nz, nx, nB) * fsnap[:, :, nt - 2]
    assert np.array_equal(model, want)


def _b_profiles_3d(v, nB, dt):
    """src/fwdTimeCpmlFor3dAw.m:46-70 in closed form: dampPml with u in grid
    index units and L = nB; x from the first y slice v(:, ., 1), y from the top
    row v(1, ., .) replicated over z, z from v(., ., 1) replicated over y."""
    nz, nx, ny = v.shape
    lr = math.log(1e-6)
    d = lambda vv, u: -(3 * vv) / (2 * nB) * lr * (u / nB) ** 2
    bx, by, bz = np.ones(v.shape), np.ones(v.shape), np.ones(v.shape)
    for q in range(nB):
        bx[:, q, :] = np.exp(-d(v[:, q, 0], nB - q) * dt)[:, None]
        bx[:, nx - nB + q, :] = np.exp(-d(v[:, nx - nB + q, 0], q + 1) * dt)[:, None]
        by[:, :, q] = np.exp(-d(v[0, :, q], nB - q) * dt)[None, :]
        by[:, :, ny - nB + q] = np.exp(-d(v[0, :, ny - nB + q], q + 1) * dt)[None, :]
        bz[nz - nB + q, :, :] = np.exp(-d(v[nz - nB + q, :, 0], q + 1) * dt)[:, None]
    return bx, by, bz


@pytest.mark.parametrize("cell", [(5, 6, 2), (5, 6, 12), (9, 6, 1), (4, 1, 13), (8, 12, 7)])
def test_cpml_3d_strip_impulse_two_steps_closed_form(cell):
    """The 2-D impulse check in 3-D, N = 1, in the y (front / rear), x and z
    strips and a corner: each axis contributes b(h) at half point h+1/2 and
    b(i) at the integer point, with the 3-D files' index-unit profiles."""
    nz, nx, ny, nB = 11, 14, 15, 4
    i0, j0, k0 = cell
    dz, dx, dy, dt = 10.0, 9.0, 8.0, 1e-3
    v = np.random.default_rng(61).uniform(1500, 3000, (nz, nx, ny))
    _, s = oracle.fwd_cpml_3d(v, 2, 1, nB, dz, dx, dy, dt, src_iz=[i0], src_ix=[j0],
                              src_iy=[k0], wavelet=np.array([1.0, 0.0]), want_snapshot=True)
    bx, by, bz = _b_profiles_3d(v, nB, dt)
    oxb, oyb, ozb = oracle.cpml_profiles_3d(v, nB, dt)
    np.testing.assert_allclose(oxb, bx, rtol=1e-14)
    np.testing.assert_allclose(oyb, by, rtol=1e-14)
    np.testing.assert_allclose(ozb, bz, rtol=1e-14)
    vd2 = (v * dt) ** 2
    a = vd2[i0, j0, k0]
    p1 = np.zeros(v.shape)
    p1[i0, j0, k0] = a
    assert np.array_equal(s[..., 0], p1)
    P = np.zeros(v.shape)
    for axis, b, d in ((0, bz[:, j0, k0], dz), (1, bx[i0, :, k0], dx), (2, by[i0, j0, :], dy)):
        c = [i0, j0, k0][axis]
        n = v.shape[axis]

        def at(off):
            idx = [i0, j0, k0]
            idx[axis] = c + off
            return tuple(idx)
        P[at(0)] += -a * b[c] * (b[c] + (b[c - 1] if c > 0 else 1.0)) / d ** 2
        if c > 0:
            P[at(-1)] += a * b[c - 1] ** 2 / d ** 2
        if c + 1 < n:
            P[at(1)] += a * b[c + 1] * b[c] / d ** 2
    np.testing.assert_allclose(s[..., 1], vd2 * P + 2 * p1, rtol=1e-11, atol=1e-12 * abs(a))
