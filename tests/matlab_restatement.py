"""Vectorised numpy restatement of the reference's MATLAB propagators.

An independent second restatement used to pin the C oracle: it follows the
array-slicing formulation of the MATLAB files (not the loop/index structure of
the C-MEX files the oracle follows), with the same floating-point operation
order, so the two agree bit for bit when both are right.

  fwd_cpml_2d   src/fd_matlab/fwdTimeCpmlFor2dAw.m:46-113, diffOperator.m:19-68
  rvs_cpml_2d   src/fd_matlab/rvsTimeCpmlFor2dAw.m:30-127
  fwd_cpml_3d   src/fwdTimeCpmlFor3dAw.m:37-160
  fwd_sponge_2d src/fwdTimeSpongeFor2dAw.m:37-106
  rvs_sponge_2d src/rvsTimeSpongeFor2dAw.m:37-111

Index sets are written 0-based: MATLAB ``a:b`` (1-based, inclusive) becomes
``range(a-1, b)``.  Intended for tiny grids only.
"""
import math

import numpy as np

# libm scalar exp/pow (numpy array exp/pow may use SIMD kernels that differ
# from libm in the last bit; the reference calls libm).
_exp = np.frompyfunc(math.exp, 1, 1)


def lexp(a):
    return np.asarray(_exp(np.asarray(a, dtype=np.float64)), dtype=np.float64)


def diff_operator(data, c, d, dim):
    """src/fd_matlab/diffOperator.m: staggered first derivative along ``dim``
    (0-based axis); output is shorter by k = 2N-1 along that axis."""
    order = len(c)
    k = 2 * order - 1
    n = data.shape[dim]
    out_shape = list(data.shape)
    out_shape[dim] = n - k
    dd = np.zeros(out_shape)
    for ii in range(1, order + 1):
        # idx1 = ((order+1):(n-k+order)) + (ii-1); idx2 = ((order+1):(n-k+order)) - ii
        idx1 = np.arange(order + 1, n - k + order + 1) + (ii - 1) - 1
        idx2 = np.arange(order + 1, n - k + order + 1) - ii - 1
        a = np.take(data, idx1, axis=dim)
        b = np.take(data, idx2, axis=dim)
        dd = dd + c[ii - 1] * (a - b) / d
    return dd


def damp_pml(u, v, L):
    """src/dampPml.m: d = d0 .* (u/L).^2 with d0 = -(3 v)/(2 L) log(R)."""
    R = 1e-6
    d0 = -(3 * v) / (2 * L) * math.log(R)
    q = u / L
    return d0 * (q * q)


def damp_pml_cmex(u, v, L):
    """finiteDifference.c:253-280 order: d0*(u/L)*(u/L)."""
    logR = math.log(1e-6)
    d0 = -(3.0 * v) / (2 * L) * logR
    return d0 * (u / L) * (u / L)


def fwd_cpml_2d(v, source, N, nB, dz, dx, dt, coeff, cmex_damping=True):
    """Returns (data[nx, nt], snapshot[nz, nx, nt]).  ``cmex_damping`` selects
    the C-MEX order of the damping product (the MEX file is what MATLAB runs)."""
    return _cpml_2d(v, source, None, N, nB, dz, dx, dt, coeff, cmex_damping)


def rvs_cpml_2d(v, data, N, nB, dz, dx, dt, coeff, cmex_damping=True):
    """Returns (model[nz+2k, nx+2k], snapshot[nz, nx, nt])."""
    return _cpml_2d(v, None, data, N, nB, dz, dx, dt, coeff, cmex_damping)


def _cpml_2d(v, source, data_in, N, nB, dz, dx, dt, coeff, cmex_damping):
    v = np.asarray(v, dtype=np.float64)
    nz, nx = v.shape
    reverse = data_in is not None
    nt = data_in.shape[1] if reverse else source.shape[2]
    k = 2 * N - 1
    fdm = np.zeros((nz + 2 * k, nx + 2 * k, 3))
    dp = damp_pml_cmex if cmex_damping else damp_pml
    ixb = np.arange(1, nB + 1)
    x_left = dp(np.tile(ixb[::-1] * dx, (nz, 1)), v[:, :nB], nB * dx)
    x_right = dp(np.tile(ixb * dx, (nz, 1)), v[:, nx - nB:], nB * dx)
    xdamp = np.concatenate([x_left, np.zeros((nz, nx - 2 * nB)), x_right], axis=1)
    xb = lexp(-xdamp * dt)
    z_down = dp(np.tile((ixb * dz)[:, None], (1, nx)), v[nz - nB:, :], nB * dz)
    zdamp = np.concatenate([np.zeros((nz - nB, nx)), z_down], axis=0)
    zb = lexp(-zdamp * dt)

    zPhi = np.zeros((nz + 2 * k, nx))
    xPhi = np.zeros((nz, nx + 2 * k))
    zA = np.zeros((nz + 2 * k, nx))
    xA = np.zeros((nz, nx + 2 * k))
    zPsi = np.zeros((nz + k, nx))
    xPsi = np.zeros((nz, nx + k))
    zP = np.zeros((nz + k, nx))
    xP = np.zeros((nz, nx + k))
    vdtSq = (v * dt) * (v * dt)
    izi = np.arange(k, nz + k)                 # (1+k):(nz+k)
    ixi = np.arange(k, nx + k)
    izl = np.arange(N - 1, nz + 2 * k - N)     # nDiffOrder:(nz+2k-nDiffOrder)
    ixl = np.arange(N - 1, nx + 2 * k - N)
    data = np.zeros((nx, nt))
    snapshot = np.zeros((nz, nx, nt))
    order = range(nt - 1, -1, -1) if reverse else range(nt)
    for it in order:
        cur = fdm[:, :, 1]
        zPhi[izi, :] = zb * zPhi[izi, :] + (zb - 1) * diff_operator(cur[np.ix_(izl + 1, ixi)], coeff, dz, 0)
        xPhi[:, ixi] = xb * xPhi[:, ixi] + (xb - 1) * diff_operator(cur[np.ix_(izi, ixl + 1)], coeff, dx, 1)
        zA[izl, :] = diff_operator(cur[:, ixi], coeff, dz, 0) + zPhi[izl, :]
        xA[:, ixl] = diff_operator(cur[izi, :], coeff, dx, 1) + xPhi[:, ixl]
        dzA = diff_operator(zA[izl, :], coeff, dz, 0)
        dxA = diff_operator(xA[:, ixl], coeff, dx, 1)
        zPsi[izi, :] = zb * zPsi[izi, :] + (zb - 1) * dzA
        xPsi[:, ixi] = xb * xPsi[:, ixi] + (xb - 1) * dxA
        zP[izi, :] = dzA + zPsi[izi, :]
        xP[:, ixi] = dxA + xPsi[:, ixi]
        if reverse:
            s = np.zeros((nz, nx))
            s[0, :] = data_in[:, it]
        else:
            s = source[:, :, it]
        fdm[np.ix_(izi, ixi, [2])] = (vdtSq * (zP[izi, :] + xP[:, ixi] + s)
                                      + 2 * cur[np.ix_(izi, ixi)]
                                      - fdm[np.ix_(izi, ixi, [0])][:, :, 0])[:, :, None]
        fdm[:, :, 0] = fdm[:, :, 1]
        fdm[:, :, 1] = fdm[:, :, 2]
        data[:, it] = fdm[k, ixi, 1]
        snapshot[:, :, it] = fdm[np.ix_(izi, ixi, [1])][:, :, 0]
    if reverse:
        return fdm[:, :, 0].copy(), snapshot
    return data, snapshot


def fwd_cpml_3d(v, source, N, nB, dz, dx, dy, dt, coeff):
    """src/fwdTimeCpmlFor3dAw.m.  Returns (data[nx, ny, nt], snapshot[nz, nx, ny, nt])."""
    v = np.asarray(v, dtype=np.float64)
    nz, nx, ny, nt = source.shape
    k = 2 * N - 1
    fdm = np.zeros((nz + 2 * k, nx + 2 * k, ny + 2 * k, 3))
    ixb = np.arange(1, nB + 1)
    # index-unit damping (the reference passes u in samples and L = nBoundary)
    xl0 = damp_pml(np.tile(ixb[::-1].astype(float), (nz, 1)), v[:, :nB, 0], nB)
    xr0 = damp_pml(np.tile(ixb.astype(float), (nz, 1)), v[:, nx - nB:, 0], nB)
    xdamp = np.concatenate([np.repeat(xl0[:, :, None], ny, 2), np.zeros((nz, nx - 2 * nB, ny)),
                            np.repeat(xr0[:, :, None], ny, 2)], axis=1)
    yf0 = damp_pml(np.tile(ixb[::-1].astype(float)[None, None, :], (1, nx, 1)), v[0:1, :, :nB], nB)
    yr0 = damp_pml(np.tile(ixb.astype(float)[None, None, :], (1, nx, 1)), v[0:1, :, ny - nB:], nB)
    ydamp = np.concatenate([np.repeat(yf0, nz, 0), np.zeros((nz, nx, ny - 2 * nB)),
                            np.repeat(yr0, nz, 0)], axis=2)
    zd0 = damp_pml(np.tile(ixb.astype(float)[:, None], (1, nx)), v[nz - nB:, :, 0], nB)
    zdamp = np.concatenate([np.zeros((nz - nB, nx, ny)), np.repeat(zd0[:, :, None], ny, 2)], axis=0)
    xb = lexp(-xdamp * dt)
    yb = lexp(-ydamp * dt)
    zb = lexp(-zdamp * dt)
    zPhi = np.zeros((nz + 2 * k, nx, ny)); xPhi = np.zeros((nz, nx + 2 * k, ny)); yPhi = np.zeros((nz, nx, ny + 2 * k))
    zA = np.zeros_like(zPhi); xA = np.zeros_like(xPhi); yA = np.zeros_like(yPhi)
    zPsi = np.zeros((nz + k, nx, ny)); xPsi = np.zeros((nz, nx + k, ny)); yPsi = np.zeros((nz, nx, ny + k))
    zP = np.zeros_like(zPsi); xP = np.zeros_like(xPsi); yP = np.zeros_like(yPsi)
    vdtSq = (v * dt) * (v * dt)
    izi = np.arange(k, nz + k); ixi = np.arange(k, nx + k); iyi = np.arange(k, ny + k)
    izl = np.arange(N - 1, nz + 2 * k - N); ixl = np.arange(N - 1, nx + 2 * k - N)
    iyl = np.arange(N - 1, ny + 2 * k - N)
    data = np.zeros((nx, ny, nt))
    snapshot = np.zeros((nz, nx, ny, nt))
    for it in range(nt):
        cur = fdm[:, :, :, 1]
        zPhi[izi] = zb * zPhi[izi] + (zb - 1) * diff_operator(cur[np.ix_(izl + 1, ixi, iyi)], coeff, dz, 0)
        xPhi[:, ixi] = xb * xPhi[:, ixi] + (xb - 1) * diff_operator(cur[np.ix_(izi, ixl + 1, iyi)], coeff, dx, 1)
        yPhi[:, :, iyi] = yb * yPhi[:, :, iyi] + (yb - 1) * diff_operator(cur[np.ix_(izi, ixi, iyl + 1)], coeff, dy, 2)
        zA[izl] = diff_operator(cur[np.ix_(np.arange(nz + 2 * k), ixi, iyi)], coeff, dz, 0) + zPhi[izl]
        xA[:, ixl] = diff_operator(cur[np.ix_(izi, np.arange(nx + 2 * k), iyi)], coeff, dx, 1) + xPhi[:, ixl]
        yA[:, :, iyl] = diff_operator(cur[np.ix_(izi, ixi, np.arange(ny + 2 * k))], coeff, dy, 2) + yPhi[:, :, iyl]
        dzA = diff_operator(zA[izl], coeff, dz, 0)
        dxA = diff_operator(xA[:, ixl], coeff, dx, 1)
        dyA = diff_operator(yA[:, :, iyl], coeff, dy, 2)
        zPsi[izi] = zb * zPsi[izi] + (zb - 1) * dzA
        xPsi[:, ixi] = xb * xPsi[:, ixi] + (xb - 1) * dxA
        yPsi[:, :, iyi] = yb * yPsi[:, :, iyi] + (yb - 1) * dyA
        zP[izi] = dzA + zPsi[izi]
        xP[:, ixi] = dxA + xPsi[:, ixi]
        yP[:, :, iyi] = dyA + yPsi[:, :, iyi]
        new = (vdtSq * (zP[izi] + xP[:, ixi] + yP[:, :, iyi] + source[:, :, :, it])
               + 2 * cur[np.ix_(izi, ixi, iyi)] - fdm[:, :, :, 0][np.ix_(izi, ixi, iyi)])
        nxt = np.zeros(fdm.shape[:3])
        nxt[np.ix_(izi, ixi, iyi)] = new
        fdm[:, :, :, 0] = fdm[:, :, :, 1]
        fdm[:, :, :, 1] = nxt
        data[:, :, it] = fdm[k][np.ix_(ixi, iyi, [1])][:, :, 0]
        snapshot[:, :, :, it] = fdm[:, :, :, 1][np.ix_(izi, ixi, iyi)]
    return data, snapshot


def _sponge(v, source, data_in, nB, dz, dx, dt):
    v = np.asarray(v, dtype=np.float64)
    nz, nx = v.shape
    reverse = data_in is not None
    nt = data_in.shape[1] if reverse else source.shape[2]
    fdm = np.zeros((nz + 2, nx + 2, 3))
    iz = np.arange(nB)
    alpha = 0.015
    e = lexp(-((alpha * (nB - iz)) * (alpha * (nB - iz))))
    boundary = np.array([math.pow(x, 10) for x in e])
    ax = (v * dt / dx) * (v * dt / dx)
    az = (v * dt / dz) * (v * dt / dz)
    vdtSq = (v * dt) * (v * dt)
    b = 2 - 2 * ax - 2 * az
    izi = np.arange(1, nz + 1)
    ixi = np.arange(1, nx + 1)
    izb = np.arange(1, nz - nB + 1)            # (1:nz-nB)+1, 0-based
    data = np.zeros((nx, nt))
    snapshot = np.zeros((nz, nx, nt))
    order = range(nt - 1, -1, -1) if reverse else range(nt)
    for it in order:
        c = fdm[:, :, 1]
        p = fdm[:, :, 0]
        if reverse:
            s = np.zeros((nz, nx))
            s[0, :] = data_in[:, it]
        else:
            s = source[:, :, it]
        new = (b * c[np.ix_(izi, ixi)] - p[np.ix_(izi, ixi)]
               + ax * (c[np.ix_(izi, ixi + 1)] + c[np.ix_(izi, ixi - 1)])
               + az * (c[np.ix_(izi + 1, ixi)] + c[np.ix_(izi - 1, ixi)])
               - vdtSq * s)
        fdm[np.ix_(izi, ixi, [2])] = new[:, :, None]
        fdm[:, :, 0] = fdm[:, :, 1]
        fdm[:, :, 1] = fdm[:, :, 2]
        for ixb1 in range(2, nB + 2):          # MATLAB ixb = (1:nBoundary)+1
            col = ixb1 - 1                     # 0-based padded column
            for lev in (0, 1):
                fdm[izb, col, lev] = boundary[ixb1 - 2] * fdm[izb, col, lev]
            ixb2 = nx - nB + ixb1
            for lev in (0, 1):
                fdm[izb, ixb2 - 1, lev] = boundary[nx - ixb2 + 1] * fdm[izb, ixb2 - 1, lev]
            izb2 = nz - nB + ixb1
            for lev in (0, 1):
                fdm[izb2 - 1, :, lev] = boundary[nz - izb2 + 1] * fdm[izb2 - 1, :, lev]
        data[:, it] = fdm[1, ixi, 1]
        snapshot[:, :, it] = fdm[np.ix_(izi, ixi, [1])][:, :, 0]
    if reverse:
        return fdm[:, :, 0].copy(), snapshot
    return data, snapshot


def fwd_sponge_2d(v, source, nB, dz, dx, dt):
    return _sponge(v, source, None, nB, dz, dx, dt)


def rvs_sponge_2d(v, data, nB, dz, dx, dt):
    return _sponge(v, None, data, nB, dz, dx, dt)
