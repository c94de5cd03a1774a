"""Seeded synthetic inputs and the reference's setup helpers (host side, numpy).

Setup helpers, same names and argument meaning as the reference's MATLAB files:
  dCoef(N)                      src/dCoef.m:34-48 (exact rational solution)
  dampPml(u, v, L)              src/dampPml.m:16-18
  extBoundary(V, nBoundary, d)  src/extBoundary.m:22-39
  ricker(f, n, dt, t0)          src/ricker.m:51-75
  layerModelGenerator(...)      src/layerModelGenerator.m:30-38

Seeded velocity models for the benchmark workloads (layered cakes, a smooth
Gaussian random perturbation, a dipping interface) and the CFL number of the
staggered scheme.
"""
from fractions import Fraction
import math

import numpy as np


def dcoef_exact(N):
    """Staggered coefficients as exact fractions: c_j = (1/(2j-1)) * prod_{m!=j}
    (2m-1)^2 / ((2m-1)^2 - (2j-1)^2), the solution of the system in src/dCoef.m."""
    out = []
    for j in range(1, N + 1):
        f = Fraction(1, 2 * j - 1)
        for m in range(1, N + 1):
            if m != j:
                a = (2 * m - 1) ** 2
                f *= Fraction(a, a - (2 * j - 1) ** 2)
        out.append(f)
    return out


def dCoef(N, kind="s"):
    """dCoef(N,'s') as float64 (each coefficient correctly rounded)."""
    if kind != "s":
        raise ValueError("only the staggered ('s') coefficients are used by the propagators")
    if not 1 <= int(N) <= 8:
        raise ValueError("nDiffOrder must be in 1..8")
    return np.array([float(c) for c in dcoef_exact(int(N))])


def dampPml(u, v, L):
    """d = d0 .* (u/L).^2, d0 = -(3 v)/(2 L) log(1e-6)  (src/dampPml.m:16-18)."""
    d0 = -(3 * np.asarray(v, dtype=np.float64)) / (2 * L) * math.log(1e-6)
    q = np.asarray(u, dtype=np.float64) / L
    return d0 * (q * q)


def extBoundary(V, nBoundary, dim=2):
    """Edge-replicating pad: left/right/bottom in 2-D, plus front/back in 3-D;
    never the top (free surface).  src/extBoundary.m:22-39."""
    V = np.asarray(V, dtype=np.float64)
    nb = int(nBoundary)
    if dim == 2:
        W = np.concatenate([np.repeat(V[:, :1], nb, 1), V, np.repeat(V[:, -1:], nb, 1)], axis=1)
        return np.concatenate([W, np.repeat(W[-1:, :], nb, 0)], axis=0)
    if dim == 3:
        W = np.concatenate([np.repeat(V[:, :1, :], nb, 1), V, np.repeat(V[:, -1:, :], nb, 1)], axis=1)
        W = np.concatenate([np.repeat(W[:, :, :1], nb, 2), W, np.repeat(W[:, :, -1:], nb, 2)], axis=2)
        return np.concatenate([W, np.repeat(W[-1:, :, :], nb, 0)], axis=0)
    raise ValueError("dim can only be 2 or 3")


def ricker(f=20.0, n=100, dt=0.001, t0=None):
    """Causal Ricker wavelet (1 - 2 pi^2 f^2 tau^2) exp(-pi^2 f^2 tau^2), tau = t - t0,
    t0 = 1/f by default (src/ricker.m:51-75).  Samples t = k*dt, k = 0..n-1."""
    if t0 is None:
        t0 = 1.0 / f
    t = np.arange(int(n)) * dt
    tau = t - t0
    return (1 - 2 * tau * tau * f ** 2 * math.pi ** 2) * np.exp(-tau ** 2 * math.pi ** 2 * f ** 2)


def layerModelGenerator(Vel, LayerSize, Width, Length=1):
    """Layered cake of sum(LayerSize) rows (src/layerModelGenerator.m:30-38)."""
    out = np.zeros((int(sum(LayerSize)), int(Width), int(Length)))
    top = 0
    for vel, size in zip(Vel, LayerSize):
        out[top:top + int(size)] = vel
        top += int(size)
    return out[:, :, 0] if Length == 1 else out


def layered_2d(nz, nx, n_layers, vmin, vmax, seed):
    """Horizontal layers: interface depths sorted random rows, velocities U[vmin, vmax]."""
    rng = np.random.default_rng(seed)
    tops = np.sort(rng.choice(np.arange(1, nz), size=n_layers - 1, replace=False))
    vel = rng.uniform(vmin, vmax, size=n_layers)
    sizes = np.diff(np.concatenate([[0], tops, [nz]]))
    return layerModelGenerator(vel, sizes, nx)


def gaussian_perturbation(shape, sigma_rel, corr_cells, seed):
    """Zero-mean smooth random field with standard deviation ``sigma_rel``: white
    noise smoothed by a Gaussian filter of ``corr_cells`` cells, then rescaled."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    w = gaussian_filter(rng.standard_normal(shape), corr_cells, mode="reflect")
    w -= w.mean()
    return w * (sigma_rel / w.std())


def smooth_copy(V, npoints):
    """Migration-velocity copy: Gaussian blur with standard deviation ``npoints`` cells."""
    from scipy.ndimage import gaussian_filter
    return gaussian_filter(np.asarray(V, dtype=np.float64), npoints, mode="nearest")


def cake_3d_dipping(nz, nx, ny, n_layers, vmin, vmax, seed, dz=10.0, dx=10.0, dy=10.0):
    """Layered cake (index order v[iz, ix, iy]) with one dipping planar interface of
    random strike and dip; below the plane the velocity is one more random draw."""
    rng = np.random.default_rng(seed)
    tops = np.sort(rng.choice(np.arange(1, nz), size=n_layers - 1, replace=False))
    vel = rng.uniform(vmin, vmax, size=n_layers)
    v = np.empty((nz, nx, ny))
    sizes = np.diff(np.concatenate([[0], tops, [nz]]))
    top = 0
    for vv, sz in zip(vel, sizes):
        v[top:top + sz] = vv
        top += sz
    strike = rng.uniform(0, 2 * math.pi)
    dip = rng.uniform(math.radians(5), math.radians(30))
    z0 = rng.uniform(0.4, 0.7) * nz * dz
    zz = np.arange(nz)[:, None, None] * dz
    xx = (np.arange(nx)[None, :, None] - nx / 2) * dx
    yy = (np.arange(ny)[None, None, :] - ny / 2) * dy
    plane = z0 + math.tan(dip) * (math.cos(strike) * xx + math.sin(strike) * yy)
    below = zz > plane
    v = np.where(below, rng.uniform(vmin, vmax), v)
    return v


def cfl_number(vmax, dt, spacings, N):
    """vmax*dt*sqrt(sum 1/d^2) * sum|c_k|.  The staggered 2N-order leapfrog scheme
    is stable for values below 1 (Nyquist-mode eigenvalue of D^-D^+ per axis is
    (2 sum|c_k| / d)^2)."""
    s = float(sum(abs(c) for c in dcoef_exact(N)))
    return vmax * dt * math.sqrt(sum(1.0 / (d * d) for d in spacings)) * s
