"""Pins of the oracle's frequency-domain CPML Helmholtz (freqCpmlFor2dAw,
src/freqCpmlFor2dAw.m), each chosen so that a plausible mistake fails it:

- no damping: A = m w^2 I + the composite D-D+ Laplacian (Dirichlet outside),
  entry by entry against the exact rational composite weights (a wrong
  coefficient sum, a swapped dz/dx or a wrong neighbour offset fails), and
  A is exactly symmetric;
- damping: every entry equals c(o) (1j w / (s (1j w + d)))^2 with d from the
  closed form of dampPml evaluated independently here (a wrong u, L, strip side,
  or the row's own versus the neighbour's damping fails), and no top layer;
- the solution: A U = -S to rounding, and equal to an independent sparse direct
  solver (SuperLU through scipy) on the same matrix, with and without damping,
  at a frequency where the matrix is indefinite (pivoting exercised).
"""
from fractions import Fraction

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from sssi_amd import synth

from test_oracle import _composite_weights


def _model(nz, nx, seed):
    v = synth.layered_2d(nz, nx, 3, 1500.0, 3000.0, seed=seed)
    return 1.0 / v ** 2  # squared slowness, the function's model argument


def _csr(m, w, N, nB, dz, dx):
    rp, col, val = oracle.freq_cpml_2d_matrix(m, w, N, nB, dz, dx)
    n = m.size
    return sp.csr_matrix((val, col, rp), shape=(n, n))


def _dampPml(u, v, L):
    # src/dampPml.m: d0 = -(3 v) / (2 L) log(1e-6); d = d0 (u / L)^2
    return -(3 * v) / (2 * L) * np.log(1e-6) * (u / L) ** 2


@pytest.mark.parametrize("N", [1, 2, 4])
def test_undamped_matrix_is_mass_plus_composite_laplacian(N):
    nz, nx, dz, dx, w = 9, 11, 10.0, 7.5, 2 * np.pi * 12
    m = _model(nz, nx, 3)
    A = _csr(m, w, N, 0, dz, dx).toarray()
    L = 2 * N - 1
    cw = _composite_weights(N)
    ref = np.zeros_like(A)
    for j in range(nx):
        for i in range(nz):
            r = j * nz + i
            ref[r, r] = m[i, j] * w * w + float(cw[L]) / dz ** 2 + float(cw[L]) / dx ** 2
            for o in range(-L, L + 1):
                if o == 0:
                    continue
                if 0 <= i + o < nz:
                    ref[r, r + o] = float(cw[L + o]) / dz ** 2
                if 0 <= j + o < nx:
                    ref[r, r + o * nz] = float(cw[L + o]) / dx ** 2
    assert np.array_equal(A != 0, ref != 0)
    assert np.allclose(A, ref, rtol=1e-14, atol=0)
    assert np.array_equal(A, A.T)
    # the weights annihilate constants exactly (rationals)
    assert sum(cw) == Fraction(0)


def test_damped_entries_follow_the_closed_form():
    nz, nx, nB, N, dz, dx, w = 14, 17, 4, 2, 10.0, 12.0, 2 * np.pi * 8
    m = _model(nz, nx, 5)
    A = _csr(m, w, N, nB, dz, dx).toarray()
    v = np.sqrt(1.0 / m)
    xd = np.zeros((nz, nx))
    zd = np.zeros((nz, nx))
    for j in range(nB):
        xd[:, j] = _dampPml((nB - j) * dx, v[:, j], nB * dx)
        xd[:, nx - nB + j] = _dampPml((j + 1) * dx, v[:, nx - nB + j], nB * dx)
    for i in range(nB):
        zd[nz - nB + i, :] = _dampPml((i + 1) * dz, v[nz - nB + i, :], nB * dz)
    L = 2 * N - 1
    cw = [float(x) for x in _composite_weights(N)]
    for j in range(nx):
        for i in range(nz):
            r = j * nz + i
            fz = (1j * w / (dz * (1j * w + zd[i, j]))) ** 2
            fx = (1j * w / (dx * (1j * w + xd[i, j]))) ** 2
            want = m[i, j] * w * w + cw[L] * fz + cw[L] * fx
            assert abs(A[r, r] - want) <= 1e-13 * abs(want)
            for o in (-L, -1, 1, L):
                if 0 <= i + o < nz:
                    assert abs(A[r, r + o] - cw[L + o] * fz) <= 1e-13 * abs(cw[L + o] * fz)
                if 0 <= j + o < nx:
                    assert abs(A[r, r + o * nz] - cw[L + o] * fx) <= 1e-13 * abs(cw[L + o] * fx)
    # top rows of interior columns are undamped: real off-diagonals there
    r = (nx // 2) * nz + 1
    assert A[r, r + 1].imag == 0 and A[r, r + nz].imag == 0


@pytest.mark.parametrize("nB, f", [(0, 6.0), (5, 15.0), (5, 40.0)])
def test_solution_matches_sparse_direct_solver(nB, f):
    nz, nx, N, dz, dx = 22, 26, 2, 10.0, 10.0
    m = _model(nz, nx, 7)
    w = 2 * np.pi * f
    rng = np.random.default_rng(3)
    S = np.zeros((nz * nx, 3), dtype=np.complex128)
    S[(nx // 2) * nz + 1, 0] = 1.0
    S[5 * nz + 7, 1] = 0.5 - 2j
    S[:, 2] = rng.standard_normal(nz * nx) + 1j * rng.standard_normal(nz * nx)
    U = oracle.freq_cpml_2d_solve(m, S, w, N, nB, dz, dx)
    A = _csr(m, w, N, nB, dz, dx).tocsc()
    for s in range(3):
        r = A @ U[:, s] + S[:, s]
        assert np.linalg.norm(r) <= 1e-12 * np.linalg.norm(S[:, s]) * 10
        ref = spla.spsolve(A, -S[:, s])
        assert np.linalg.norm(U[:, s] - ref) <= 1e-9 * np.linalg.norm(ref)
    if f == 40.0:
        # the undamped part of the operator is indefinite at this frequency
        Ad = _csr(m, w, N, 0, dz, dx).toarray().real
        ev = np.linalg.eigvalsh(Ad)
        assert ev.min() < 0 < ev.max()


def test_solution_is_linear_in_the_source():
    nz, nx, N, nB = 16, 18, 3, 4
    m = _model(nz, nx, 9)
    w = 2 * np.pi * 11
    rng = np.random.default_rng(4)
    a = rng.standard_normal(nz * nx) + 1j * rng.standard_normal(nz * nx)
    b = rng.standard_normal(nz * nx) + 1j * rng.standard_normal(nz * nx)
    U = oracle.freq_cpml_2d_solve(m, np.stack([a, b, a + (2 - 1j) * b], axis=1), w, N, nB,
                                  10.0, 10.0)
    lin = U[:, 0] + (2 - 1j) * U[:, 1]
    assert np.linalg.norm(U[:, 2] - lin) <= 1e-12 * np.linalg.norm(lin)
