"""Frequency-domain CPML Helmholtz (freqCpmlFor2dAw, src/freqCpmlFor2dAw.m) on
the GPU against the oracle: the impedance matrix bit for bit (CSR), and the
solution bit for bit (the device factorisation is the oracle's banded LU,
panel by panel, with every element updated in the same order), for orders
1..8, with and without damping, at an indefinite frequency, for a band wider
than the grid (nDiffOrder 8 on 5 rows) and narrower than a panel (1 row...
3 rows at nDiffOrder 1), with several right-hand sides in both source shapes;
plus the residual against scipy's CSR product and the argument errors."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
import sssi_amd
from sssi_amd import helmholtz, synth

pytestmark = pytest.mark.gpu


def _model(nz, nx, seed):
    v = synth.layered_2d(nz, nx, 3, 1500.0, 3000.0, seed=seed)
    return 1.0 / v ** 2


CASES = [  # nz, nx, N, nB, f
    (22, 26, 2, 5, 15.0),
    (22, 26, 2, 0, 6.0),
    (18, 30, 4, 4, 40.0),
    (3, 20, 1, 1, 9.0),
    (5, 7, 8, 2, 20.0),
    (40, 44, 3, 8, 25.0),
    (180, 12, 3, 2, 10.0),  # band 900: two rows per thread in the column replay and solve
]


@pytest.mark.parametrize("nz, nx, N, nB, f", CASES)
def test_matrix_and_solution_match_oracle_bit_for_bit(nz, nx, N, nB, f):
    print(f"helmholtz {nz}x{nx} N={N} nB={nB} f={f}", flush=True)
    m = _model(nz, nx, nz + nx)
    w = 2 * np.pi * f
    A = helmholtz.impedance_matrix(m, w, N, nB, 10.0, 12.5)
    rp, col, val = oracle.freq_cpml_2d_matrix(m, w, N, nB, 10.0, 12.5)
    assert np.array_equal(A.indptr, rp) and np.array_equal(A.indices, col)
    assert np.array_equal(A.data, val)
    rng = np.random.default_rng(nz)
    n = nz * nx
    S = np.zeros((n, 3), dtype=np.complex128)
    S[(nx // 2) * nz, 0] = 1.0
    S[:, 1] = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    S[n // 3, 2] = 0.25 - 1j
    U = helmholtz.solve(m, S, w, N, nB, 10.0, 12.5)
    Uo = oracle.freq_cpml_2d_solve(m, S, w, N, nB, 10.0, 12.5)
    assert np.array_equal(U, Uo), np.max(np.abs(U - Uo))
    for s in range(3):
        r = A @ U[:, s] + S[:, s]
        assert np.linalg.norm(r) <= 1e-11 * np.linalg.norm(S[:, s])


@pytest.mark.parametrize("nc", ["1", "4"])
@pytest.mark.parametrize("nz, nx, N, nB, f", [CASES[0], CASES[2], CASES[5], CASES[6]])
def test_replay_columns_per_workgroup_match_oracle(nz, nx, N, nB, f, nc, monkeypatch):
    """The column replay with one or four columns per workgroup (SSSI_FREQ_NC;
    two, the default, is what the other tests run) gives the oracle's solution
    bit for bit, including column counts that are not a multiple of the
    workgroup's and a band that needs two rows per thread."""
    monkeypatch.setenv("SSSI_FREQ_NC", nc)
    m = _model(nz, nx, nz + nx)
    w = 2 * np.pi * f
    n = nz * nx
    S = np.zeros((n, 2), dtype=np.complex128)
    S[(nx // 2) * nz, 0] = 1.0
    S[:, 1] = np.random.default_rng(nx).standard_normal(n)
    print(f"helmholtz nc={nc} {nz}x{nx} N={N} nB={nB}", flush=True)
    U = helmholtz.solve(m, S, w, N, nB, 10.0, 12.5)
    assert np.array_equal(U, oracle.freq_cpml_2d_solve(m, S, w, N, nB, 10.0, 12.5))


def test_drop_in_shapes_and_outputs():
    nz, nx = 20, 24
    m = _model(nz, nx, 5)
    w = 2 * np.pi * 12
    src = np.zeros((nz, nx, 2), dtype=np.complex128)
    src[1, 12, 0] = 1.0
    src[10, 3, 1] = 2j
    A, snap = sssi_amd.freqCpmlFor2dAw(m, src, w, 2, 4, 10.0, 10.0)
    assert isinstance(A, sp.csr_matrix) and A.shape == (nz * nx, nz * nx)
    assert snap.shape == src.shape
    flat = helmholtz.solve(m, src.reshape(nz * nx, 2, order="F"), w, 2, 4, 10.0, 10.0)
    assert np.array_equal(snap.reshape(nz * nx, 2, order="F"), flat)
    A1 = sssi_amd.freqCpmlFor2dAw(m, src, w, 2, 4, 10.0, 10.0, nargout=1)
    assert (A1 != A).nnz == 0
    _, none = sssi_amd.freqCpmlFor2dAw(m, np.zeros((0,)), w, 2, 4, 10.0, 10.0)
    assert none is None


def test_argument_errors():
    m = _model(10, 12, 1)
    S = np.ones((120, 1), dtype=np.complex128)
    for args in [(m, S, 1.0, 0, 2, 10.0, 10.0), (m, S, 1.0, 2, 7, 10.0, 10.0),
                 (-m, S, 1.0, 2, 2, 10.0, 10.0), (m, S, 1.0, 2, 2, 0.0, 10.0)]:
        with pytest.raises(sssi_amd.SssiError) as e:
            helmholtz.solve(*args)
        assert e.value.code == -1
