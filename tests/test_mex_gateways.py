"""MEX gateways (sssi_amd/mex) compiled against a test double of the MEX API.

CPU: every gateway compiles warning-free and rejects bad calls with a MATLAB
error identifier before touching the device (argument count, class,
dimensions, integer ranges, the reference's unchecked 2-D source, options).
The Kirchhoff gateways (eikonal2d_mex, migrate_mex) reject sources outside
the grid, non-vector or non-integer positions and a shot whose width differs
from the tables'.
GPU: the gateway outputs equal the C-ABI drop-ins bit for bit, the struct
(point) source equals the dense source, snapStride slices equal the full
snapshot's, nargout = 1 skips the snapshot, and the gateways match the FP64
oracle (rel-L2 <= 1e-5).
"""
import numpy as np
import pytest

import oracle
import sssi_amd
from sssi_amd import synth
from tests import mexmock
from tests.mexmock import MexError, call


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


@pytest.fixture(scope="module", autouse=True)
def built():
    mexmock.build()


def small_2d(nt=120):
    v = synth.layered_2d(40, 70, 3, 1500, 3000, seed=7)
    w = synth.ricker(25, nt, 1e-3)
    src = np.zeros((40, 70, nt), order="F")
    src[0, 35, :] = w
    return v, src, w


@pytest.mark.parametrize("name,nargs", [("fwdTimeCpmlFor2dAw_mex", 7),
                                        ("rvsTimeCpmlFor2dAw_mex", 7),
                                        ("fwdTimeCpmlFor3dAw_mex", 8),
                                        ("fwdTimeSpongeFor2dAw_mex", 6),
                                        ("rvsTimeSpongeFor2dAw_mex", 6),
                                        ("eikonal2d_mex", 5), ("migrate_mex", 7),
                                        ("freqCpmlFor2dAw_mex", 7)])
def test_argument_count(name, nargs):
    with pytest.raises(MexError) as e:
        call(name, *([1.0] * (nargs - 1)))
    assert e.value.ident == "sssi:nargin"
    with pytest.raises(MexError) as e:
        call(name, *([1.0] * (nargs + 2)))
    assert e.value.ident == "sssi:nargin"


def test_fwd2d_validation():
    v, src, _ = small_2d(10)
    args = [v, src, 2, 10, 10.0, 10.0, 1e-3]
    bad = list(args)
    bad[0] = mexmock.Int32Scalar()
    with pytest.raises(MexError) as e:
        call("fwdTimeCpmlFor2dAw_mex", *bad)
    assert e.value.ident == "sssi:class"
    bad = list(args)
    bad[1] = src[:, :-1, :]
    with pytest.raises(MexError) as e:
        call("fwdTimeCpmlFor2dAw_mex", *bad)
    assert e.value.ident == "sssi:size"
    for k, val in [(2, 9), (2, 0), (2, 2.5), (3, -1)]:
        bad = list(args)
        bad[k] = val
        with pytest.raises(MexError) as e:
            call("fwdTimeCpmlFor2dAw_mex", *bad)
        assert e.value.ident == "sssi:range"
    with pytest.raises(MexError) as e:
        call("fwdTimeCpmlFor2dAw_mex", *args, {"snapStride": 2.0})
    assert e.value.ident == "sssi:opts"
    with pytest.raises(MexError) as e:
        call("fwdTimeCpmlFor2dAw_mex", *args, {"snapStride": 0.0})
    assert e.value.ident == "sssi:range"
    # library-side validation (boundary thicker than the model) surfaces as sssi:run
    bad = list(args)
    bad[3] = 40
    with pytest.raises(MexError) as e:
        call("fwdTimeCpmlFor2dAw_mex", *bad)
    assert e.value.ident == "sssi:run" and "nBoundary" in e.value.msg


def test_struct_source_validation():
    v, _, w = small_2d(10)
    with pytest.raises(MexError) as e:
        call("fwdTimeCpmlFor2dAw_mex", v, {"iz": [0.0], "wavelet": w}, 2, 10, 10.0, 10.0, 1e-3)
    assert e.value.ident == "sssi:source" and "ix" in e.value.msg
    with pytest.raises(MexError) as e:
        call("fwdTimeCpmlFor2dAw_mex", v, {"iz": [0.0, 1.0], "ix": [3.0, 4.0], "wavelet": w},
             2, 10, 10.0, 10.0, 1e-3)
    assert e.value.ident == "sssi:source"
    v3 = np.full((8, 9, 10), 2000.0)
    with pytest.raises(MexError) as e:
        call("fwdTimeCpmlFor3dAw_mex", v3, {"iz": [0.0], "ix": [3.0], "wavelet": w},
             2, 2, 10.0, 10.0, 10.0, 1e-3)
    assert e.value.ident == "sssi:source" and "iy" in e.value.msg
    # indices must be whole, non-negative and finite before they reach the library
    for bad in (1.5, -1.0, float("nan"), float("inf")):
        with pytest.raises(MexError) as e:
            call("fwdTimeCpmlFor2dAw_mex", v, {"iz": [0.0], "ix": [bad], "wavelet": w[:, None]},
                 2, 10, 10.0, 10.0, 1e-3)
        assert e.value.ident == "sssi:source" and "integers" in e.value.msg


def test_rvs_validation():
    v, _, _ = small_2d(10)
    with pytest.raises(MexError) as e:
        call("rvsTimeCpmlFor2dAw_mex", v, np.zeros((69, 10)), 2, 10, 10.0, 10.0, 1e-3)
    assert e.value.ident == "sssi:size"
    with pytest.raises(MexError) as e:
        call("rvsTimeSpongeFor2dAw_mex", v, np.zeros((71, 10)), 10, 10.0, 10.0, 1e-3)
    assert e.value.ident == "sssi:size"


@pytest.mark.gpu
def test_fwd2d_gateway_equals_abi_and_oracle():
    v, src, w = small_2d()
    d_mex, s_mex = call("fwdTimeCpmlFor2dAw_mex", v, src, 3, 10, 10.0, 10.0, 1e-3)
    d_api, s_api = sssi_amd.fwdTimeCpmlFor2dAw(v, src, 3, 10, 10.0, 10.0, 1e-3)
    assert d_mex.shape == (70, 120) and s_mex.shape == (40, 70, 120)
    assert np.array_equal(d_mex, d_api) and np.array_equal(s_mex, s_api)
    d_o, s_o = oracle.fwd_cpml_2d(v, 120, 3, 10, 10.0, 10.0, 1e-3, source=src,
                                  want_snapshot=True)
    assert rel_l2(d_mex, d_o) < 1e-5 and rel_l2(s_mex, s_o) < 1e-5
    # point source, snapshot stride, nargout = 1
    pts = {"iz": [0.0], "ix": [35.0], "wavelet": w.reshape(-1, 1)}
    d_p, s_p = call("fwdTimeCpmlFor2dAw_mex", v, pts, 3, 10, 10.0, 10.0, 1e-3,
                    {"snapStride": 7.0})
    assert np.array_equal(d_p, d_mex)
    assert s_p.shape == (40, 70, 18) and np.array_equal(s_p, s_mex[:, :, ::7])
    out = call("fwdTimeCpmlFor2dAw_mex", v, src, 3, 10, 10.0, 10.0, 1e-3, nargout=1)
    assert len(out) == 1 and np.array_equal(out[0], d_mex)


@pytest.mark.gpu
def test_fwd2d_gateway_two_dimensional_source_is_one_step():
    v, src, _ = small_2d(1)
    d2, s2 = call("fwdTimeCpmlFor2dAw_mex", v, src[:, :, 0], 2, 10, 10.0, 10.0, 1e-3)
    d3, s3 = sssi_amd.fwdTimeCpmlFor2dAw(v, src, 2, 10, 10.0, 10.0, 1e-3)
    assert d2.shape == (70, 1) and s2.shape == (40, 70)
    assert np.array_equal(d2, d3) and np.array_equal(s2, s3[:, :, 0])


@pytest.mark.gpu
def test_rvs2d_gateway_equals_abi():
    v, src, _ = small_2d()
    data, _ = sssi_amd.fwdTimeCpmlFor2dAw(v, src, 3, 10, 10.0, 10.0, 1e-3, nargout=1)
    m_mex, s_mex = call("rvsTimeCpmlFor2dAw_mex", v, data, 3, 10, 10.0, 10.0, 1e-3)
    m_api, s_api = sssi_amd.rvsTimeCpmlFor2dAw(v, data, 3, 10, 10.0, 10.0, 1e-3)
    assert m_mex.shape == (40 + 10, 70 + 10)
    assert np.array_equal(m_mex, m_api) and np.array_equal(s_mex, s_api)
    m_s, s_s = call("rvsTimeCpmlFor2dAw_mex", v, data, 3, 10, 10.0, 10.0, 1e-3,
                    {"snapStride": 5.0})
    assert np.array_equal(m_s, m_mex) and np.array_equal(s_s, s_mex[:, :, ::5])


@pytest.mark.gpu
def test_sponge_gateways_equal_abi():
    v, src, _ = small_2d()
    d_mex, s_mex = call("fwdTimeSpongeFor2dAw_mex", v, src, 10, 10.0, 10.0, 1e-3)
    d_api, s_api = sssi_amd.fwdTimeSpongeFor2dAw(v, src, 10, 10.0, 10.0, 1e-3)
    assert np.array_equal(d_mex, d_api) and np.array_equal(s_mex, s_api)
    m_mex, r_mex = call("rvsTimeSpongeFor2dAw_mex", v, d_mex, 10, 10.0, 10.0, 1e-3)
    m_api, r_api = sssi_amd.rvsTimeSpongeFor2dAw(v, d_mex, 10, 10.0, 10.0, 1e-3)
    assert m_mex.shape == (42, 72)
    assert np.array_equal(m_mex, m_api) and np.array_equal(r_mex, r_api)


@pytest.mark.gpu
def test_fwd3d_gateway_equals_abi_and_oracle():
    v = synth.cake_3d_dipping(20, 24, 22, 3, 1500, 2500, seed=3)
    nt = 60
    w = synth.ricker(30, nt, 1e-3)
    src = np.zeros((20, 24, 22, nt), order="F")
    src[0, 12, 11, :] = w
    d_mex, s_mex = call("fwdTimeCpmlFor3dAw_mex", v, src, 2, 5, 10.0, 10.0, 10.0, 1e-3)
    d_api, s_api = sssi_amd.fwdTimeCpmlFor3dAw(v, src, 2, 5, 10.0, 10.0, 10.0, 1e-3)
    assert d_mex.shape == (24, 22, nt) and s_mex.shape == (20, 24, 22, nt)
    assert np.array_equal(d_mex, d_api) and np.array_equal(s_mex, s_api)
    d_o, _ = oracle.fwd_cpml_3d(v, nt, 2, 5, 10.0, 10.0, 10.0, 1e-3, src_iz=[0], src_ix=[12],
                                src_iy=[11], wavelet=w)
    assert rel_l2(d_mex, d_o) < 1e-5
    pts = {"iz": [0.0], "ix": [12.0], "iy": [11.0], "wavelet": w.reshape(-1, 1)}
    d_p, s_p = call("fwdTimeCpmlFor3dAw_mex", v, pts, 2, 5, 10.0, 10.0, 10.0, 1e-3,
                    {"snapStride": 4.0})
    assert np.array_equal(d_p, d_mex) and np.array_equal(s_p, s_mex[..., ::4])


def test_kirchhoff_gateway_validation():
    V = np.full((10, 12), 2000.0)
    for k, val, ident in [(2, 0, "sssi:range"), (2, 11, "sssi:range"), (3, 13, "sssi:range"),
                          (4, -1, "sssi:range"), (1, 0.5, "sssi:range")]:
        args = [V, 10.0, 1, 1, 3]
        args[k] = val
        with pytest.raises(MexError) as e:
            call("eikonal2d_mex", *args, nargout=1)
        assert e.value.ident == ident
    with pytest.raises(MexError) as e:
        call("eikonal2d_mex", np.ones((1, 12)), 10.0, 1, 1, 3, nargout=1)
    assert e.value.ident == "sssi:size"
    tt = np.zeros((10, 12, 12))
    xs = np.arange(1, 13, dtype=float)
    shot = np.zeros((50, 12))
    good = [tt, xs, xs, shot, 1e-3, 10, 1]
    for k, val, ident in [(1, np.ones((2, 2)), "sssi:size"), (2, xs - 0.5, "sssi:range"),
                          (3, np.zeros((50, 11)), "sssi:size"), (5, 9, "sssi:size"),
                          (6, 13, "sssi:range"), (6, 0, "sssi:range")]:
        args = list(good)
        args[k] = val
        with pytest.raises(MexError) as e:
            call("migrate_mex", *args, nargout=1)
        assert e.value.ident == ident


@pytest.mark.gpu
def test_kirchhoff_gateways_equal_abi_and_oracle():
    V = synth.smooth_copy(synth.layered_2d(24, 30, 3, 1500, 3000, seed=4), 2)
    T = call("eikonal2d_mex", V, 10.0, 1.0, 7.0, 20.0, nargout=1)[0]
    assert np.array_equal(T, sssi_amd.eikonal2d(V, 10.0, 1, 7, 20))
    assert np.array_equal(T, oracle.eikonal2d(V, 10.0, 1, 7, 20))
    xsr = np.arange(1, 31)
    tt = np.stack([oracle.eikonal2d(V, 10.0, 1, int(x), 20) for x in xsr], axis=2)
    shot = np.random.default_rng(6).standard_normal((700, 30))
    xr = xsr[::2]
    m = call("migrate_mex", tt, xr.astype(float), xsr.astype(float), shot, 1e-3, 24, 9,
             nargout=1)[0]
    assert np.array_equal(m, sssi_amd.migrate(tt, xr, xsr, shot, 1e-3, 24, 9))
    assert np.array_equal(m, oracle.kirchhoff_migrate(tt, xr, xsr, shot, 1e-3, 24, 9))
    with pytest.raises(MexError) as e:
        call("migrate_mex", tt, xr.astype(float), xsr.astype(float), shot[:40], 1e-3, 24, 9,
             nargout=1)
    assert e.value.ident == "sssi:run"


def test_freq_gateway_validation():
    m = np.full((6, 8), 1.0 / 2000.0 ** 2)
    good = [m, np.ones((48, 2)), 60.0, 2, 2, 10.0, 10.0]
    for k, val, ident in [(1, np.ones((47, 1)), "sssi:size"), (3, 9, "sssi:range"),
                          (3, 0, "sssi:range"), (0, np.ones((2, 3, 4)), "sssi:size"),
                          (2, np.ones(2), "sssi:size"), (1, mexmock.Int32Scalar(), "sssi:class")]:
        args = list(good)
        args[k] = val
        with pytest.raises(MexError) as e:
            call("freqCpmlFor2dAw_mex", *args)
        assert e.value.ident == ident


@pytest.mark.gpu
def test_freq_gateway_equals_abi():
    from sssi_amd import helmholtz
    nz, nx = 18, 22
    m = 1.0 / synth.layered_2d(nz, nx, 3, 1500, 3000, seed=2) ** 2
    w = 2 * np.pi * 14
    rng = np.random.default_rng(5)
    src = rng.standard_normal((nz, nx, 2)) + 1j * rng.standard_normal((nz, nx, 2))
    A, snap = call("freqCpmlFor2dAw_mex", m, src, w, 3.0, 4.0, 10.0, 12.0)
    A_ref = helmholtz.impedance_matrix(m, w, 3, 4, 10.0, 12.0)
    assert A.shape == A_ref.shape and (A != A_ref).nnz == 0
    assert snap.shape == src.shape
    assert np.array_equal(snap, helmholtz.solve(m, src, w, 3, 4, 10.0, 12.0))
    # a real source is the complex one with zero imaginary part; (n, nsrc) shape kept
    real = src.real.reshape(nz * nx, 2, order="F")
    _, s_real = call("freqCpmlFor2dAw_mex", m, real, w, 3.0, 4.0, 10.0, 12.0)
    assert s_real.shape == real.shape
    assert np.array_equal(s_real, helmholtz.solve(m, real + 0j, w, 3, 4, 10.0, 12.0))
    # an empty source gives A and []
    A1, s_empty = call("freqCpmlFor2dAw_mex", m, np.zeros((0, 0)), w, 3.0, 4.0, 10.0, 12.0)
    assert (A1 != A_ref).nnz == 0 and s_empty.size == 0
