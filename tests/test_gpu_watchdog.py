"""Non-finite watchdog: a time step beyond the scheme's stability limit makes
the wavefield overflow to Inf/NaN (ordinary FP32 overflow, no fault), and every
entry point then returns SSSI_EUNSTABLE (-6) instead of handing back NaN/Inf
traces silently as the reference does (fwdTimeCpmlFor2dAw_mex.c:183-290 has no
check).  Stable runs on the same handles are unaffected."""
import numpy as np
import pytest

import sssi_amd
from sssi_amd import synth
from sssi_amd.rtm import Rtm2D
from sssi_amd.serve import Server
from tests import mexmock
from tests.mexmock import MexError, call

pytestmark = pytest.mark.gpu

EUNSTABLE = -6
NZ, NX, NT, N, NB, DX = 48, 96, 400, 2, 10, 10.0
V0 = 3000.0
DT_BAD = 0.9 * DX / V0     # Courant number 0.9 sqrt(2) (9/8 + 1/24) = 1.48
DT_OK = 0.3 * DX / V0


def _v():
    return np.full((NZ, NX), V0)


def _w(dt):
    return synth.ricker(25.0, NT, dt)


def test_fwd_cpml_2d_unstable_returns_eunstable():
    print("watchdog: fwd cpml 2d", flush=True)
    with pytest.raises(sssi_amd.SssiError) as e:
        sssi_amd.fwd_cpml_2d_points(_v(), NT, N, NB, DX, DX, DT_BAD, [0], [NX // 2], _w(DT_BAD))
    assert e.value.code == EUNSTABLE and "stability" in str(e.value)
    d, _ = sssi_amd.fwd_cpml_2d_points(_v(), NT, N, NB, DX, DX, DT_OK, [0], [NX // 2], _w(DT_OK))
    assert np.all(np.isfinite(d))


def test_rvs_and_sponge_unstable_return_eunstable():
    print("watchdog: rvs cpml / sponge", flush=True)
    data = np.random.default_rng(1).standard_normal((NX, NT))
    with pytest.raises(sssi_amd.SssiError) as e:
        sssi_amd.rvs_cpml_2d_ex(_v(), data, N, NB, DX, DX, DT_BAD, want_snapshot=False)
    assert e.value.code == EUNSTABLE
    dtb = 0.9 * DX / V0  # 5-point scheme: limit v dt/dx <= 1/sqrt(2)
    with pytest.raises(sssi_amd.SssiError) as e:
        sssi_amd.fwd_sponge_2d_points(_v(), NT, NB, DX, DX, dtb, [0], [NX // 2], _w(dtb))
    assert e.value.code == EUNSTABLE


def test_handle_reports_at_sync_and_clears_on_reset():
    print("watchdog: handle", flush=True)
    P = sssi_amd.Propagator(_v(), N, NB, DX, DX, DT_BAD)
    P.set_points([0], [NX // 2], _w(DT_BAD))
    P.run(NT)
    with pytest.raises(sssi_amd.SssiError) as e:
        P.sync()
    assert e.value.code == EUNSTABLE
    P.reset()
    P.run(5)       # a few steps stay finite even at this dt
    P.sync()
    P.close()


def test_3d_unstable_returns_eunstable():
    print("watchdog: 3d", flush=True)
    v = np.full((24, 32, 32), V0)
    dt = 0.9 * DX / V0
    w = synth.ricker(25.0, 300, dt)
    with pytest.raises(sssi_amd.SssiError) as e:
        sssi_amd.fwd_cpml_3d_points(v, 300, N, 4, DX, DX, DX, dt, [0], [16], [16], w)
    assert e.value.code == EUNSTABLE


def test_rtm_unstable_returns_eunstable():
    print("watchdog: rtm", flush=True)
    V = _v()
    r = Rtm2D(V, V * 0.95, N, NB, DX, DX, DT_BAD, NT, img_stride=5, ck_format="bf16")
    r.add_shot(0, NX, NX // 2, _w(DT_BAD))
    with pytest.raises(sssi_amd.SssiError) as e:
        r.precompute_data(batched=True)
        r.migrate(batched=True)
    assert e.value.code == EUNSTABLE
    r.close()


def test_served_job_unstable_fails_with_eunstable_others_unaffected():
    print("watchdog: server", flush=True)
    with Server(hbm_budget_bytes=4e9, max_active=4) as s:
        bad = s.submit_fwd2d(_v(), NT, N, NB, DX, DX, DT_BAD, 0, NX // 2, _w(DT_BAD))
        good = s.submit_fwd2d(_v(), NT, N, NB, DX, DX, DT_OK, 0, NX // 2, _w(DT_OK))
        with pytest.raises(sssi_amd.SssiError) as e:
            s.wait(bad)
        assert e.value.code == EUNSTABLE
        assert s.wait(good)
        d = s.result(good)
        ref, _ = sssi_amd.fwd_cpml_2d_points(_v(), NT, N, NB, DX, DX, DT_OK, [0], [NX // 2],
                                             _w(DT_OK))
        assert np.array_equal(d, ref)


def test_mex_gateway_raises_unstable_identifier():
    print("watchdog: mex", flush=True)
    mexmock.build()
    src = np.zeros((NZ, NX, NT), order="F")
    src[0, NX // 2, :] = _w(DT_BAD)
    with pytest.raises(MexError) as e:
        call("fwdTimeCpmlFor2dAw_mex", _v(), src, N, NB, DX, DX, DT_BAD, nargout=1)
    assert e.value.ident == "sssi:unstable"
