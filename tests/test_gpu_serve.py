"""GPU tests of the serving loop (continuous batching, HBM-budget admission).

- every result is identical, bit for bit, whether its job runs alone or in a
  mix of co-resident jobs of other kinds, sizes and step counts;
- a served forward job equals the direct (one-job) propagation bit for bit,
  and a served RTM shot matches the FP64 oracle (BF16 checkpoints: <= 1e-3);
- admission never lets the resident jobs exceed the device-memory budget;
- a job larger than the budget is rejected at submission.
"""
import time

import numpy as np
import pytest

import oracle
import sssi_amd
from sssi_amd import synth

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


def make_jobs():
    jobs = []
    v1 = synth.layered_2d(60, 130, 3, 1500, 3000, seed=1)
    jobs.append(("fwd2d", dict(v=v1, nt=200, nDiffOrder=2, nBoundary=10, dz=10.0, dx=10.0,
                               dt=1e-3, src_iz=0, src_ix=65, wavelet=synth.ricker(20, 200, 1e-3))))
    v2 = synth.layered_2d(90, 200, 4, 1500, 3500, seed=2)
    jobs.append(("fwd2d", dict(v=v2, nt=300, nDiffOrder=4, nBoundary=15, dz=10.0, dx=10.0,
                               dt=8e-4, src_iz=3, src_ix=100, wavelet=synth.ricker(25, 300, 8e-4))))
    V = synth.extBoundary(synth.layered_2d(40, 100, 3, 1500, 3000, seed=3), 10, 2)
    VS = synth.smooth_copy(V, 3)
    jobs.append(("rtm2d", dict(V=V, VS=VS, nt=250, nDiffOrder=3, nBoundary=10, dz=10.0, dx=10.0,
                               dt=1e-3, src_ix=60, wavelet=synth.ricker(20, 250, 1e-3),
                               img_stride=2, ck_format="bf16")))
    v3 = synth.cake_3d_dipping(24, 40, 36, 3, 1500, 2500, seed=4)
    jobs.append(("fwd3d", dict(v=v3, nt=80, nDiffOrder=2, nBoundary=5, dz=10.0, dx=10.0,
                               dy=10.0, dt=1e-3, src=(0, 20, 18),
                               wavelet=synth.ricker(25, 80, 1e-3))))
    v5 = synth.layered_2d(70, 100, 3, 1500, 2500, seed=5)
    jobs.append(("fwd2d", dict(v=v5, nt=150, nDiffOrder=2, nBoundary=8, dz=10.0, dx=10.0,
                               dt=1e-3, src_iz=0, src_ix=50, wavelet=synth.ricker(30, 150, 1e-3))))
    return jobs


def submit(server, kind, kw):
    return getattr(server, "submit_" + kind)(**kw)


def test_result_alone_equals_result_in_mix():
    jobs = make_jobs()
    with sssi_amd.Server(hbm_budget_bytes=8e9, max_active=16) as s:
        ids = []
        for kind, kw in jobs:
            ids.append(submit(s, kind, kw))
            time.sleep(0.005)   # arrivals while earlier jobs are running
        s.drain()
        mix = [s.result(i) for i in ids]
    for (kind, kw), r in zip(jobs, mix):
        with sssi_amd.Server(hbm_budget_bytes=8e9, max_active=1) as s1:
            jid = submit(s1, kind, kw)
            assert s1.wait(jid, timeout=60.0)
            alone = s1.result(jid)
        assert np.array_equal(alone, r), kind


def test_served_jobs_match_direct_path_and_oracle():
    jobs = make_jobs()
    with sssi_amd.Server(hbm_budget_bytes=8e9, max_active=16) as s:
        ids = [submit(s, kind, kw) for kind, kw in jobs]
        s.drain()
        res = [s.result(i) for i in ids]
    kw = jobs[1][1]
    d, _ = sssi_amd.fwd_cpml_2d_points(kw["v"], kw["nt"], kw["nDiffOrder"], kw["nBoundary"],
                                       kw["dz"], kw["dx"], kw["dt"], [kw["src_iz"]],
                                       [kw["src_ix"]], kw["wavelet"])
    assert np.array_equal(d, res[1])
    kw = jobs[2][1]
    o = oracle.rtm_shot_cpml_2d(kw["V"], kw["VS"], kw["nt"], [0], [kw["src_ix"]], kw["wavelet"],
                                kw["nDiffOrder"], kw["nBoundary"], kw["dz"], kw["dx"], kw["dt"],
                                img_stride=kw["img_stride"], eps=1e-3)
    assert rel_l2(res[2], o["image"]) < 1e-3
    kw = jobs[3][1]
    d3, _ = oracle.fwd_cpml_3d(kw["v"], kw["nt"], kw["nDiffOrder"], kw["nBoundary"], kw["dz"],
                               kw["dx"], kw["dy"], kw["dt"], src_iz=[0], src_ix=[20], src_iy=[18],
                               wavelet=kw["wavelet"])
    assert rel_l2(res[3], d3) < 1e-5


def test_admission_respects_budget():
    jobs = make_jobs()
    with sssi_amd.Server(hbm_budget_bytes=8e9, max_active=16) as probe:
        sizes = []
        for kind, kw in jobs:
            jid = submit(probe, kind, kw)
            sizes.append(probe.info(jid)["bytes"])
        probe.drain()
    budget = max(sizes) * 1.5   # never room for all, always room for the largest
    with sssi_amd.Server(hbm_budget_bytes=budget, max_active=16) as s:
        ids = [submit(s, kind, kw) for kind, kw in jobs for _ in range(2)]
        s.drain()
        infos = [s.info(i) for i in ids]
    events = sorted([(f["t_admit"], f["bytes"]) for f in infos] +
                    [(f["t_done"], -f["bytes"]) for f in infos], key=lambda e: (e[0], e[1]))
    used = 0.0
    for _, b in events:
        used += b
        assert used <= budget + 1e-6
    assert all(f["t_done"] >= f["t_admit"] >= f["t_arrive"] for f in infos)


def test_oversized_job_is_rejected():
    kind, kw = make_jobs()[1]
    with sssi_amd.Server(hbm_budget_bytes=1e5, max_active=4) as s:
        with pytest.raises(sssi_amd.SssiError) as e:
            submit(s, kind, kw)
        assert e.value.code == -5


def test_served_rtm_with_cached_data_true_equals_computing_it():
    """A served RTM shot given its cached dataTrue (as the reference's apps load
    it instead of recomputing, mainRtmTimeCpmlFor2dAw.m:170-184) skips the
    fwd(V) propagation and produces the same image bit for bit."""
    kind, kw = make_jobs()[2]
    nB = kw["nBoundary"]
    V, VS = kw["V"], kw["VS"]
    pz, px = V.shape[0] - nB, V.shape[1] - 2 * nB
    r = sssi_amd.Rtm2D(V[:pz, nB:nB + px], VS[:pz, nB:nB + px], kw["nDiffOrder"], nB, kw["dz"],
                       kw["dx"], kw["dt"], kw["nt"], img_stride=kw["img_stride"],
                       ck_format=kw["ck_format"])
    r.add_shot(0, px, kw["src_ix"] - nB, kw["wavelet"])
    r.precompute_data()
    dtrue = r.shot(0)["data_true"]
    r.close()
    with sssi_amd.Server(hbm_budget_bytes=8e9, max_active=4) as s:
        a = s.submit_rtm2d(**kw)
        b = s.submit_rtm2d(**kw, data_true=dtrue)
        s.drain()
        ia, ib = s.result(a), s.result(b)
    assert np.any(ia != 0) and np.array_equal(ia, ib)


def test_borrowed_rtm_inputs_equal_copied_ones():
    """Borrowed submissions (the models and dataTrue read at job setup, held
    by the Python object until the job is done) give the copied submissions'
    images bit for bit, also when the arrays are temporaries that only the
    server object keeps alive."""
    kind, kw = make_jobs()[2]
    nB = kw["nBoundary"]
    V, VS = kw["V"], kw["VS"]
    pz, px = V.shape[0] - nB, V.shape[1] - 2 * nB
    r = sssi_amd.Rtm2D(V[:pz, nB:nB + px], VS[:pz, nB:nB + px], kw["nDiffOrder"], nB, kw["dz"],
                       kw["dx"], kw["dt"], kw["nt"], img_stride=kw["img_stride"],
                       ck_format=kw["ck_format"])
    r.add_shot(0, px, kw["src_ix"] - nB, kw["wavelet"])
    r.precompute_data()
    dtrue = r.shot(0)["data_true"]
    r.close()
    rest = {k: v for k, v in kw.items() if k not in ("V", "VS")}
    with sssi_amd.Server(hbm_budget_bytes=8e9, max_active=4) as s:
        ids = [s.submit_rtm2d(**kw, borrow=False),
               s.submit_rtm2d(**kw, data_true=dtrue, borrow=False),
               s.submit_rtm2d(V=np.array(V, order="C"), VS=np.array(VS, order="C"), **rest),
               s.submit_rtm2d(V=np.array(V, order="C"), VS=np.array(VS, order="C"), **rest,
                              data_true=np.array(dtrue, order="C"))]
        s.drain()
        im = [s.result(i) for i in ids]
    assert np.any(im[0] != 0)
    for k in range(1, 4):
        assert np.array_equal(im[0], im[k]), k
