"""The benchmark workloads at their real sizes, on the GPU.

2-D 120 x 240: CPML forward, 120 x 240 extended grid, nDiffOrder 2, nt 1000,
  every-10th snapshots, against the FP64 oracle (data rel-L2 <= 1e-5,
  snapshot stack rel-L2 <= 2e-5), through the point-source entry and through
  the MEX gateway with the reference's dense (nz, nx, nt) source, bit for bit.
RTM 600 x 6000: the 64-shot RTM batch on the 600 x 6000 model with BF16
  checkpoints: every per-shot image and the stack equal, bit for bit, the
  one-shot-per-launch migration of the same shots; two shots and their stack
  against the FP64 oracle (rel-L2 <= 1e-3).
3-D 420 x 440 x 440: CPML forward on the 420 x 440 x 440 grid for the full 2000
  steps: exact odd symmetry and time shift (no oracle at this size).
Serving mix: the seeded 60-job serving mix at its real sizes through the server:
  each served result equals its direct propagation bit for bit (2-D forward:
  the drop-in entry; RTM shot: a one-shot migration; 3-D forward: the drop-in
  entry), and the smallest job of each kind matches the FP64 oracle.
"""
import numpy as np
import pytest

import oracle
import sssi_amd
from sssi_amd import synth, workloads

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


@pytest.mark.timeout(300)
def test_cpml2d_120x240_matches_oracle_and_gateway():
    from tests import mexmock
    wl = workloads.cpml2d_120x240()
    V, nt, k = wl["V"], wl["nt"], wl["snap_stride"]
    args = (V, nt, wl["N"], wl["nB"], wl["dz"], wl["dx"], wl["dt"])
    print(f"2-D 120x240: {V.shape} N={wl['N']} nt={nt} snap every {k}", flush=True)
    d, s = sssi_amd.fwd_cpml_2d_points(*args, wl["src_iz"], wl["src_ix"], wl["wavelet"],
                                       want_snapshot=True, snap_stride=k)
    dr, sr = oracle.fwd_cpml_2d(*args, src_iz=wl["src_iz"], src_ix=wl["src_ix"],
                                wavelet=wl["wavelet"], want_snapshot=True, snap_stride=k)
    e_d, e_s = rel_l2(d, dr), rel_l2(s, sr)
    e_slice = max(rel_l2(s[..., q], sr[..., q]) for q in range(1, s.shape[2]))
    print(f"2-D 120x240: data rel-L2 {e_d:.3e}, snapshot stack {e_s:.3e}, worst slice "
          f"{e_slice:.3e}", flush=True)
    assert e_d < 1e-5
    assert e_s < 2e-5
    # the MEX gateway: the reference's dense (nz, nx, nt) source for the data,
    # and the point-source struct with opts.snapStride for the snapshots
    mexmock.build()
    src = np.zeros(V.shape + (nt,), order="F")
    src[wl["src_iz"][0], wl["src_ix"][0], :] = wl["wavelet"]
    (dg,) = mexmock.call("fwdTimeCpmlFor2dAw_mex", V, src, wl["N"], wl["nB"], wl["dz"],
                         wl["dx"], wl["dt"], nargout=1)
    assert np.array_equal(dg, d)
    pts = {"iz": np.array([float(wl["src_iz"][0])]), "ix": np.array([float(wl["src_ix"][0])]),
           "wavelet": np.asfortranarray(wl["wavelet"][:, None])}
    dg, sg = mexmock.call("fwdTimeCpmlFor2dAw_mex", V, pts, wl["N"], wl["nB"], wl["dz"],
                          wl["dx"], wl["dt"], {"snapStride": float(k)})
    assert np.array_equal(dg, d) and np.array_equal(sg, s)


@pytest.mark.timeout(600)
def test_rtm_600x6000_full_batch_equals_one_shot_per_launch_and_oracle():
    wl = workloads.rtm_600x6000()
    shots = wl["shots"]
    kw = (wl["V"], wl["VS"], wl["N"], wl["nB"], wl["d"], wl["d"], wl["dt"], wl["nt"])
    print(f"RTM 600x6000: {len(shots)} shots, bf16 checkpoints", flush=True)
    r = sssi_amd.Rtm2D(*kw, img_stride=wl["img_stride"], ck_format="bf16")
    for xb, xc, sx in shots:
        r.add_shot(xb, xc, sx, wl["wavelet"])
    print(f"RTM 600x6000: device {r.device_bytes() / 1e9:.1f} GB", flush=True)
    r.precompute_data(batched=True)
    r.migrate(batched=True)
    stack_b = r.stack()
    img_b = [r.shot(q)["image"] for q in range(len(shots))]
    r.migrate(batched=False)
    stack_p = r.stack()
    for q in range(len(shots)):
        assert np.array_equal(r.shot(q)["image"], img_b[q]), q
    assert np.any(stack_b != 0) and np.array_equal(stack_b, stack_p)
    r.close()
    # two shots (the narrowest windows) and their stack against the FP64 oracle
    pick = sorted(range(len(shots)), key=lambda q: shots[q][1])[:2]
    r2 = sssi_amd.Rtm2D(*kw, img_stride=wl["img_stride"], ck_format="bf16")
    stack_o = np.zeros(wl["V"].shape)
    nB = wl["nB"]
    for n, q in enumerate(pick):
        xb, xc, sx = shots[q]
        r2.add_shot(xb, xc, sx, wl["wavelet"])
        Vw = synth.extBoundary(wl["V"][:, xb:xb + xc], nB, 2)
        VSw = synth.extBoundary(wl["VS"][:, xb:xb + xc], nB, 2)
        o = oracle.rtm_shot_cpml_2d(Vw, VSw, wl["nt"], [0], [sx + nB], wl["wavelet"], wl["N"],
                                    nB, wl["d"], wl["d"], wl["dt"], img_stride=wl["img_stride"],
                                    eps=1e-3)
        e = rel_l2(img_b[q], o["image"])
        print(f"RTM 600x6000: shot {q} ({Vw.shape[0]}x{Vw.shape[1]}) image rel-L2 {e:.3e}",
              flush=True)
        assert e < 1e-3
        stack_o[:, xb:xb + xc] += o["image"][:wl["V"].shape[0], nB:nB + xc]
    r2.migrate(batched=True)
    for n, q in enumerate(pick):  # a shot's image does not depend on its batch
        assert np.array_equal(r2.shot(n)["image"], img_b[q])
    e = rel_l2(r2.stack(), stack_o)
    print(f"RTM 600x6000: 2-shot stack rel-L2 {e:.3e}", flush=True)
    assert e < 1e-3
    r2.close()


@pytest.mark.timeout(600)
def test_cpml3d_420x440x440_full_length_odd_symmetry_and_time_shift():
    wl = workloads.cpml3d_420x440x440()
    nt = wl["nt"]
    args = (wl["V"], nt, wl["N"], wl["nB"], wl["d"], wl["d"], wl["d"], wl["dt"])
    iz, ix, iy = wl["src"]
    w = wl["wavelet"]
    print(f"3-D 420x440x440: {wl['V'].shape} nt={nt}", flush=True)
    d, _ = sssi_amd.fwd_cpml_3d_points(*args, [iz], [ix], [iy], w)
    assert np.any(d != 0) and np.all(np.isfinite(d))
    dn, _ = sssi_amd.fwd_cpml_3d_points(*args, [iz], [ix], [iy], -w)
    assert np.array_equal(dn, -d)
    del dn
    k = 23
    ws = np.concatenate([np.zeros(k), w[:-k]])
    dsh, _ = sssi_amd.fwd_cpml_3d_points(*args, [iz], [ix], [iy], ws)
    assert not np.any(dsh[:, :, :k])
    assert np.array_equal(dsh[:, :, k:], d[:, :, :-k])


def _direct(kind, kw):
    if kind == "fwd2d":
        d, _ = sssi_amd.fwd_cpml_2d_points(kw["v"], kw["nt"], kw["nDiffOrder"], kw["nBoundary"],
                                           kw["dz"], kw["dx"], kw["dt"], [kw["src_iz"]],
                                           [kw["src_ix"]], kw["wavelet"])
        return d
    if kind == "fwd3d":
        iz, ix, iy = kw["src"]
        d, _ = sssi_amd.fwd_cpml_3d_points(kw["v"], kw["nt"], kw["nDiffOrder"], kw["nBoundary"],
                                           kw["dz"], kw["dx"], kw["dy"], kw["dt"], [iz], [ix],
                                           [iy], kw["wavelet"])
        return d
    nB = kw["nBoundary"]
    V, VS = kw["V"], kw["VS"]
    pz, px = V.shape[0] - nB, V.shape[1] - 2 * nB
    r = sssi_amd.Rtm2D(V[:pz, nB:nB + px], VS[:pz, nB:nB + px], kw["nDiffOrder"], nB, kw["dz"],
                       kw["dx"], kw["dt"], kw["nt"], img_stride=kw["img_stride"],
                       ck_format=kw["ck_format"])
    r.add_shot(0, px, kw["src_ix"] - nB, kw["wavelet"])
    r.migrate(batched=False)
    img = r.shot(0)["image"]
    r.close()
    return img


def _oracle(kind, kw):
    if kind == "fwd2d":
        d, _ = oracle.fwd_cpml_2d(kw["v"], kw["nt"], kw["nDiffOrder"], kw["nBoundary"], kw["dz"],
                                  kw["dx"], kw["dt"], src_iz=[kw["src_iz"]],
                                  src_ix=[kw["src_ix"]], wavelet=kw["wavelet"])
        return d
    if kind == "fwd3d":
        iz, ix, iy = kw["src"]
        d, _ = oracle.fwd_cpml_3d(kw["v"], kw["nt"], kw["nDiffOrder"], kw["nBoundary"], kw["dz"],
                                  kw["dx"], kw["dy"], kw["dt"], src_iz=[iz], src_ix=[ix],
                                  src_iy=[iy], wavelet=kw["wavelet"])
        return d
    o = oracle.rtm_shot_cpml_2d(kw["V"], kw["VS"], kw["nt"], [0], [kw["src_ix"]], kw["wavelet"],
                                kw["nDiffOrder"], kw["nBoundary"], kw["dz"], kw["dx"], kw["dt"],
                                img_stride=kw["img_stride"], eps=1e-3)
    return o["image"]


@pytest.mark.timeout(900)
def test_serving_mix_served_mix_equals_direct_and_oracle():
    jobs = workloads.serving_mix(60)
    kinds = [k for k, _ in jobs]
    print("serving mix: " + ", ".join(f"{k} x{kinds.count(k)}" for k in sorted(set(kinds))),
          flush=True)
    with sssi_amd.Server(hbm_budget_bytes=200e9, max_active=64) as s:
        ids = [getattr(s, "submit_" + kind)(**kw) for kind, kw in jobs]
        s.drain()
        served = [s.result(i) for i in ids]
    print("serving mix: served", flush=True)
    for q, ((kind, kw), res) in enumerate(zip(jobs, served)):
        ref = _direct(kind, kw)
        assert ref.shape == res.shape, (q, kind)
        assert np.array_equal(ref, res), (q, kind, rel_l2(res, ref))
    print("serving mix: every served result equals its direct run", flush=True)
    for kind, tol in (("fwd2d", 1e-5), ("rtm2d", 1e-3), ("fwd3d", 1e-5)):
        cand = [q for q in range(len(jobs)) if jobs[q][0] == kind]
        q = min(cand, key=lambda q: workloads.job_cells(*jobs[q]))
        e = rel_l2(served[q], _oracle(*jobs[q]))
        print(f"serving mix: job {q} ({kind}) vs FP64 oracle rel-L2 {e:.3e}", flush=True)
        assert e < tol
