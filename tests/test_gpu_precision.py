"""FP32 precision of the HIP propagators at production sizes, against the
FP64 oracle: seismograms within rel-L2 1e-5; wavefield snapshots within 5e-5.

Snapshot error grows with the step count: the rounding of p(t+1) = p(t) +
[p(t) - p(t-1) + v^2 dt^2 lap] enters the next step's time difference, and a
leapfrog recursion amplifies such per-step errors by about sqrt(nt)/(w dt)
(w dt = 0.08 at 25 Hz, 0.5 ms), so the 4000-step 2-D snapshots sit at
1.7e-5 .. 3.1e-5 while the surface seismogram stays at 5.8e-6.

- 2-D: the 1000 x 3000 (+40-cell CPML, 1040 x 3080 padded) layered model
  with a 5 % Gaussian perturbation, nDiffOrder 5, dt 0.5 ms, 4000 steps, one
  surface shot (the benchmark workload, full length);
- 3-D: a 120^3 (+10-cell CPML) dipping-layer model, nDiffOrder 4, 600 steps.
The oracle runs multi-threaded (about 20 s for the 2-D case on 16 cores).
"""
import numpy as np
import pytest

import oracle
import sssi_amd
from sssi_amd import synth

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


def test_2d_full_length_benchmark_workload():
    import bench
    wl = bench.workload()
    args = (wl["V"], wl["nt"], wl["N"], wl["nB"], wl["dz"], wl["dx"], wl["dt"])
    stride = 1000
    d_gpu, s_gpu = sssi_amd.fwd_cpml_2d_points(*args, wl["src_iz"], wl["src_ix"], wl["wavelet"],
                                               want_snapshot=True, snap_stride=stride)
    d_ref, s_ref = oracle.fwd_cpml_2d(*args, src_iz=wl["src_iz"], src_ix=wl["src_ix"],
                                      wavelet=wl["wavelet"], want_snapshot=True,
                                      snap_stride=stride)
    e_data = rel_l2(d_gpu, d_ref)
    e_snap = [rel_l2(s_gpu[:, :, k], s_ref[:, :, k]) for k in range(1, s_ref.shape[2])]
    print(f"2-D 1040x3080 N=5 nt=4000: data rel-L2 {e_data:.3e}, snapshots "
          + ", ".join(f"{e:.3e}" for e in e_snap), flush=True)
    assert e_data < 1e-5
    assert max(e_snap) < 5e-5


def test_3d_precision():
    n0, nB, N, nt, dt = 120, 10, 4, 600, 1e-3
    v = synth.cake_3d_dipping(n0, n0, n0, 6, 1500.0, 4000.0, seed=11)
    V = synth.extBoundary(v, nB, 3)
    w = synth.ricker(20.0, nt, dt)
    c = (n0 + 2 * nB) // 2
    args = (V, nt, N, nB, 10.0, 10.0, 10.0, dt)
    d_gpu, s_gpu = sssi_amd.fwd_cpml_3d_points(*args, [0], [c], [c], w, want_snapshot=True,
                                               snap_stride=300)
    d_ref, s_ref = oracle.fwd_cpml_3d(*args, src_iz=[0], src_ix=[c], src_iy=[c], wavelet=w,
                                      want_snapshot=True, snap_stride=300)
    e_data = rel_l2(d_gpu, d_ref)
    e_snap = [rel_l2(s_gpu[..., k], s_ref[..., k]) for k in range(1, s_ref.shape[3])]
    print(f"3-D 130x140x140 N=4 nt=600: data rel-L2 {e_data:.3e}, snapshots "
          + ", ".join(f"{e:.3e}" for e in e_snap), flush=True)
    assert e_data < 1e-5
    assert max(e_snap) < 5e-5


@pytest.mark.parametrize("fmt,tol", [("fp32", 1e-5), ("bf16", 1e-3)])
def test_rtm_image_precision_at_benchmark_size(fmt, tol):
    """One shot of the 64-shot RTM workload (600 x 6000 model, window of the
    first shot, nDiffOrder 4, 3000 steps, imaging every 5 steps) against the
    FP64 oracle."""
    import bench
    wl = bench.workload_rtm(1)
    xb, xc, s = wl["shots"][0]
    r = sssi_amd.Rtm2D(wl["V"], wl["VS"], wl["N"], wl["nB"], wl["d"], wl["d"], wl["dt"],
                       wl["nt"], img_stride=5, ck_format=fmt)
    r.add_shot(xb, xc, s, wl["wavelet"])
    r.migrate(batched=True)
    g = r.shot(0)
    r.close()
    Vw = synth.extBoundary(wl["V"][:, xb:xb + xc], wl["nB"], 2)
    VSw = synth.extBoundary(wl["VS"][:, xb:xb + xc], wl["nB"], 2)
    o = oracle.rtm_shot_cpml_2d(Vw, VSw, wl["nt"], [0], [s + wl["nB"]], wl["wavelet"], wl["N"],
                                wl["nB"], wl["d"], wl["d"], wl["dt"], img_stride=5, eps=1e-3)
    e = rel_l2(g["image"], o["image"])
    print(f"RTM {Vw.shape[0]}x{Vw.shape[1]} nt={wl['nt']} {fmt} checkpoints: image rel-L2 {e:.3e}",
          flush=True)
    assert e < tol
