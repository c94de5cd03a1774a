"""Size-independent properties at the benchmark sizes, exact bit for bit.

- Odd symmetry: the step is linear and negation is exact in IEEE arithmetic
  (subnormals included), so data(-w) == -data(w) for any wavelet w.
- Time shift: the wavefield and the CPML memory variables stay exactly zero
  until the first nonzero source sample, after which the same operations run
  on the same values, so delaying the wavelet by k steps delays the
  seismogram by k steps.
- Batching: a shot's RTM image does not depend on the shots it shares
  launches with (batched == one shot per launch) on the full 600 x 6000 model.
- The sponge pair also matches the FP64 oracle over the first 100 steps of the
  benchmark shot.

Sizes: 2-D 1040 x 3080 (1000 x 3000 + 40-cell strips), nDiffOrder 5, 4000 steps;
3-D 420 x 440 x 440, nDiffOrder 4, 200 steps.
"""
import numpy as np
import pytest

import bench
import sssi_amd

pytestmark = pytest.mark.gpu


def test_2d_odd_symmetry_and_time_shift_at_benchmark_size():
    wl = bench.workload()
    args = (wl["V"], wl["nt"], wl["N"], wl["nB"], wl["dz"], wl["dx"], wl["dt"])
    w = wl["wavelet"]
    d, _ = sssi_amd.fwd_cpml_2d_points(*args, wl["src_iz"], wl["src_ix"], w)
    dn, _ = sssi_amd.fwd_cpml_2d_points(*args, wl["src_iz"], wl["src_ix"], -w)
    assert np.any(d != 0)
    assert np.array_equal(dn, -d)
    k = 37
    ws = np.concatenate([np.zeros(k), w[:-k]])
    dsh, _ = sssi_amd.fwd_cpml_2d_points(*args, wl["src_iz"], wl["src_ix"], ws)
    assert not np.any(dsh[:, :k])
    assert np.array_equal(dsh[:, k:], d[:, :-k])


def test_sponge_odd_symmetry_time_shift_and_oracle_at_benchmark_size():
    """The Cerjan-sponge pair on the 2-D benchmark model and shot: exact odd
    symmetry and time shift over 4000 steps, and the first 100 steps of the
    seismogram against the FP64 oracle."""
    import oracle
    wl = bench.workload()
    w = wl["wavelet"]
    args = (wl["V"], wl["nt"], wl["nB"], wl["dz"], wl["dx"], wl["dt"])
    d, _ = sssi_amd.fwd_sponge_2d_points(*args, wl["src_iz"], wl["src_ix"], w)
    dn, _ = sssi_amd.fwd_sponge_2d_points(*args, wl["src_iz"], wl["src_ix"], -w)
    assert np.any(d != 0)
    assert np.array_equal(dn, -d)
    k = 37
    ws = np.concatenate([np.zeros(k), w[:-k]])
    dsh, _ = sssi_amd.fwd_sponge_2d_points(*args, wl["src_iz"], wl["src_ix"], ws)
    assert not np.any(dsh[:, :k])
    assert np.array_equal(dsh[:, k:], d[:, :-k])
    n = 100
    d_ref, _ = oracle.fwd_sponge_2d(wl["V"], n, wl["nB"], wl["dz"], wl["dx"], wl["dt"],
                                    src_iz=wl["src_iz"], src_ix=wl["src_ix"], wavelet=w[:n])
    assert np.linalg.norm(d[:, :n] - d_ref) / np.linalg.norm(d_ref) < 1e-5


def test_3d_odd_symmetry_and_time_shift_at_benchmark_size():
    wl = bench.workload_3d()
    nt = 200
    args = (wl["V"], nt, wl["N"], wl["nB"], wl["d"], wl["d"], wl["d"], wl["dt"])
    iz, ix, iy = wl["src"]
    w = wl["wavelet"][:nt]
    d, _ = sssi_amd.fwd_cpml_3d_points(*args, [iz], [ix], [iy], w)
    dn, _ = sssi_amd.fwd_cpml_3d_points(*args, [iz], [ix], [iy], -w)
    assert np.any(d != 0)
    assert np.array_equal(dn, -d)
    k = 11
    ws = np.concatenate([np.zeros(k), w[:-k]])
    dsh, _ = sssi_amd.fwd_cpml_3d_points(*args, [iz], [ix], [iy], ws)
    assert not np.any(dsh[:, :, :k])
    assert np.array_equal(dsh[:, :, k:], d[:, :, :-k])


def test_rtm_batched_equals_single_on_full_model():
    wl = bench.workload_rtm(4)
    images = []
    for batched in (True, False):
        r = sssi_amd.Rtm2D(wl["V"], wl["VS"], wl["N"], wl["nB"], wl["d"], wl["d"], wl["dt"],
                           wl["nt"], img_stride=5, ck_format="bf16")
        for xb, xc, s in wl["shots"]:
            r.add_shot(xb, xc, s, wl["wavelet"])
        r.migrate(batched=batched)
        images.append((r.stack(), [r.shot(q)["image"] for q in range(len(wl["shots"]))]))
        r.close()
    assert np.any(images[0][0] != 0)
    assert np.array_equal(images[0][0], images[1][0])
    for a, b in zip(images[0][1], images[1][1]):
        assert np.array_equal(a, b)
