"""Seeded synthetic workloads: the benchmark cases and the Kirchhoff and Helmholtz ones.

Each builder returns plain numpy inputs (models indexed [iz, ix(, iy)], Ricker
wavelets, source positions) so that the same inputs drive the GPU path, the
tests and the benchmark.

  cpml2d_120x240()       2-D CPML forward, 100 x 200 + nB 20 (120 x 240), N 2, nt 1000,
                         dt 1 ms, 20 Hz Ricker at (0, 100 + nB), every-10th snapshots
  cpml2d_1040x3080()     2-D CPML forward, 1000 x 3000 + nB 40 (1040 x 3080), N 5,
                         nt 4000, dt 0.5 ms, 25 Hz Ricker at the surface centre
  rtm_600x6000(n)        2-D RTM, n shots (64) of width U[400, 1200] on a 600 x 6000
                         layered model and its 10-cell Gaussian blur, N 4, nB 30,
                         nt 3000, dt 1 ms, 15 Hz, imaging every 5 steps
  cpml3d_420x440x440()   3-D CPML forward, 400^3 + nB 20 (420 x 440 x 440), N 4,
                         nt 2000, dt 1 ms, 15 Hz Ricker at the surface centre
  serving_mix(n)         serving mix: 70 % 2-D forward / RTM-shot jobs (200..2000 x
                         200..4000, N 2..5, nt 500..3000), 30 % 3-D forward (128^3 ..
                         256^3, nt 500..1500); layered 1500-4500 m/s; 10-30 Hz Ricker
  kirchhoff(n)   Kirchhoff migration in the geometry of the reference's
                 Kirchhoff app (shots and receivers at every surface column,
                 tables for every position): 300 x 600, d 10 m, velocity
                 increasing with depth 1800-4000 m/s (10-cell blur), the app's
                 dt and nt, iMax (nz + nx) / 2; n shots (16) spread over the
                 surface with seeded synthetic records
  freq_fwi()     frequency-domain CPML Helmholtz in the geometry of the
                 reference's frequency FWI app: 100 x 100 model + nB 20
                 (120 x 140 squared slowness), N 2, d 10 m, 100 shots (one per
                 surface column of the physical model, unit source), 10 Hz
"""
import numpy as np

from . import synth


def cpml2d_120x240():
    nz0, nx0, nB, N = 100, 200, 20, 2
    d, dt, nt = 10.0, 1e-3, 1000
    v = synth.layered_2d(nz0, nx0, 5, 1500.0, 4500.0, seed=1)
    V = synth.extBoundary(v, nB, 2)
    w = synth.ricker(20.0, nt, dt)
    return dict(V=V, nB=nB, N=N, dz=d, dx=d, dt=dt, nt=nt, wavelet=w, src_iz=[0],
                src_ix=[nx0 // 2 + nB], snap_stride=10)


def cpml2d_1040x3080():
    nz0, nx0, nB, N = 1000, 3000, 40, 5
    dz = dx = 5.0
    dt, nt = 5e-4, 4000
    v = synth.layered_2d(nz0, nx0, 20, 1500.0, 5000.0, seed=2)
    v = v * (1.0 + synth.gaussian_perturbation(v.shape, 0.05, 10.0, seed=3))
    v = np.clip(v, 1500.0, 5000.0)
    V = synth.extBoundary(v, nB, 2)
    w = synth.ricker(25.0, nt, dt)
    return dict(V=V, nB=nB, N=N, dz=dz, dx=dx, dt=dt, nt=nt, wavelet=w, src_iz=[0],
                src_ix=[nx0 // 2 + nB])


def rtm_600x6000(nshots=64, seed=5):
    nz, nx, nB, N = 600, 6000, 30, 4
    d, dt, nt = 10.0, 1e-3, 3000
    V = synth.layered_2d(nz, nx, 12, 1500.0, 4500.0, seed=seed)
    VS = synth.smooth_copy(V, 10)
    rng = np.random.default_rng(seed + 1)
    w = synth.ricker(15.0, nt, dt)
    shots = []
    for _ in range(nshots):
        xc = int(rng.integers(400, 1201))
        xb = int(rng.integers(0, nx - xc + 1))
        shots.append((xb, xc, xc // 2))
    return dict(V=V, VS=VS, nB=nB, N=N, d=d, dt=dt, nt=nt, wavelet=w, shots=shots,
                img_stride=5)


def cpml3d_420x440x440():
    n0, nB, N = 400, 20, 4
    d, dt, nt = 10.0, 1e-3, 2000
    v = synth.cake_3d_dipping(n0, n0, n0, 8, 1500.0, 4500.0, seed=4)
    # the 8th-order 3-D scheme at dt = 1 ms, d = 10 m is stable below 4490 m/s
    v = np.clip(v, 1500.0, 4450.0)
    V = synth.extBoundary(v, nB, 3)
    w = synth.ricker(15.0, nt, dt)
    c = (n0 + 2 * nB) // 2
    return dict(V=V, nB=nB, N=N, d=d, dt=dt, nt=nt, wavelet=w, src=(0, c, c))


def serving_mix(njobs=60, seed=7):
    """Job list [(kind, kwargs)] for sssi_amd.Server.submit_<kind>.  RTM-shot
    jobs carry an extBoundary-extended model (nB = 20) so that each one is also
    a one-shot Rtm2D migration of the physical model over its whole width."""
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(njobs):
        f = float(rng.uniform(10, 30))
        if rng.random() < 0.7:
            nz, nx = int(rng.integers(200, 2001)), int(rng.integers(200, 4001))
            N = int(rng.integers(2, 6))
            nt = int(rng.integers(500, 3001))
            v = synth.layered_2d(nz, nx, 6, 1500.0, 4500.0, seed=seed * 1000 + k)
            nB = 20
            dt = 0.5 * 10.0 / (4500.0 * np.sqrt(2) * float(np.sum(np.abs(synth.dCoef(N)))))
            w = synth.ricker(f, nt, dt)
            if rng.random() < 0.5:
                jobs.append(("fwd2d", dict(v=v, nt=nt, nDiffOrder=N, nBoundary=nB, dz=10.0,
                                           dx=10.0, dt=dt, src_iz=0, src_ix=nx // 2, wavelet=w)))
            else:
                pz, px = nz - nB, nx - 2 * nB
                vp = v[:pz, :px]
                jobs.append(("rtm2d", dict(V=synth.extBoundary(vp, nB, 2),
                                           VS=synth.extBoundary(synth.smooth_copy(vp, 5), nB, 2),
                                           nt=nt, nDiffOrder=N, nBoundary=nB, dz=10.0, dx=10.0,
                                           dt=dt, src_ix=nB + px // 2, wavelet=w, img_stride=5,
                                           ck_format="bf16")))
        else:
            n = int(rng.integers(128, 257))
            nt = int(rng.integers(500, 1501))
            N = int(rng.integers(2, 6))
            v = synth.cake_3d_dipping(n, n, n, 4, 1500.0, 4500.0, seed=seed * 1000 + k)
            dt = 0.5 * 10.0 / (4500.0 * np.sqrt(3) * float(np.sum(np.abs(synth.dCoef(N)))))
            jobs.append(("fwd3d", dict(v=v, nt=nt, nDiffOrder=N, nBoundary=16, dz=10.0, dx=10.0,
                                       dy=10.0, dt=dt, src=(0, n // 2, n // 2),
                                       wavelet=synth.ricker(f, nt, dt))))
    return jobs


def job_cells(kind, kw):
    """Cell updates of a serving-mix job (RTM shots propagate three times)."""
    if kind == "fwd3d":
        return int(np.prod(kw["v"].shape)) * kw["nt"]
    v = kw["v"] if kind == "fwd2d" else kw["V"]
    return int(np.prod(v.shape)) * kw["nt"] * (3 if kind == "rtm2d" else 1)


def kirchhoff(nshots=16, seed=31):
    """The app's rules (applications/kirchhoff_2d_acoustic/
    mainKirchhoffCpmlFor2dAw.m:51-85, :111, :115): xShotGrid = xRecGrid =
    1..nx, xShotAndRecGrid their union, dt = 0.75 dz / vmax / sqrt(2),
    nt = round(sqrt((dx nx)^2 + (dz nz)^2) * 2 / vmin / dt + 1)."""
    nz, nx, d = 300, 600, 10.0
    vmin, vmax = 1800.0, 4000.0
    v = np.sort(synth.layered_2d(nz, nx, 10, vmin, vmax, seed=seed), axis=0)
    V = np.clip(synth.smooth_copy(v, 10), vmin, vmax)
    dt = 0.75 * (d / vmax / np.sqrt(2))
    nt = int(round(np.sqrt((d * nx) ** 2 + (d * nz) ** 2) * 2 / vmin / dt + 1))
    xsr = np.arange(1, nx + 1)
    xShot = np.unique(np.round(np.linspace(1, nx, nshots)).astype(int))
    rng = np.random.default_rng(seed + 1)
    shots = np.asfortranarray(
        rng.standard_normal((nt, nx, xShot.size), dtype=np.float32).astype(np.float64))
    return dict(V=V, d=d, dt=dt, nt=nt, iMax=(nz + nx) // 2, xsr=xsr, xRec=xsr, xShot=xShot,
                ixs=xShot, shots=shots)


def freq_fwi(f=10.0, seed=41):
    """applications/fwi_2d_acoustic/mainFwiFreqCpml2dAw.m:93-127, :211-221:
    nBoundary 20 (model extended down and sideways), nDiffOrder 2, one source
    per physical surface column, solved per frequency."""
    nz0, nx0, nB, N, d = 100, 100, 20, 2, 10.0
    v = synth.layered_2d(nz0, nx0, 6, 1500.0, 4500.0, seed=seed)
    # extended down and sideways as the app's model (nz + nB, nx + 2 nB)
    ve = np.pad(v, ((0, nB), (nB, nB)), mode="edge")
    m = 1.0 / ve ** 2
    nz, nx = m.shape
    xs = np.arange(nB, nB + nx0)
    S = np.zeros((nz * nx, xs.size), dtype=np.complex128)
    S[xs * nz + 0, np.arange(xs.size)] = 1.0
    return dict(m=m, S=S, w=2 * np.pi * f, N=N, nB=nB, d=d)
