"""Benchmark: 2-D acoustic CPML forward modelling, single shot, on one MI355X.

Workload (FP32 on the GPU): 1000 x 3000 physical grid, dz = dx = 5 m, seeded
20-layer model (1500-5000 m/s) plus a smooth Gaussian random perturbation
(sigma 5 %, 50 m correlation), clipped to 1500-5000 m/s; nBoundary = 40 CPML
cells (extended grid 1040 x 3080); nDiffOrder = 5 (10th order in space);
dt = 0.5 ms; nt = 4000; 25 Hz Ricker at the surface centre; receivers at all
3080 surface columns.

One benchmark "step" is one whole propagation (4000 time steps) of that shot
with the model, damping profiles, source and wavefields already resident in
HBM.  `value` is Gcell-updates/s over the K timed steps.

Prints one JSON line with the metric, a roofline object for the time-step
kernel (algorithmic bytes per launch / average launch duration measured with
HIP events on the kernel's stream) and a CPU baseline (the FP64 oracle on a
bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def workload():
    from sssi_amd import synth
    nz0, nx0, nB, N = 1000, 3000, 40, 5
    dz = dx = 5.0
    dt = 5e-4
    nt = 4000
    v = synth.layered_2d(nz0, nx0, 20, 1500.0, 5000.0, seed=2)
    v = v * (1.0 + synth.gaussian_perturbation(v.shape, 0.05, 10.0, seed=3))
    v = np.clip(v, 1500.0, 5000.0)
    V = synth.extBoundary(v, nB, 2)
    w = synth.ricker(25.0, nt, dt)
    return dict(V=V, nB=nB, N=N, dz=dz, dx=dx, dt=dt, nt=nt, wavelet=w, src_iz=[0],
                src_ix=[nx0 // 2 + nB])


def traffic_from_profiles():
    """HBM bytes per launch of the time-step kernel, from the committed rocprof
    counter summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor)."""
    path = os.path.join(ROOT, "profiles", "cpml2d_step_bench_traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(wl, steps=60):
    import oracle
    threads = oracle.num_threads()
    t0 = time.perf_counter()
    oracle.fwd_cpml_2d(wl["V"], steps, wl["N"], wl["nB"], wl["dz"], wl["dx"], wl["dt"],
                       src_iz=wl["src_iz"], src_ix=wl["src_ix"], wavelet=wl["wavelet"][:steps])
    el = time.perf_counter() - t0
    cells = wl["V"].size
    return {"value": round(cells * steps / el / 1e9, 4), "unit": "Gcell-updates/s",
            "cores": threads, "kind": "port",
            "sample": f"FP64 oracle, same {wl['V'].shape[0]}x{wl['V'].shape[1]} grid and shot, "
                      f"first {steps} of {wl['nt']} time steps, {threads} OpenMP threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); multi-GPU runs are "
              "not part of this build.", file=sys.stderr)
        sys.exit(2)

    import sssi_amd
    wl = workload()
    V = wl["V"]
    nt = wl["nt"]
    P = sssi_amd.Propagator(V, wl["N"], wl["nB"], wl["dz"], wl["dx"], wl["dt"])
    P.set_points(wl["src_iz"], wl["src_ix"], wl["wavelet"])
    bytes_step, cells = P.alg_bytes_per_step()

    for _ in range(args.warmup):
        P.reset()
        P.run(nt)
    P.sync()

    kern_ms = 0.0
    launches = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        P.reset()
        ms, n = P.run_timed(nt)
        kern_ms += ms
        launches += n
    P.sync()
    wall = time.perf_counter() - t0
    data = P.data(nt)
    if not np.all(np.isfinite(data)):
        print("non-finite receiver data", file=sys.stderr)
        sys.exit(3)

    updates = cells * nt * args.steps
    value = updates / wall / 1e9
    avg_launch_s = kern_ms / 1e3 / launches
    achieved = bytes_step / avg_launch_s / 1e9
    out = {
        "metric": "Gcell-updates/s (2D acoustic CPML forward, FP32)",
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded layered model + Gaussian perturbation, Ricker source)",
        "config": {"workload": "2D CPML forward, 1040x3080 ext grid (1000x3000 + nB=40), "
                               "nDiffOrder=5, nt=4000, dt=0.5ms, 1 shot, 3080 receivers",
                   "nz": int(V.shape[0]), "nx": int(V.shape[1]), "nt": nt,
                   "nDiffOrder": wl["N"], "nBoundary": wl["nB"], "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles(),
                     "kernel": "cpml2d_step", "alg_bytes_per_launch": bytes_step,
                     "avg_launch_us": round(avg_launch_s * 1e6, 3), "launches": launches},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
