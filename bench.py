"""Benchmark: the acoustic CPML propagators and RTM on one MI355X.

Headline (the default line's top-level fields): 2-D CPML forward modelling,
single shot.  1000 x 3000 physical grid, dz = dx = 5 m, seeded 20-layer model
(1500-5000 m/s) plus a smooth Gaussian random perturbation (sigma 5 %, 50 m
correlation), clipped to 1500-5000 m/s; nBoundary = 40 CPML cells (extended
grid 1040 x 3080); nDiffOrder = 5 (10th order in space); dt = 0.5 ms;
nt = 4000; 25 Hz Ricker at the surface centre; receivers at all 3080 surface
columns.  One benchmark "step" is one whole propagation (4000 time steps) of
that shot with the model, damping profiles, source and wavefields already
resident in HBM.  `value` is Gcell-updates/s over the K timed steps.  The line
carries a roofline object for the time-step kernel (algorithmic bytes per
launch / average launch duration measured with HIP events on the kernel's
stream) and a CPU baseline (the FP64 oracle, one shot per host core, on a
bounded sample of the same workload).

Sub-objects of the same line, each with its own fixed step counts:
  cfg4_3d     3-D CPML forward, 420 x 440 x 440, nDiffOrder 4, nt 2000
              (2 timed passes, roofline from its own event timing)
  cfg3_rtm    64-shot RTM: plain FP32 one shot per launch, batched FP32,
              batched 16-bit checkpoints, served (shots/hour)
  cfg5_serve  the 60-job Poisson serving mix: p50/p99 latency and throughput,
              continuous batching vs one job at a time
  kirchhoff   Kirchhoff path in the reference app's geometry: eikonal
              traveltime tables for every surface position (Gcell-updates/s,
              roofline of the pass kernel) and migrated shots/hour
  freq        frequency-domain CPML Helmholtz, one frequency of the FWI app's
              geometry (all shots): factorisation + solve
`--workload` runs one of them alone (and the sponge / copy calibrations).
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
    from sssi_amd import workloads
    return workloads.cpml2d_1040x3080()


def traffic_from_profiles(name="cpml2d_step_bench_traffic.json"):
    """HBM bytes per launch of the time-step kernel, from the committed rocprof
    counter summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor)."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(wl, steps=None, cores=None):
    """The FP64 oracle on the host cores, one shot per core: each core runs its
    own copy of the workload's shot single-threaded for the first `steps` time
    steps (about 10 s of CPU work), all cores at once; the value is the
    aggregate cell-update rate."""
    import threading
    import oracle
    cores = cores or min(16, os.cpu_count() or 1)
    nz, nx = wl["V"].shape
    if steps is None:
        steps = max(5, int(10.0 * 6e7 / (nz * nx)))  # ~10 s at ~0.06 Gcell/s per core
    steps = min(steps, wl["nt"])
    errors = []

    def shot():
        try:
            oracle.set_threads(1)
            oracle.fwd_cpml_2d(wl["V"], steps, wl["N"], wl["nB"], wl["dz"], wl["dx"], wl["dt"],
                               src_iz=wl["src_iz"], src_ix=wl["src_ix"],
                               wavelet=wl["wavelet"][:steps])
        except Exception as e:  # reported below
            errors.append(e)

    threads = [threading.Thread(target=shot) for _ in range(cores)]
    t0 = time.perf_counter()
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    el = time.perf_counter() - t0
    if errors:
        raise errors[0]
    return {"value": round(cores * nz * nx * steps / el / 1e9, 4), "unit": "Gcell-updates/s",
            "cores": cores, "kind": "port",
            "sample": f"FP64 oracle, one shot per core on {cores} host cores at once, each the "
                      f"same {nz}x{nx} grid and shot for the first {steps} of {wl['nt']} time "
                      f"steps, single-threaded ({el:.1f} s)"}


def workload_3d():
    from sssi_amd import workloads
    return workloads.cpml3d_420x440x440()


def bench_3d(args):
    """3-D CPML forward, 420 x 440 x 440 ext grid, nDiffOrder = 4, nt = 2000."""
    import sssi_amd
    wl = workload_3d()
    V = wl["V"]
    nt = wl["nt"] if args.nt is None else args.nt
    sssi_amd.set_autotune(not args.no_tune)
    P = sssi_amd.Propagator(V, wl["N"], wl["nB"], wl["d"], wl["d"], wl["dt"], dy=wl["d"])
    ty, zc = P.layout()
    iz, ix, iy = wl["src"]
    P.set_points([iz], [ix], wl["wavelet"], src_iy=[iy])
    bytes_step, cells = P.alg_bytes_per_step()
    for _ in range(args.warmup):
        P.reset()
        P.run(min(nt, 50))
    P.sync()
    kern_ms, launches = 0.0, 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        P.reset()
        ms, n = P.run_timed(nt)
        kern_ms += ms
        launches += n
    wall = time.perf_counter() - t0
    P.close()
    avg = kern_ms / 1e3 / launches
    ach = bytes_step / avg / 1e9
    return {
        "metric": "Gcell-updates/s (3D acoustic CPML forward, FP32)",
        "value": round(cells * nt * args.steps / wall / 1e9, 3), "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded 8-layer cake with a dipping interface, Ricker source)",
        "config": {"workload": f"3D CPML forward, {V.shape[0]}x{V.shape[1]}x{V.shape[2]} ext "
                               f"grid, nDiffOrder=4, nt={nt}, dt=1ms, 1 shot",
                   "parallelism": "1 GPU",
                   "layout": {"block_y": ty, "z_chunk": zc, "autotuned": not args.no_tune}},
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles("cpml3d_step_bench_traffic.json"),
                     "kernel": "cpml3d_step", "alg_bytes_per_launch": bytes_step,
                     "avg_launch_us": round(avg * 1e6, 3), "launches": launches},
    }


def bench_sponge(args):
    """2-D Cerjan-sponge forward on the headline 2-D model and shot (1040 x 3080 ext grid,
    nBoundary = 40, nt = 4000)."""
    import sssi_amd
    wl = workload()
    V = wl["V"]
    nt = wl["nt"] if args.nt is None else args.nt
    P = sssi_amd.Propagator(V, 1, wl["nB"], wl["dz"], wl["dx"], wl["dt"], abc="sponge")
    P.set_points(wl["src_iz"], wl["src_ix"], wl["wavelet"])
    bytes_step, cells = P.alg_bytes_per_step()
    for _ in range(args.warmup):
        P.reset()
        P.run(nt)
    P.sync()
    kern_ms, launches = 0.0, 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        P.reset()
        ms, n = P.run_timed(nt)
        kern_ms += ms
        launches += n
    wall = time.perf_counter() - t0
    avg = kern_ms / 1e3 / launches
    ach = bytes_step / avg / 1e9
    return {
        "metric": "Gcell-updates/s (2D acoustic sponge forward, FP32)",
        "value": round(cells * nt * args.steps / wall / 1e9, 3), "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded layered model + Gaussian perturbation, Ricker source)",
        "config": {"workload": f"2D sponge forward, {V.shape[0]}x{V.shape[1]} ext grid, "
                               f"nBoundary={wl['nB']}, nt={nt}, 1 shot", "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles("sponge2d_step_bench_traffic.json"),
                     "kernel": "sponge2d_step", "alg_bytes_per_launch": bytes_step,
                     "avg_launch_us": round(avg * 1e6, 3), "launches": launches},
    }


def workload_rtm(nshots=64, seed=5):
    from sssi_amd import workloads
    return workloads.rtm_600x6000(nshots, seed)


def bench_rtm(args):
    """RTM: 64 shots of uneven aperture on the 600 x 6000 model.

    Modes: plain (one shot per launch, FP32 checkpoints); batched (one launch per
    time step covers every shot) with FP32 and with 16-bit scaled checkpoints,
    which separates the two effects; served (every shot a job of the serving
    loop, 16-bit checkpoints).  dataTrue = fwd(V) is cached per shot as the
    reference's apps cache it (mainRtmTimeCpmlFor2dAw.m:170-184): shots/hour
    leaves it out (the served jobs load the cache), and is also given with it
    computed inside the timed region (plain: precompute + migrate; served: jobs
    that compute their own).  The plain path is also timed end to end per shot
    (a handle per shot window, as each served job has: setup, dataTrue,
    migration, image to the host) on 16 of the shots."""
    import sssi_amd
    from sssi_amd import synth
    wl = workload_rtm(args.shots)
    nshot = len(wl["shots"])
    res = {}
    data_true = []
    for name, batched, fmt in (("plain_fp32", False, "fp32"), ("batched_fp32", True, "fp32"),
                               ("batched_" + args.ck, True, args.ck)):
        r = sssi_amd.Rtm2D(wl["V"], wl["VS"], wl["N"], wl["nB"], wl["d"], wl["d"], wl["dt"],
                           wl["nt"], img_stride=wl["img_stride"], ck_format=fmt)
        for xb, xc, sx in wl["shots"]:
            r.add_shot(xb, xc, sx, wl["wavelet"])
        t0 = time.perf_counter()
        r.precompute_data(batched=batched)
        t_true = time.perf_counter() - t0
        t0 = time.perf_counter()
        ms, launches = r.migrate(batched=batched)
        wall = time.perf_counter() - t0
        res[name] = {"shots_per_hour": round(nshot / wall * 3600, 1),
                     "shots_per_hour_with_data_true": round(nshot / (wall + t_true) * 3600, 1),
                     "seconds": round(wall, 3), "data_true_seconds": round(t_true, 3),
                     "event_ms": round(ms, 2), "launches": launches,
                     "device_GB": round(r.device_bytes() / 1e9, 2)}
        if not data_true:  # the cache the served shots load, as the apps load theirs
            data_true = [np.asfortranarray(r.shot(q)["data_true"]) for q in range(nshot)]
        r.close()
    # served: each shot a job of the serving loop (extended window models), all
    # submitted at once; with the cached dataTrue, and computing it
    nB = wl["nB"]
    jobs = []
    for xb, xc, sx in wl["shots"]:
        # column-major, as a MATLAB caller holds them (no conversion in the timed region)
        jobs.append(dict(V=np.asfortranarray(synth.extBoundary(wl["V"][:, xb:xb + xc], nB, 2)),
                         VS=np.asfortranarray(synth.extBoundary(wl["VS"][:, xb:xb + xc], nB, 2)),
                         nt=wl["nt"], nDiffOrder=wl["N"], nBoundary=nB, dz=wl["d"], dx=wl["d"],
                         dt=wl["dt"], src_ix=sx + nB, wavelet=wl["wavelet"],
                         img_stride=wl["img_stride"], ck_format=args.ck))
    for name, cached in (("served_" + args.ck, True), ("served_" + args.ck + "_computing_data_true",
                                                        False)):
        with sssi_amd.Server(hbm_budget_bytes=args.budget_gb * 1e9, max_active=nshot) as s:
            t0 = time.perf_counter()
            ids = [s.submit_rtm2d(**kw, data_true=dt if cached else None)
                   for kw, dt in zip(jobs, data_true)]
            s.drain()
            wall = time.perf_counter() - t0
            info = [s.info(i) for i in ids]
        lat = np.array([f["latency"] for f in info])
        t_a = min(f["t_arrive"] for f in info)
        key = "shots_per_hour" if cached else "shots_per_hour_with_data_true"
        res[name] = {key: round(nshot / wall * 3600, 1), "seconds": round(wall, 3),
                     "p50_s": round(float(np.percentile(lat, 50)), 3),
                     "p99_s": round(float(np.percentile(lat, 99)), 3),
                     "last_arrival_s": round(max(f["t_arrive"] for f in info) - t_a, 3),
                     "last_admission_s": round(max(f["t_admit"] for f in info) - t_a, 3),
                     "first_done_s": round(min(f["t_done"] for f in info) - t_a, 3),
                     "last_done_s": round(max(f["t_done"] for f in info) - t_a, 3)}
    # the plain path end to end per shot, as the served jobs are: a handle per
    # shot on its own window (model upload and conversion, state allocation),
    # dataTrue, migration, the image back on the host; 16 of the 64 shots
    nsub = min(16, nshot)
    t0 = time.perf_counter()
    for xb, xc, sx in wl["shots"][:nsub]:
        r1 = sssi_amd.Rtm2D(np.asfortranarray(wl["V"][:, xb:xb + xc]),
                            np.asfortranarray(wl["VS"][:, xb:xb + xc]), wl["N"], wl["nB"],
                            wl["d"], wl["d"], wl["dt"], wl["nt"], img_stride=wl["img_stride"],
                            ck_format="fp32")
        r1.add_shot(0, xc, sx, wl["wavelet"])
        r1.precompute_data(batched=False)
        r1.migrate(batched=False)
        r1.stack()
        r1.close()
    e2e = time.perf_counter() - t0
    res["plain_fp32_end_to_end"] = {
        "shots_per_hour_with_data_true": round(nsub / e2e * 3600, 1), "seconds": round(e2e, 3),
        "shots": nsub}
    sv = res["served_" + args.ck]["shots_per_hour"]
    b = res["batched_" + args.ck]["shots_per_hour"]
    p = res["plain_fp32"]["shots_per_hour"]
    svc = res["served_" + args.ck + "_computing_data_true"]["shots_per_hour_with_data_true"]
    pw = res["plain_fp32"]["shots_per_hour_with_data_true"]
    return {
        "metric": "RTM shots/hour (2D CPML, 64 shots, served + batched + 16-bit checkpoints)",
        "value": sv, "unit": "shots/hour", "n_gpus": 1, "steps": 1, "warmup": 0,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded 600x6000 layered model, 10-cell Gaussian-blurred copy)",
        "config": {"workload": "2D RTM, 64 shots, nz=600, nx~U[400,1200], nB=30, "
                               "nDiffOrder=4, nt=3000, imaging every 5 steps, dataTrue cached",
                   "parallelism": "1 GPU"},
        "speedup_vs_plain_fp32": round(sv / p, 3),
        "batched_speedup_vs_plain_fp32": round(b / p, 3),
        "speedup_batching_alone": round(res["batched_fp32"]["shots_per_hour"] / p, 3),
        "served_speedup_vs_plain_fp32_both_computing_data_true": round(svc / pw, 3),
        "served_speedup_vs_plain_fp32_end_to_end": round(
            svc / res["plain_fp32_end_to_end"]["shots_per_hour_with_data_true"], 3),
        "modes": res,
    }


def serve_jobs(njobs, seed=7):
    from sssi_amd import workloads
    return workloads.serving_mix(njobs, seed)


def bench_serve(args):
    """Poisson arrivals of the seeded job mix at a fixed rate; p50/p99 latency and
    throughput with continuous batching vs one job at a time."""
    import threading
    import sssi_amd
    jobs = serve_jobs(args.jobs)
    rng = np.random.default_rng(11)
    gaps = rng.exponential(1.0 / args.rate, size=len(jobs))
    out = {}
    for name, max_active in (("batched", 64), ("one_at_a_time", 1)):
        with sssi_amd.Server(hbm_budget_bytes=args.budget_gb * 1e9, max_active=max_active) as s:
            ids = []
            t0 = time.perf_counter()

            def feeder():
                t = t0
                for (kind, kw), g in zip(jobs, gaps):
                    t += g
                    d = t - time.perf_counter()
                    if d > 0:
                        time.sleep(d)
                    ids.append(getattr(s, "submit_" + kind)(**kw))

            th = threading.Thread(target=feeder)
            th.start()
            th.join()
            s.drain()
            wall = time.perf_counter() - t0
            lat = np.array([s.info(i)["latency"] for i in ids])
            st = s.stats()
        out[name] = {"p50_s": round(float(np.percentile(lat, 50)), 4),
                     "p99_s": round(float(np.percentile(lat, 99)), 4),
                     "jobs_per_s": round(len(jobs) / wall, 3),
                     "gcell_updates_per_s": round(st["cell_updates"] / wall / 1e9, 2),
                     "wall_s": round(wall, 3), "launches": st["launches"], "ticks": st["ticks"]}
    b, p = out["batched"], out["one_at_a_time"]
    return {
        "metric": "serving p50/p99 latency and throughput (mixed 2D/3D forward + RTM jobs)",
        "value": b["jobs_per_s"], "unit": "jobs/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded layered models, Ricker sources, Poisson arrivals)",
        "config": {"workload": f"{len(jobs)} jobs, Poisson rate {args.rate}/s, 70% 2D fwd/RTM "
                               "200x200..2000x4000, 30% 3D fwd 128^3..256^3",
                   "parallelism": "1 GPU"},
        "p50_s": b["p50_s"], "p99_s": b["p99_s"],
        "throughput_vs_one_at_a_time": round(b["jobs_per_s"] / p["jobs_per_s"], 3),
        "modes": out,
    }


def bench_copy(args):
    """Device-to-device copy of a known size (hipMemcpy's blit kernel), the
    calibration run for the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = 256 << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 0, ctypes.c_size_t(n))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    wall = time.perf_counter() - t0
    hip.hipFree(a)
    hip.hipFree(b)
    # the achievable streaming rate for a stencil step's minimal access pattern
    # (3 FP32 reads + 1 write per cell) at the 2-D headline grid's size (MALL
    # resident), the 3-D grid's and a 4x larger one
    import sssi_amd
    from sssi_amd._lib import _call
    stream = {}
    for label, cells in (("2d_1040x3200", 1040 * 3200), ("3d_420x440x440", 420 * 440 * 440),
                         ("4x_3d", 4 * 420 * 440 * 440)):
        us, gbs = ctypes.c_double(), ctypes.c_double()
        sssi_amd.lib()
        _call("sssi_calib_stream", cells - cells % 4, 50, ctypes.byref(us), ctypes.byref(gbs))
        stream[label] = {"cells": cells, "us": round(us.value, 2), "GB_per_s": round(gbs.value, 1),
                         "frac_of_8TBs": round(gbs.value / HBM_PEAK_GBS, 4)}
    return {"metric": "device copy GB/s (calibration)", "value": round(2 * n * args.steps / wall / 1e9, 1),
            "unit": "GB/s", "bytes_read_per_copy": n, "bytes_written_per_copy": n,
            "steps": args.steps, "stream_triad": stream}


def bench_kirchhoff(args):
    import oracle
    import sssi_amd
    from sssi_amd import workloads
    wl = workloads.kirchhoff()
    V, shots = wl["V"], wl["shots"]
    nz, nx = V.shape
    ntab = wl["xsr"].size
    nshot = shots.shape[2]
    t0 = time.perf_counter()
    k = sssi_amd.Kirchhoff2D(V, wl["d"], wl["xsr"], wl["iMax"], wl["dt"])
    tables_wall = time.perf_counter() - t0
    eik_ms, passes, launches, _ = k.timing()
    t0 = time.perf_counter()
    k.load_shots(shots)  # records resident, as forward-modelled data on the device would be
    load_s = time.perf_counter() - t0
    st = k.migrate_loaded(wl["xRec"], wl["ixs"])  # warm-up
    steps = max(1, args.steps)
    mig_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        st = k.migrate_loaded(wl["xRec"], wl["ixs"])
        mig_ms += k.timing()[3]
    mig_wall = time.perf_counter() - t0
    tt = None if args.no_cpu else k.travel_time()
    k.close()
    if not np.all(np.isfinite(st)):
        raise RuntimeError("non-finite stack")
    cell_updates = nz * nx * ntab * passes
    # per launch (H passes on overlapping tiles): each table read and written once, slowness
    launch_bytes = nz * nx * (16 * ntab + 8)
    avg_launch_s = eik_ms / 1e3 / launches
    sums = nz * nx * wl["xRec"].size * nshot  # trace samples summed per migrate call
    out = {
        "metric": "Kirchhoff migrated shots/hour (tables and records resident in HBM)",
        "value": round(nshot * steps / mig_wall * 3600, 1),
        "unit": "shots/h",
        "kernel_shots_per_hour": round(nshot * steps / (mig_ms / 1e3) * 3600, 1),
        "records_upload_s": round(load_s, 3),
        "migrate_gsums_per_s": round(sums * steps / (mig_ms / 1e3) / 1e9, 2),
        "steps": steps,
        "dtype": "f64",
        "config": {"workload": f"Kirchhoff 2D, {nz}x{nx}, d=10m, {ntab} tables (every surface "
                               f"position), {wl['xRec'].size} receivers, {nshot} shots per "
                               f"migrate call, nt={wl['nt']}, iMax={wl['iMax']}",
                   "nz": nz, "nx": nx, "tables": ntab, "receivers": int(wl["xRec"].size),
                   "shots": nshot, "nt": wl["nt"], "iMax": wl["iMax"]},
        "eikonal": {"tables_wall_s": round(tables_wall, 3), "kernel_s": round(eik_ms / 1e3, 3),
                    "passes": passes, "gcell_updates_per_s": round(cell_updates / (eik_ms / 1e3)
                                                                   / 1e9, 2)},
        "roofline": {"bound": "hbm", "achieved": round(launch_bytes / avg_launch_s / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(launch_bytes / avg_launch_s / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": "eik_tile<4,32>",
                     "alg_bytes_per_launch": launch_bytes,
                     "avg_launch_us": round(avg_launch_s * 1e6, 2), "launches": launches,
                     "note": "4 passes per launch on 64x32 tiles: the kernel is bound by FP64 "
                             "VALU issue (the update with a correctly rounded sqrt, plus halo "
                             "recompute), not by HBM"},
    }
    if not args.no_cpu:
        cores = min(16, os.cpu_count() or 1)
        oracle.set_threads(cores)
        n_tab = 12  # ~5 s: full-length tables of 12 surface positions
        t0 = time.perf_counter()
        for x in np.linspace(1, nx, n_tab).round().astype(int):
            oracle.eikonal2d(V, wl["d"], 1, int(x), wl["iMax"])
        e_el = time.perf_counter() - t0
        nmig = 6 * nshot  # ~5 s: the loaded shots six times over
        t0 = time.perf_counter()
        for s in range(nmig):
            oracle.kirchhoff_migrate(tt, wl["xRec"], wl["xsr"], shots[:, :, s % nshot], wl["dt"],
                                     nz, int(wl["ixs"][s % nshot]))
        m_el = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": round(nmig / m_el * 3600, 1), "unit": "shots/h", "cores": cores,
            "kind": "port",
            "eikonal_gcell_updates_per_s": round(nz * nx * n_tab * passes / e_el / 1e9, 4),
            "sample": f"FP64 oracle on {cores} OpenMP threads: migration of {nmig} shots "
                      f"against {wl['xRec'].size} receivers ({m_el:.1f} s); eikonal tables of "
                      f"{n_tab} surface positions, {passes} passes each ({e_el:.1f} s)"}
    return out


def bench_freq(args):
    """Frequency-domain CPML Helmholtz: one frequency of the FWI app's
    geometry (120 x 140, N 2, 100 shots), factorisation + solve on the device."""
    import oracle
    import sssi_amd
    from sssi_amd import helmholtz, workloads
    wl = workloads.freq_fwi()
    m, S = wl["m"], wl["S"]
    nz, nx = m.shape
    args_ = (m, S, wl["w"], wl["N"], wl["nB"], wl["d"], wl["d"])
    helmholtz.solve(*args_)  # warm-up
    steps = max(1, args.steps)
    f_ms = s_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        U = helmholtz.solve(*args_)
        a, b = helmholtz.last_timing()
        f_ms += a
        s_ms += b
    wall = time.perf_counter() - t0
    if not np.all(np.isfinite(U)):
        raise RuntimeError("non-finite solution")
    kl = (2 * wl["N"] - 1) * nz
    out = {
        "metric": "frequency-domain CPML solves per second (all shots of one frequency)",
        "value": round(steps / wall, 3), "unit": "frequencies/s",
        "ms_per_frequency": round(wall / steps * 1e3, 2),
        "factor_ms": round(f_ms / steps, 2), "solve_ms": round(s_ms / steps, 2),
        "steps": steps, "dtype": "c128",
        "config": {"workload": f"freqCpmlFor2dAw, {nz}x{nx} (100x100 + nB 20), nDiffOrder 2, "
                               f"{S.shape[1]} shots, 10 Hz; banded LU, bandwidth {kl}",
                   "nz": nz, "nx": nx, "shots": int(S.shape[1]), "bandwidth": kl},
    }
    if not args.no_cpu:
        cores = min(16, os.cpu_count() or 1)
        oracle.set_threads(cores)
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            oracle.freq_cpml_2d_solve(*args_)
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": round(reps / el, 3), "unit": "frequencies/s", "cores": cores, "kind": "port",
            "sample": f"FP64 oracle (same banded LU, OpenMP over the update's columns and the "
                      f"right-hand sides) on {cores} threads, {reps} frequencies ({el:.1f} s)"}
    return out


def _sub(fn, args, **over):
    """Run one sub-benchmark with its own fixed step counts; a failure is
    recorded instead of ending the run."""
    a = argparse.Namespace(**vars(args))
    for k, v in over.items():
        setattr(a, k, v)
    t0 = time.perf_counter()
    try:
        out = fn(a)
    except Exception as e:  # noqa: BLE001 -- reported in the JSON line
        out = {"error": f"{type(e).__name__}: {e}"}
    out["bench_wall_s"] = round(time.perf_counter() - t0, 1)
    print(f"[bench] {fn.__name__}: {out.get('value')} {out.get('unit', '')} "
          f"({out['bench_wall_s']} s)", file=sys.stderr, flush=True)
    return out


def bench_2d(args):
    import sssi_amd
    wl = workload()
    V = wl["V"]
    nt = wl["nt"]
    sssi_amd.set_autotune(not args.no_tune)
    P = sssi_amd.Propagator(V, wl["N"], wl["nB"], wl["dz"], wl["dx"], wl["dt"])
    rz, ng = P.layout()
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
    P.close()
    if not np.all(np.isfinite(data)):
        raise RuntimeError("non-finite receiver data")

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
                   "nDiffOrder": wl["N"], "nBoundary": wl["nB"], "parallelism": "1 GPU",
                   "layout": {"tile_rows": rz, "march_groups": ng, "autotuned": not args.no_tune,
                              "graph": os.environ.get("SSSI_GRAPH", "0") != "0"}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles(),
                     "kernel": "cpml2d_single", "alg_bytes_per_launch": bytes_step,
                     "avg_launch_us": round(avg_launch_s * 1e6, 3), "launches": launches},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-sub", action="store_true",
                    help="default workload: the headline 2-D line only, without the config "
                         "3 / 4 / 5 sub-benchmarks")
    ap.add_argument("--no-tune", action="store_true",
                    help="2d/3d: default kernel layout instead of the autotuned one")
    ap.add_argument("--workload", default="2d",
                    choices=["2d", "3d", "sponge", "rtm", "serve", "copy", "kirchhoff", "freq"],
                    help="2d: the headline 2-D forward line with the sub-benchmarks (default); "
                         "the others: one workload alone")
    ap.add_argument("--nt", type=int, default=None,
                    help="3d / sponge: time steps (default 2000 / 4000)")
    ap.add_argument("--shots", type=int, default=64, help="rtm: number of shots")
    ap.add_argument("--ck", default="bf16", choices=["fp16", "bf16"], help="rtm: checkpoint type")
    ap.add_argument("--jobs", type=int, default=60, help="serve: number of jobs")
    ap.add_argument("--rate", type=float, default=20.0, help="serve: Poisson arrival rate (1/s)")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="serve: HBM budget (GB)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); multi-GPU runs are "
              "not part of this build.", file=sys.stderr)
        sys.exit(2)
    one = {"3d": bench_3d, "sponge": bench_sponge, "rtm": bench_rtm, "serve": bench_serve,
           "copy": bench_copy, "kirchhoff": bench_kirchhoff, "freq": bench_freq}
    if args.workload in one:
        print(json.dumps(one[args.workload](args)), flush=True)
        return
    out = bench_2d(args)
    print(f"[bench] headline: {out['value']} Gcell-updates/s", file=sys.stderr, flush=True)
    if not args.no_sub:
        # configs 4, 3 and 5, each with its own fixed step counts
        out["cfg4_3d"] = _sub(bench_3d, args, steps=2, warmup=1, nt=None)
        out["cfg3_rtm"] = _sub(bench_rtm, args, shots=64)
        out["cfg5_serve"] = _sub(bench_serve, args, jobs=60, rate=20.0)
        out["kirchhoff"] = _sub(bench_kirchhoff, args, steps=20)
        out["freq"] = _sub(bench_freq, args, steps=3)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
