"""CPU checks of the benchmark's contract: the one-GPU guard, the workload
definitions (sizes, seeding) and the committed traffic summaries that the
benchmark line quotes."""
import json
import os
import subprocess
import sys

import numpy as np

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

ALG_2D = 55_391_520     # bytes per launch, 2-D CPML 1040 x 3080, nB = 40
ALG_3D = 1_674_886_400  # bytes per launch, 3-D CPML 420 x 440 x 440, nB = 20
ALG_SP = 51_263_520     # bytes per launch, 2-D sponge 1040 x 3080


def test_multi_gpu_request_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_2d_workload_is_the_headline_configuration():
    wl = bench.workload()
    nz, nx = wl["V"].shape
    assert (nz, nx) == (1040, 3080)  # 1000 x 3000 plus the 40-cell extension
    assert (wl["N"], wl["nB"], wl["nt"]) == (5, 40, 4000)
    assert wl["dt"] == 5e-4 and wl["dz"] == wl["dx"] == 5.0
    # 16 B per cell, 20 B per damped-direction strip cell, 4 B per receiver
    nB = wl["nB"]
    assert 16 * nz * nx + 20 * (nB * nx + nz * 2 * nB) + 4 * nx == ALG_2D
    assert 16 * nz * nx + 4 * nx == ALG_SP  # the sponge workload's model and shot
    assert np.array_equal(bench.workload()["V"], wl["V"])  # seeded


def test_3d_and_rtm_workloads():
    wl = bench.workload_3d()
    nz, nx, ny = wl["V"].shape
    assert (nz, nx, ny) == (420, 440, 440)
    assert (wl["N"], wl["nB"]) == (4, 20)
    nB = wl["nB"]
    strips = nB * nx * ny + nz * ny * 2 * nB + nz * 2 * nB * nx
    assert 16 * nz * nx * ny + 20 * strips + 4 * nx * ny == ALG_3D

    r = bench.workload_rtm(64)
    assert r["V"].shape == (600, 6000) and r["nt"] == 3000 and r["N"] == 4
    assert len(r["shots"]) == 64
    for xb, xc, s in r["shots"]:
        assert 400 <= xc <= 1200 and 0 <= xb and xb + xc <= 6000 and s == xc // 2
    assert r["shots"] == bench.workload_rtm(64)["shots"]  # seeded


def test_committed_traffic_summaries():
    for name, alg in (("cpml2d_step_bench_traffic.json", ALG_2D),
                      ("cpml3d_step_bench_traffic.json", ALG_3D),
                      ("sponge2d_step_bench_traffic.json", ALG_SP)):
        with open(os.path.join(ROOT, "profiles", name)) as f:
            d = json.load(f)
        t = bench.traffic_from_profiles(name)
        scaled = d["fetch_kib"] * 1024 * d["fetch_scale"] + d["write_kib"] * 1024 * d["write_scale"]
        assert abs(t - scaled) < 1024  # the summary shows KiB to one decimal
        assert 1.0 <= t / alg < 2.0
