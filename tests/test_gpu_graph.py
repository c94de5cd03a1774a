"""The hipGraph-captured time loop (64 step launches reading a device-resident
step base, replayed) against direct launches: identical results bit for bit
-- a forward run with snapshots and receivers, a run continued across calls
(sssi_prop_run_steps; each call re-captures), a reverse run, batched RTM
phases, and autotuned layouts.
SSSI_GRAPH=1 switches capture on (default off); it is read at every run.
"""
import os

import numpy as np
import pytest

import sssi_amd
from sssi_amd import synth

pytestmark = pytest.mark.gpu


def with_graph(on, fn):
    old = os.environ.get("SSSI_GRAPH")
    os.environ["SSSI_GRAPH"] = "1" if on else "0"
    try:
        return fn()
    finally:
        if old is None:
            os.environ.pop("SSSI_GRAPH", None)
        else:
            os.environ["SSSI_GRAPH"] = old


def model(nz, nx, seed):
    v = synth.layered_2d(nz, nx, 5, 1500.0, 3500.0, seed)
    return np.clip(v * (1 + synth.gaussian_perturbation((nz, nx), 0.05, 4, seed + 1)), 1500, 3500)


def test_forward_graph_equals_direct():
    v = model(150, 330, 3)
    nt, dt = 333, 5e-4
    w = synth.ricker(25, nt, dt)

    def run():
        return sssi_amd.fwd_cpml_2d_points(v, nt, 4, 15, 5.0, 5.0, dt, [0, 60], [165, 40],
                                           np.stack([w, -w], axis=1), want_snapshot=True,
                                           snap_stride=9)
    d0, s0 = with_graph(False, run)
    d1, s1 = with_graph(True, run)
    assert np.array_equal(d0, d1) and np.array_equal(s0, s1)


def test_continued_runs_equal_one_run():
    v = model(120, 260, 5)
    nt, dt = 400, 5e-4
    w = synth.ricker(20, nt, dt)

    def run(splits):
        P = sssi_amd.Propagator(v, 3, 12, 5.0, 5.0, dt)
        P.set_points([0], [130], w)
        t = 0
        for n in splits:
            P.run_steps(t, n)
            t += n
        P.sync()
        return P.data(nt), P.field()
    d0, f0 = with_graph(False, lambda: run([nt]))
    for splits in ([nt], [150, 250], [130, 140, 130]):
        d1, f1 = with_graph(True, lambda: run(splits))
        assert np.array_equal(d0, d1), splits
        assert np.array_equal(f0, f1), splits


def test_reverse_graph_equals_direct():
    v = model(100, 240, 7)
    nt, dt = 300, 5e-4
    data = np.outer(np.linspace(0.5, 1.5, 240), synth.ricker(20, nt, dt))

    def run():
        return sssi_amd.rvs_cpml_2d_ex(v, data, 5, 10, 5.0, 5.0, dt, want_snapshot=True,
                                       snap_stride=7)
    m0, s0 = with_graph(False, run)
    m1, s1 = with_graph(True, run)
    assert np.array_equal(m0, m1) and np.array_equal(s0, s1)


@pytest.mark.parametrize("batched", [True, False])
def test_rtm_graph_equals_direct(batched):
    V = model(80, 500, 9)
    VS = synth.smooth_copy(V, 4)
    nt, dt = 260, 5e-4
    w = synth.ricker(25, nt, dt)

    def run():
        r = sssi_amd.Rtm2D(V, VS, 3, 15, 5.0, 5.0, dt, nt, img_stride=3, ck_format="bf16")
        for xb, xc, s in [(0, 200, 100), (150, 300, 120), (380, 120, 60)]:
            r.add_shot(xb, xc, s, w)
        r.migrate(batched=batched)
        return r.stack()
    a = with_graph(False, run)
    b = with_graph(True, run)
    assert np.array_equal(a, b)


def test_autotuned_layouts_match_oracle_and_are_cached():
    """The autotuner picks a layout per shape; the picked layout passes the
    same oracle tolerance, and a second handle of the shape reuses it."""
    import oracle
    v = model(140, 400, 11)
    nt, dt = 200, 5e-4
    w = synth.ricker(25, nt, dt)
    sssi_amd.set_autotune(True)
    try:
        P = sssi_amd.Propagator(v, 4, 20, 5.0, 5.0, dt)
        lay = P.layout()
        assert lay in {(4, 0), (4, 2), (4, 3), (4, 4), (4, 8), (8, 0), (8, 2), (8, 4)}
        P.set_points([0], [200], w)
        P.run(nt)
        d = P.data(nt)
        Q = sssi_amd.Propagator(v, 4, 20, 5.0, 5.0, dt)
        assert Q.layout() == lay
        v3 = synth.cake_3d_dipping(40, 70, 66, 3, 1500, 3000, seed=2)
        R = sssi_amd.Propagator(v3, 2, 6, 10.0, 10.0, 1e-3, dy=10.0)
        ty, zc = R.layout()
        assert ty in (4, 8, 16) and zc in (32, 48, 64, 80, 128)
    finally:
        sssi_amd.set_autotune(False)
    d_ref, _ = oracle.fwd_cpml_2d(v, nt, 4, 20, 5.0, 5.0, dt, src_iz=[0], src_ix=[200],
                                  wavelet=w)
    assert np.linalg.norm(d - d_ref) / np.linalg.norm(d_ref) < 1e-5


def test_invalid_tune_cache_entries_are_ignored(tmp_path):
    """Entries of SSSI_TUNE_CACHE that name no valid layout (a hand-edited or
    damaged file) are ignored and the shape is tuned afresh.  The file is read
    once per process, so the handles are made in a child process."""
    import json
    import subprocess
    import sys
    cache = tmp_path / "tune.txt"
    # "dim nz nx ny N nB a b" for the two shapes below, with layouts no kernel has
    cache.write_text("2 64 96 1 3 8 7 3\n3 24 40 36 2 6 5 0\n")
    code = (
        "import json, numpy as np, sssi_amd\n"
        "from sssi_amd import synth\n"
        "sssi_amd.set_autotune(True)\n"
        "v2 = synth.layered_2d(64, 96, 3, 1500.0, 2500.0, seed=1)\n"
        "p2 = sssi_amd.Propagator(v2, 3, 8, 10.0, 10.0, 1e-3)\n"
        "v3 = synth.cake_3d_dipping(24, 40, 36, 3, 1500.0, 2500.0, seed=2)\n"
        "p3 = sssi_amd.Propagator(v3, 2, 6, 10.0, 10.0, 1e-3, dy=10.0)\n"
        "p2.run(5); p3.run(5); p2.sync(); p3.sync()\n"
        "print(json.dumps([list(p2.layout()), list(p3.layout())]))\n")
    env = dict(os.environ, SSSI_TUNE_CACHE=str(cache))
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                       timeout=240, cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    assert r.returncode == 0, r.stderr[-3000:]
    (rz, ng), (ty, zc) = json.loads(r.stdout.strip().splitlines()[-1])
    assert (rz, ng) in {(4, 0), (4, 2), (4, 3), (4, 4), (4, 8), (8, 0), (8, 2), (8, 4)}
    assert ty in (4, 8, 16) and 8 <= zc <= 4096
    # the re-tuned layouts were appended to the cache
    assert len(cache.read_text().splitlines()) == 4
