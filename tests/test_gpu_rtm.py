"""GPU parity of the batched RTM (through the C ABI) against the FP64 oracle.

Tolerances (relative L2 against the FP64 oracle on the same seeded inputs):
  FP32 checkpoints            image and stack <= 1e-5
  FP16 / BF16 checkpoints     image and stack <= 1e-3
Batched (all shots per launch) and one-shot-per-launch runs give identical
images, bit for bit: a shot's result does not depend on the shots it shares
launches with.
"""
import numpy as np
import pytest

import oracle
import sssi_amd
from sssi_amd import synth

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


NZ, NX, NB, N, NT, DT, DX = 40, 170, 10, 3, 320, 1e-3, 10.0
SHOTS = [(0, 80, 40), (50, 100, 30), (95, 75, 37)]


@pytest.fixture(scope="module")
def model():
    V = synth.layered_2d(NZ, NX, 4, 1500.0, 3000.0, seed=21)
    V = V * (1 + synth.gaussian_perturbation(V.shape, 0.03, 3, seed=22))
    V = np.clip(V, 1500.0, 3000.0)
    VS = synth.smooth_copy(V, 3)
    w = synth.ricker(20.0, NT, DT)
    return V, VS, w


def oracle_shot(V, VS, w, xb, xc, src, stride, eps=1e-3):
    Vw = synth.extBoundary(V[:, xb:xb + xc], NB, 2)
    VSw = synth.extBoundary(VS[:, xb:xb + xc], NB, 2)
    return oracle.rtm_shot_cpml_2d(Vw, VSw, NT, [0], [src + NB], w, N, NB, DX, DX, DT,
                                   img_stride=stride, eps=eps)


@pytest.fixture(scope="module")
def reference(model):
    V, VS, w = model
    out = {}
    for stride in (1, 3):
        shots = [oracle_shot(V, VS, w, xb, xc, s, stride) for xb, xc, s in SHOTS]
        stack = np.zeros((NZ, NX))
        for (xb, xc, _), o in zip(SHOTS, shots):
            stack[:, xb:xb + xc] += o["image"][:NZ, NB:NB + xc]
        out[stride] = (shots, stack)
    return out


def run_gpu(model, stride, fmt, batched):
    V, VS, w = model
    r = sssi_amd.Rtm2D(V, VS, N, NB, DX, DX, DT, NT, img_stride=stride, ck_format=fmt)
    for xb, xc, s in SHOTS:
        r.add_shot(xb, xc, s, w)
    ms, launches = r.migrate(batched=batched)
    assert launches > 0 and ms > 0
    return r


@pytest.mark.parametrize("stride", [1, 3])
def test_rtm_fp32_matches_oracle(model, reference, stride):
    r = run_gpu(model, stride, "fp32", batched=True)
    shots, stack = reference[stride]
    for k, o in enumerate(shots):
        g = r.shot(k)
        assert rel_l2(g["data_true"], o["data_true"]) < 1e-5
        assert rel_l2(g["num"], o["num"]) < 1e-5, rel_l2(g["num"], o["num"])
        assert rel_l2(g["image"], o["image"]) < 1e-5, rel_l2(g["image"], o["image"])
    assert rel_l2(r.stack(), stack) < 1e-5


@pytest.mark.parametrize("fmt", ["fp16", "bf16"])
def test_rtm_half_checkpoints_within_1e3(model, reference, fmt):
    r = run_gpu(model, 3, fmt, batched=True)
    shots, stack = reference[3]
    for k, o in enumerate(shots):
        assert rel_l2(r.shot(k)["image"], o["image"]) < 1e-3
    assert rel_l2(r.stack(), stack) < 1e-3


@pytest.mark.parametrize("fmt", ["fp32", "fp16"])
def test_batched_equals_one_shot_per_launch(model, fmt):
    a = run_gpu(model, 3, fmt, batched=True)
    b = run_gpu(model, 3, fmt, batched=False)
    assert np.array_equal(a.stack(), b.stack())
    for k in range(len(SHOTS)):
        assert np.array_equal(a.shot(k)["image"], b.shot(k)["image"])


def test_shot_alone_equals_shot_in_batch(model):
    V, VS, w = model
    alone = sssi_amd.Rtm2D(V, VS, N, NB, DX, DX, DT, NT, img_stride=3, ck_format="bf16")
    alone.add_shot(*SHOTS[1], w)
    alone.migrate(batched=True)
    mix = run_gpu(model, 3, "bf16", batched=True)
    assert np.array_equal(alone.shot(0)["image"], mix.shot(1)["image"])
