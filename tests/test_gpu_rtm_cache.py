"""Durable per-shot RTM results with resume by existence (the reference's apps
save dataTrue<xs>.mat / rtmsnapshot<xs>.mat / StackedRTM.mat and skip what
exists, applications/rtm_2d_acoustic/mainRtmTimeCpmlFor2dAw.m:170-184,
:259-264, mainRtmTimeSpongeFor2dAw.m:343-346):

- a migration interrupted after some shots and resumed from the cache equals
  the uninterrupted migration bit for bit (stack and every per-shot image);
- the resumed run loads, and does not recompute, the finished shots;
- the files are MAT-files that scipy.io reads, with the reference's variable
  names and shapes;
- a cache made from other inputs is refused (SSSI_ESTATE)."""
import os

import numpy as np
import pytest
import scipy.io

import sssi_amd
from sssi_amd import synth

pytestmark = pytest.mark.gpu


def _setup():
    nz, nx, nt, N, nB = 60, 320, 300, 3, 10
    V = synth.layered_2d(nz, nx, 4, 1500.0, 3500.0, seed=21)
    VS = synth.smooth_copy(V, 3)
    w = synth.ricker(25.0, nt, 1e-3)
    shots = [(0, 120, 60), (40, 150, 75), (100, 110, 30), (170, 150, 80), (200, 120, 100)]
    return V, VS, w, shots, (N, nB, 10.0, 10.0, 1e-3, nt)


def _handle(V, VS, p, shots, w, cache):
    r = sssi_amd.Rtm2D(V, VS, *p, img_stride=3, ck_format="bf16")
    if cache is not None:
        r.set_cache(cache)
    for xb, xc, s in shots:
        r.add_shot(xb, xc, s, w)
    return r


@pytest.mark.parametrize("batched", [True, False])
def test_resumed_migration_equals_uninterrupted(tmp_path, batched):
    V, VS, w, shots, p = _setup()
    nB, nt = p[1], p[5]
    print(f"rtm cache: batched={batched}", flush=True)
    a = _handle(V, VS, p, shots, w, tmp_path / "a")
    a.migrate(batched=batched)
    stack_a = a.stack()
    img_a = [a.shot(q)["image"] for q in range(len(shots))]
    assert a.cache_stats() == {"loaded_data": 0, "loaded_images": 0,
                               "saved": 2 * len(shots) + 1}
    a.close()
    # the run that stopped after two shots: their files exist, the stack not yet
    b = _handle(V, VS, p, shots[:2], w, tmp_path / "b")
    b.migrate(batched=batched)
    b.close()
    os.remove(tmp_path / "b" / "StackedRTM.mat")
    c = _handle(V, VS, p, shots, w, tmp_path / "b")
    c.migrate(batched=batched)
    assert c.cache_stats()["loaded_data"] == 2 and c.cache_stats()["loaded_images"] == 2
    assert np.any(stack_a != 0)
    assert np.array_equal(c.stack(), stack_a)
    for q in range(len(shots)):
        assert np.array_equal(c.shot(q)["image"], img_a[q]), q
    c.close()
    # MAT-files with the reference's names and shapes, readable by scipy
    xb, xc, s = shots[0]
    m = scipy.io.loadmat(str(tmp_path / "b" / f"dataTrue{xb + s + 1}.mat"))
    assert m["dataTrue"].shape == (xc + 2 * nB, nt)
    st = scipy.io.loadmat(str(tmp_path / "b" / "StackedRTM.mat"))["StackedRTM"]
    assert np.array_equal(st, stack_a)
    img = scipy.io.loadmat(str(tmp_path / "b" / f"image{xb + s + 1}.mat"))["image"]
    assert np.array_equal(img, img_a[0])


def test_cache_from_other_inputs_is_refused(tmp_path):
    V, VS, w, shots, p = _setup()
    a = _handle(V, VS, p, shots[:1], w, tmp_path)
    a.migrate()
    a.close()
    b = _handle(V, VS * 1.01, p, shots[:1], w, tmp_path)   # another migration model
    with pytest.raises(sssi_amd.SssiError) as e:
        b.migrate()
    assert e.value.code == -4 and "other inputs" in str(e.value)
    b.close()
