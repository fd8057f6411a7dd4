"""Checkpoints on the GPU: a world saved after some steps and restored into a fresh
engine (built from the image alone) continues bit for bit as the original, for every
state field; damaged or mismatched images are refused.

Bar: bit-identical A, GA, S, V after the continued steps; step counters carried over.
"""
import numpy as np
import pytest

from lenia_amd import _lib as L
from lenia_amd import workloads as WL
from lenia_amd.engine import Engine, read_checkpoint

pytestmark = pytest.mark.gpu

ASYM = [1, 1, 1, 1]
# reference step, advection with live velocity, and mass-conserving physics mode with
# oscillators, on two shapes in one batch
WORLDS = [
    dict(shape=(32, 32, 32), seed=41),
    dict(shape=(32, 32, 32), seed=42, mode="adv", asymptotic=ASYM),
    dict(shape=(24, 24, 24), seed=43, quirks=0, mode="srt", growth="none", osc_la=1, osc_la_amt=-0.02),
]
STATE = ("A", "GA", "S", "V")


def _worlds():
    out = []
    for spec in WORLDS:
        spec = dict(spec)
        shape = spec.pop("shape")
        out.append(WL.world(shape, 3, 9, 8, **spec))
    return out


def _engine(ws):
    eng = Engine(keep_velocity=True)
    for w in ws:
        eng.add_world(w["shape"], w["C"], w["kernels"], w["cfg"])
    eng.commit()
    for i, w in enumerate(ws):
        eng.set(i, "A", w["A"])
    return eng


def test_restored_worlds_continue_bitwise():
    ws = _worlds()
    eng = _engine(ws)
    eng.step(7)
    images = [eng.checkpoint(i) for i in range(len(ws))]
    eng.step(5)
    eng.synchronize()
    ref = {(i, f): eng.get(i, f) for i in range(len(ws)) for f in STATE}
    steps = [eng.steps(i) for i in range(len(ws))]
    eng.close()
    assert steps == [12] * len(ws)

    fresh = Engine(keep_velocity=True)
    for img in images:
        fresh.add_world_from_checkpoint(img)
    fresh.commit()
    for i, img in enumerate(images):
        fresh.restore(i, img)
        assert fresh.steps(i) == 7
    # the image's fields are what a direct read-back gives
    ck = read_checkpoint(images[1])
    assert ck["steps"] == 7 and ck["fields"]["V"].any()
    fresh.step(5)
    fresh.synchronize()
    for i in range(len(ws)):
        assert fresh.steps(i) == 12
        for f in STATE:
            got = fresh.get(i, f)
            assert np.array_equal(got.view(np.uint32), ref[(i, f)].view(np.uint32)), (i, f)
    fresh.close()


def test_mismatched_and_damaged_images_are_refused():
    ws = _worlds()[:2]
    eng = _engine(ws)
    eng.step(2)
    img0, img1 = eng.checkpoint(0), eng.checkpoint(1)
    with pytest.raises(L.LeniaHipError, match="parameters|kernels"):
        eng.restore(0, img1)  # another world's image (same shape, other parameters)
    with pytest.raises(L.LeniaHipError, match="size"):
        eng.restore(0, img0[:-4])
    bad = bytearray(img0.tobytes())
    bad[0] = ord("X")
    with pytest.raises(L.LeniaHipError, match="magic"):
        eng.restore(0, bytes(bad))
    # a refused image leaves the world untouched; the right one still loads
    before = eng.get(0, "A")
    eng.restore(0, img0)
    assert np.array_equal(eng.get(0, "A"), before)
    eng.close()
