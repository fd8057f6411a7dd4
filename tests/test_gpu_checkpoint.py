"""Checkpoints on the GPU: a world's image holds the CPU checker's state (A, GA, S, V)
at the step it was saved; restored into a fresh engine (built from the image alone) it
continues bit for bit as the checker and as the uninterrupted GPU run, for every state
field; damaged or mismatched images are refused.  Layout intent: save_config /
load_config (config.cpp:75-261).

Bar: bit-identical A, GA, S, V; step counters carried over.
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


def _checker(ws, marks):
    """The CPU checker's state fields of every world after each number of steps in marks."""
    from oracle import oracle as O
    fmap = {"A": O.F_A, "GA": O.F_GA, "S": O.F_S, "V": O.F_V}
    out = {k: {} for k in marks}
    for i, w in enumerate(ws):
        ow = O.World(w["shape"], w["C"], w["kernels"], w["cfg"], 32)
        ow.set(O.F_A, w["A"])
        done = 0
        for k in marks:
            ow.step(k - done)
            done = k
            for f, of in fmap.items():
                out[k][(i, f)] = ow.get(of)
    return out


def _eq(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def test_restored_worlds_continue_bitwise():
    ws = _worlds()
    want = _checker(ws, (7, 12))
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
    # every image holds the checker's state at the save step
    for i, img in enumerate(images):
        ck = read_checkpoint(img)
        assert ck["steps"] == 7
        for f in STATE:
            assert _eq(ck["fields"][f].reshape(-1), want[7][(i, f)]), ("image", i, f)
    assert read_checkpoint(images[1])["fields"]["V"].any()
    fresh.step(5)
    fresh.synchronize()
    for i in range(len(ws)):
        assert fresh.steps(i) == 12
        for f in STATE:
            got = fresh.get(i, f)
            assert _eq(got, want[12][(i, f)]), ("restored vs checker", i, f)
            assert _eq(got, ref[(i, f)]), ("restored vs uninterrupted", i, f)
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
