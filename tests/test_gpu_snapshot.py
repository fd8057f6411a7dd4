"""Double-buffered read-backs (lh_snapshot*): a snapshot holds exactly the state after
the steps enqueued before it, however many steps are enqueued after it before it is
read; two slots alternate.  Expected values: the CPU checker's A, H and U_SUM after the
same number of steps (the main thread's GRID -> RBB, H_RB, U_SUM_RB copies,
main.cpp:225-236), and the same worlds stepped and read synchronously on the GPU.

Bar: bit-identical A, H, U_SUM."""
import numpy as np
import pytest

from lenia_amd import _lib as L
from lenia_amd import workloads as WL
from lenia_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _engine(ws):
    eng = Engine(keep_usum=True)
    for w in ws:
        eng.add_world(w["shape"], w["C"], w["kernels"], w["cfg"])
    eng.commit()
    for i, w in enumerate(ws):
        eng.set(i, "A", w["A"])
    return eng


def _eq(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _checker_states(ws, marks):
    """The CPU checker's A, H and U_SUM of every world after each number of steps in marks."""
    from oracle import oracle as O
    fmap = {"A": O.F_A, "H": O.F_H, "USUM": O.F_USUM}
    want = {k: {} for k in marks}
    for i, w in enumerate(ws):
        ow = O.World(w["shape"], w["C"], w["kernels"], w["cfg"], 32)
        ow.set(O.F_A, w["A"])
        done = 0
        for k in marks:
            ow.step(k - done)
            done = k
            for f, of in fmap.items():
                want[k][(i, f)] = ow.get(of)
    return want


def test_snapshots_hold_the_state_of_their_step():
    ws = [WL.plain32(seed=61), WL.world((24, 24, 24), 3, 9, 8, seed=62, mode="srt")]
    marks = (3, 7, 8, 12)
    want = _checker_states(ws, marks)
    ref = _engine(ws)
    done = 0
    for k in marks:  # the synchronous GPU run agrees with the checker at every mark
        ref.step(k - done)
        done = k
        for i in range(len(ws)):
            for f in ("A", "H", "USUM"):
                assert _eq(ref.get(i, f), want[k][(i, f)]), ("synchronous", k, i, f)
    ref.close()

    eng = _engine(ws)
    eng.step(3)
    eng.snapshot(0, fields=("A", "H", "USUM"))
    eng.step(4)
    eng.snapshot(1, fields=("A", "H"))
    eng.step(30)  # in flight while the snapshots are read
    for i in range(len(ws)):
        for f in ("A", "H", "USUM"):
            assert _eq(eng.snapshot_read(0, i, f), want[3][(i, f)]), (3, i, f)
        for f in ("A", "H"):
            assert _eq(eng.snapshot_read(1, i, f), want[7][(i, f)]), (7, i, f)
    assert eng.snapshot_ready(0) and eng.snapshot_ready(1)
    with pytest.raises(L.LeniaHipError, match="not in this snapshot"):
        eng.snapshot_read(1, 0, "USUM")
    eng.synchronize()

    # alternating slots on a fresh run, one world at a time
    eng2 = _engine(ws)
    eng2.step(8)
    eng2.snapshot(0, first_world=1, n_worlds=1)
    eng2.step(4)
    eng2.snapshot(1, first_world=0, n_worlds=1)
    assert _eq(eng2.snapshot_read(0, 1, "A"), want[8][(1, "A")])
    assert _eq(eng2.snapshot_read(1, 0, "A"), want[12][(0, "A")])
    with pytest.raises(L.LeniaHipError, match="world not in this snapshot"):
        eng2.snapshot_read(0, 0, "A")
    eng.close()
    eng2.close()
