"""lh_mc_case_noalloc (static device scratch, no allocation) against lh_mc_case, through
the C-ABI: two calls in a row, the second on a smaller world, on A, H and U_SUM, both
pad modes; both equal the CPU checker's case index (mesh.cpp:19-32, 173-185)."""
import numpy as np
import pytest

from lenia_amd import workloads as WL

pytestmark = pytest.mark.gpu


def test_mc_case_noalloc_matches_allocating_entry_point():
    from lenia_amd import _lib as L
    from lenia_amd.engine import Engine
    from oracle import oracle as O
    big = WL.world((128, 128, 128), 3, 9, 16, seed=61, asymptotic=[1, 1, 1, 1])
    small = WL.world((24, 24, 24), 2, 4, 6, seed=62, asymptotic=[1, 1, 1, 1])
    eng = Engine(keep_usum=True, autotune=False)
    ids = [eng.add_world(w["shape"], w["C"], w["kernels"], w["cfg"]) for w in (big, small)]
    eng.commit()
    for wid, w in zip(ids, (big, small)):
        eng.set(wid, "A", w["A"])
    eng.step(2)
    eng.synchronize()
    order = [(ids[0], big, "A", 0, True), (ids[1], small, "A", 1, True), (ids[0], big, "H", 2, False),
             (ids[1], small, "USUM", 0, False), (ids[1], small, "H", 1, True)]
    for wid, w, field, c, face_only in order:
        print(f"mc_case_noalloc world={wid} shape={w['shape']} field={field} c={c} face_only={face_only}",
              flush=True)
        (W, H, D) = w["shape"]
        got = eng.mc_case(wid, field, c, 0.5, face_only, noalloc=True)
        exp = eng.mc_case(wid, field, c, 0.5, face_only)
        assert np.array_equal(got, exp), (wid, field, c, face_only)
        f = eng.get(wid, field).reshape(w["C"], D, H, W)
        assert np.array_equal(got, O.mc_case(f[c], 0.5, face_only, 32))
        assert 0 < np.count_nonzero(got) < got.size
    # a second engine on the same GPU shares the scratch array
    eng2 = Engine(autotune=False)
    wid2 = eng2.add_world(small["shape"], small["C"], small["kernels"], small["cfg"])
    eng2.commit()
    eng2.set(wid2, "A", small["A"])
    got = eng2.mc_case(wid2, "A", 1, 0.3, True, noalloc=True)
    assert np.array_equal(got, eng2.mc_case(wid2, "A", 1, 0.3, True))
    with pytest.raises(L.LeniaHipError, match="bad channel"):
        eng2.mc_case(wid2, "A", 5, 0.5, True, noalloc=True)
    eng2.close()
    eng.close()
