"""GPU marching-cubes meshes (lh_mesh) against the CPU checker, bit for bit: vertex
positions, accumulated normals, colours and the index buffer of views 1-3
(generate_mesh_v1/v2/v3, mesh.cpp:34, 434, 1115) on stepped worlds."""
import numpy as np
import pytest

from lenia_amd import workloads as WL
from lenia_amd import _lib as L

pytestmark = pytest.mark.gpu

ASYM = [1, 1, 1, 1]
COLORS = [0xFF2040C0, 0xFF10E020, 0xFFE0A000, 0xFF808080]


def stepped(shape=(32, 32, 32), C=3, seed=50, steps=4):
    from lenia_amd.engine import Engine
    w = WL.world(shape, C, 6, 6, seed, asymptotic=ASYM)
    eng = Engine(keep_usum=True)
    wid = eng.add_world(w["shape"], C, w["kernels"], w["cfg"])
    eng.commit()
    eng.set(wid, "A", w["A"])
    eng.step(steps)
    eng.synchronize()
    return eng, wid


def fields(eng, wid, shape, C, view):
    W, H, D = shape
    names = ["A", "USUM", "H"][:view]
    return [eng.get(wid, n).reshape(C, D, H, W) for n in names]


def same(gpu, ref):
    (gv, gi), (rv, ri) = gpu, ref
    assert len(gv) == len(rv) and len(gi) == len(ri)
    assert np.array_equal(gi, ri)
    assert gv.tobytes() == rv.tobytes()


@pytest.mark.parametrize("view", [1, 2, 3])
def test_mesh_views_bitwise(view):
    from lenia_amd.engine import mc_table
    from oracle import oracle as O
    shape, C = (32, 32, 32), 3
    eng, wid = stepped(shape, C)
    tab = mc_table()
    for face_only in (True, False):
        print(f"view={view} face_only={face_only}", flush=True)
        gpu = eng.mesh(wid, view=view, iso=0.5, cell_size=3.5, face_only=face_only, colors=COLORS[:C])
        ref = O.mesh(fields(eng, wid, shape, C, view), tab, iso=0.5, cell=3.5, face_only=face_only,
                     colors=COLORS[:C])
        assert len(gpu[0]) > 100
        same(gpu, ref)
    eng.close()


def test_mesh_non_cubic_two_channels_other_iso():
    from lenia_amd.engine import mc_table
    from oracle import oracle as O
    shape, C = (44, 24, 44), 2
    eng, wid = stepped(shape, C, seed=51, steps=3)
    tab = mc_table()
    for iso in (0.2, 0.65):
        gpu = eng.mesh(wid, view=3, iso=iso, cell_size=1.25, colors=COLORS[:C])
        ref = O.mesh(fields(eng, wid, shape, C, 3), tab, iso=iso, cell=1.25, colors=COLORS[:C])
        same(gpu, ref)
    eng.close()


def test_mesh_follows_the_callers_table():
    """A caller's table (here every triangle's orientation reversed) drives both the
    triangles and the normals."""
    from lenia_amd.engine import mc_table
    from oracle import oracle as O
    shape, C = (32, 32, 32), 3
    eng, wid = stepped(shape, C, seed=52)
    tab = mc_table()
    flipped = tab.copy()
    for m in range(256):
        n = int((tab[m] >= 0).sum())
        for t in range(n // 3):
            flipped[m, 3 * t + 1], flipped[m, 3 * t + 2] = tab[m, 3 * t + 2], tab[m, 3 * t + 1]
    gpu = eng.mesh(wid, view=1, tri_table=flipped)
    ref = O.mesh(fields(eng, wid, shape, C, 1), flipped, iso=0.5, cell=3.5)
    same(gpu, ref)
    base = eng.mesh(wid, view=1)
    assert len(base[0]) == len(gpu[0])
    assert not np.array_equal(base[1], gpu[1])
    eng.close()


def test_mesh_matches_at_128_cubed():
    from lenia_amd.engine import mc_table
    from oracle import oracle as O
    shape, C = (128, 128, 128), 3
    eng, wid = stepped(shape, C, seed=53, steps=2)
    gpu = eng.mesh(wid, view=1, colors=COLORS[:C])
    ref = O.mesh(fields(eng, wid, shape, C, 1), mc_table(), iso=0.5, cell=3.5, colors=COLORS[:C])
    assert len(gpu[0]) > 1000
    same(gpu, ref)
    eng.close()


def test_mesh_errors_and_size_query():
    import ctypes as C_
    from lenia_amd.engine import Engine, mc_table
    w = WL.plain32(seed=54, asymptotic=ASYM)
    eng = Engine()  # no U_SUM kept
    wid = eng.add_world(w["shape"], 3, w["kernels"], w["cfg"])
    eng.commit()
    eng.set(wid, "A", w["A"])
    eng.step(2)
    with pytest.raises(L.LeniaHipError, match="U_SUM"):
        eng.mesh(wid, view=2)
    with pytest.raises(L.LeniaHipError, match="view"):
        eng.mesh(wid, view=4)
    bad = mc_table()
    bad[1, :3] = [0, 1, 2]  # case 1 does not cross edge 1
    with pytest.raises(L.LeniaHipError, match="triangulation table"):
        eng.mesh(wid, tri_table=bad)
    v, i = eng.mesh(wid, view=1)
    nv, ni = C_.c_int64(), C_.c_int64()
    lib = L.load()
    rc = lib.lh_mesh(eng._h, wid, 1, 0.5, 3.5, 1, None, None, None, 0, None, 0, C_.byref(nv), C_.byref(ni))
    assert rc == 0 and (nv.value, ni.value) == (len(v), len(i))
    small = np.zeros(max(1, len(v) - 1), dtype=v.dtype)
    idx = np.zeros(len(i), dtype=np.uint32)
    rc = lib.lh_mesh(eng._h, wid, 1, 0.5, 3.5, 1, None, None, small.ctypes.data_as(C_.c_void_p), len(v) - 1,
                     idx.ctypes.data_as(C_.POINTER(C_.c_uint32)), len(i), C_.byref(nv), C_.byref(ni))
    assert rc == L.EINVAL and nv.value == len(v)
    eng.close()


def test_mesh_of_a_snapshot_while_steps_run():
    """lh_mesh_snapshot meshes the fields a snapshot took (the mesh threads' RBB, U_SUM_RB,
    H_RB) while 30 later steps are in flight: bit-identical to the checker on those fields,
    and to lh_mesh of a second engine stepped to the same point."""
    from lenia_amd.engine import Engine, mc_table
    from oracle import oracle as O
    shape, C = (32, 32, 32), 3
    eng, wid = stepped(shape, C, seed=55, steps=3)
    eng.snapshot(0, 0, 1, ("A", "USUM", "H"))
    eng.step(30)
    gpu = eng.mesh(wid, view=3, colors=COLORS[:C], snapshot_slot=0)
    snap = [eng.snapshot_read(0, wid, n).reshape(C, 32, 32, 32) for n in ("A", "USUM", "H")]
    same(gpu, O.mesh(snap, mc_table(), colors=COLORS[:C]))
    ref_eng, ref_wid = stepped(shape, C, seed=55, steps=3)
    same(gpu, ref_eng.mesh(ref_wid, view=3, colors=COLORS[:C]))
    ref_eng.close()
    live = eng.mesh(wid, view=3)  # the state 30 steps later meshes differently
    assert not (len(live[1]) == len(gpu[1]) and np.array_equal(live[1], gpu[1]))
    eng.snapshot(1, 0, 1, ("A",))
    with pytest.raises(L.LeniaHipError, match="not in the snapshot"):
        eng.mesh(wid, view=2, snapshot_slot=1)
    eng.close()


@pytest.mark.parametrize("view", [1, 3])
def test_mesh_with_the_reference_table_bitwise(view):
    """The reference's own TRIANGLE_TABLE (tests/golden/tri_table.json), passed as the
    reference integration passes it: GPU mesh bit-identical to the checker."""
    import json
    import os
    from oracle import oracle as O
    ref_tab = np.array(json.load(open(os.path.join(os.path.dirname(__file__), "golden", "tri_table.json")))["rows"],
                       dtype=np.int32)
    shape, C = (32, 32, 32), 3
    eng, wid = stepped(shape, C, seed=56)
    gpu = eng.mesh(wid, view=view, tri_table=ref_tab, colors=COLORS[:C])
    ref = O.mesh(fields(eng, wid, shape, C, view), ref_tab, colors=COLORS[:C])
    assert len(gpu[0]) > 100
    same(gpu, ref)
    eng.close()


def test_entries_after_a_row_terminator_are_ignored():
    """A caller's table whose rows hold values after their -1 terminator meshes exactly as
    the same table padded with -1 (rows end at their first negative entry)."""
    import json
    import os
    ref_tab = np.array(json.load(open(os.path.join(os.path.dirname(__file__), "golden", "tri_table.json")))["rows"],
                       dtype=np.int32)
    junk = ref_tab.copy()
    for cs in range(256):
        n = int(np.argmax(junk[cs] < 0)) if (junk[cs] < 0).any() else 16
        if n + 1 < 16:
            junk[cs, n + 1:] = np.arange(16 - n - 1) % 12 + (cs % 3) * 7  # edges, junk and > 11
    assert not np.array_equal(junk, ref_tab)
    shape, C = (32, 32, 32), 3
    eng, wid = stepped(shape, C, seed=57)
    a = eng.mesh(wid, view=1, tri_table=ref_tab, colors=COLORS[:C])
    b = eng.mesh(wid, view=1, tri_table=junk, colors=COLORS[:C])
    same(a, b)
    eng.close()
