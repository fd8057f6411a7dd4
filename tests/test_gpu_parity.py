"""GPU parity: the HIP step engine (through the C-ABI) against the CPU checker.

Bar: bit-identical fp32 results for the reference-compatible step (growth table,
reference halos and index quirks); 1e-4 max-abs per channel against the fp64 checker
where growth is analytic; FP64 per-world mass conserved to 1e-5 relative in physics mode.
"""
import numpy as np
import pytest

from lenia_amd import workloads as WL
from tests.pairs import mismatch, run_pair

pytestmark = pytest.mark.gpu

ASYM = [1, 1, 1, 1]
OSC = dict(osc_la=1, osc_flow=1, osc_div_flow=1, osc_la_amt=-0.02, osc_flow_amt=0.03, osc_div_flow_amt=-0.04)

CASES = {
    "rt_sobel_torus": dict(seed=1),
    "rt_asymptotic": dict(seed=2, asymptotic=ASYM),
    "srt": dict(seed=3, mode="srt", asymptotic=ASYM),
    "adv_vel": dict(seed=4, mode="adv", asymptotic=ASYM),
    "adv_force": dict(seed=14, mode="adv", adv_vel=False, asymptotic=ASYM),
    "cd_blur_zero": dict(seed=5, gradient="central", blur=True, sobel_border="none", border="none",
                         asymptotic=ASYM),
    "cd_copy_torus": dict(seed=6, gradient="central", blur=False, asymptotic=ASYM),
    "cd_copy_wall": dict(seed=16, gradient="central", blur=False, sobel_border="wall", asymptotic=ASYM),
    "wall_mu_vel": dict(seed=7, border="wall", sobel_border="wall", mu_vel=True, asymptotic=ASYM),
    "sobel_none_rt_other": dict(seed=17, sobel_border="none", border="none", asymptotic=ASYM),
    "growth_transposition": dict(seed=18, growth="transposition", asymptotic=ASYM),
    "soft_clip": dict(seed=19, soft_clip=[1, 0, 1, 0], asymptotic=[0, 1, 1, 0]),
    "soft_clip_transposition": dict(seed=20, growth="transposition", soft_clip=[1, 1, 0, 0]),
    "growth_none_dd2": dict(seed=21, growth="none", dd=2.0, sigma=0.8, asymptotic=ASYM),
    "s_norm_half": dict(seed=22, s_norm=0.5, sigma=0.55, asymptotic=ASYM),
    # no reference quirks: analytic growth bell, true Sobel axes, full halos
    "physics_rt_growth": dict(seed=27, quirks=0),
    "physics_none_srt": dict(seed=28, quirks=0, mode="srt", growth="none"),
    "physics_transposition_clip": dict(seed=29, quirks=0, growth="transposition", soft_clip=[1, 0, 1, 0]),
    # oscillators (lenia.cpp:803-911): Laplacian, flow and divergence-of-flow terms
    "osc_all_vel": dict(seed=23, asymptotic=ASYM, **OSC, osc_vel=1),
    "osc_force_wall": dict(seed=24, border="wall", sobel_border="wall", asymptotic=ASYM, **OSC, osc_vel=0),
    "osc_adv_zero": dict(seed=25, mode="adv", border="none", sobel_border="none", quirks=0x7EF,
                         asymptotic=ASYM, **OSC, osc_vel=1),
    "osc_physics_srt": dict(seed=26, mode="srt", quirks=0, growth="none", **OSC, osc_vel=0),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_one_step_all_fields_bitwise(name):
    w = WL.plain32(**CASES[name])
    fields = ["A", "GA", "S", "V", "H"]
    if w["cfg"]["growth"] == "rt":
        fields.append("B")
    if w["cfg"]["mode"] != "adv":
        fields.append("MU")
    r = run_pair(w, 1, fields=fields)
    for f in fields:
        bad, err = mismatch(*r[f])
        assert bad == 0, f"{name}: field {f} differs in {bad} values (max abs {err:.3e})"


@pytest.mark.parametrize("name", sorted(CASES))
def test_ten_steps_bitwise(name):
    w = WL.plain32(**CASES[name])
    r = run_pair(w, 10, fields=["A", "S"])
    for f in ("A", "S"):
        bad, err = mismatch(*r[f])
        assert bad == 0, f"{name}: {f} after 10 steps differs in {bad} values (max abs {err:.3e})"


FUSED = (0, 0, False, 2)  # forced schedule: launchers' sweep defaults, spectral growth fused


@pytest.mark.parametrize("name", ["rt_sobel_torus", "rt_asymptotic", "soft_clip", "physics_rt_growth",
                                  "cd_blur_zero", "growth_transposition", "srt"])
def test_fused_growth_bitwise(name):
    """The spectral growth as one fused launch (spec_z + inv_growth with the per-kernel
    spectra kept in LDS): every field bit-identical to the checker after 3 steps, with the
    growth table and the analytic bell, asymptotic and soft-clip channels, U_SUM kept."""
    w = WL.plain32(**CASES[name])
    fields = ["A", "GA", "S", "H", "USUM"] + (["B"] if w["cfg"]["growth"] == "rt" else [])
    for keep_usum in (False, True):
        f = fields if keep_usum else [x for x in fields if x != "USUM"]
        r = run_pair(w, 3, fields=f, keep_usum=keep_usum, schedule=FUSED)
        for x in f:
            bad, err = mismatch(*r[x])
            assert bad == 0, f"{name} keep_usum={keep_usum}: {x} differs in {bad} values (max abs {err:.3e})"


@pytest.mark.parametrize("shape,C,K,R", [((8, 8, 8), 3, 4, 2), ((8, 4, 8), 2, 3, 2), ((12, 12, 12), 4, 6, 3),
                                         ((16, 16, 16), 4, 16, 4), ((16, 8, 16), 3, 5, 3),
                                         ((24, 24, 24), 3, 9, 6), ((32, 32, 32), 2, 5, 8)])
def test_fused_growth_shapes_bitwise(shape, C, K, R):
    """Every shape whose half spectrum fits in LDS, with 2 to 4 channels and up to 16
    kernels (uneven kernels per destination channel), fused against the checker."""
    w = WL.world(shape, C, K, R, seed=sum(shape) + 77, asymptotic=[1, 0, 1, 0])
    r = run_pair(w, 4, fields=["A", "H", "B"], schedule=FUSED)
    for f in ("A", "H", "B"):
        bad, err = mismatch(*r[f])
        assert bad == 0, f"{shape}: {f} differs in {bad} values (max abs {err:.3e})"


@pytest.mark.parametrize("shape,C,K,R", [((8, 8, 8), 3, 4, 2), ((8, 4, 8), 2, 3, 2), ((12, 12, 12), 2, 3, 3),
                                         ((16, 16, 16), 4, 6, 4), ((16, 8, 16), 3, 5, 3),
                                         ((24, 24, 24), 3, 9, 6), ((36, 36, 36), 3, 9, 8),
                                         ((40, 40, 40), 3, 9, 8), ((44, 24, 44), 3, 9, 8),
                                         ((48, 48, 48), 4, 16, 8)])
def test_shapes_bitwise(shape, C, K, R):
    w = WL.world(shape, C, K, R, seed=sum(shape), asymptotic=ASYM)
    r = run_pair(w, 5, fields=["A", "H", "MU"])
    for f in ("A", "H", "MU"):
        bad, err = mismatch(*r[f])
        assert bad == 0, f"{shape}: {f} differs in {bad} values (max abs {err:.3e})"


@pytest.mark.parametrize("shape,C,K,R,steps,over", [
    ((64, 64, 64), 3, 9, 8, 3, dict(asymptotic=ASYM)),
    ((64, 64, 64), 2, 4, 6, 2, dict(mode="adv", gradient="central", blur=True, sobel_border="none",
                                    border="none", asymptotic=ASYM)),
    # the 96^3 workload: C=4, K=16, R=12, central differences + blur, advection, zero border
    ((96, 96, 96), 4, 16, 12, 10, dict(mode="adv", gradient="central", blur=True, sobel_border="none",
                                       border="none", adv_vel=True, asymptotic=ASYM)),
    ((128, 128, 128), 3, 9, 16, 1, dict(asymptotic=ASYM)),
])
def test_large_worlds_bitwise(shape, C, K, R, steps, over):
    """Large worlds run the LDS-cooperative FFT kernels and 32-wide stencil tiles; the
    96^3 workload (central differences + blur, advection, zero border) for 10 steps."""
    import os
    from oracle import oracle as O
    w = WL.world(shape, C, K, R, seed=sum(shape) + steps, **over)
    fields = ["A", "H", "S"] + (["MU"] if w["cfg"]["mode"] != "adv" else []) + \
        (["V"] if w["cfg"]["mode"] == "adv" else [])
    O.set_threads(min(8, os.cpu_count() or 1))  # the checker's result does not depend on it
    try:
        r = run_pair(w, steps, fields=fields, keep_velocity=True)
    finally:
        O.set_threads(1)
    for f in fields:
        bad, err = mismatch(*r[f])
        assert bad == 0, f"{shape}: {f} differs in {bad} values (max abs {err:.3e})"


@pytest.mark.parametrize("shape,C,K,R,steps,over", [
    # radix 7, register-resident transforms; every phase, growth fused where it fits
    ((28, 28, 28), 3, 9, 6, 3, dict(asymptotic=ASYM)),
    ((28, 28, 28), 2, 4, 5, 2, dict(mode="srt", border="wall", sobel_border="wall", mu_vel=True)),
    # non-cube with radices 5 and 3
    ((20, 12, 20), 3, 6, 4, 3, dict(gradient="central", blur=True, sobel_border="none", border="none",
                                    asymptotic=ASYM)),
    # above 48: LDS-staged transforms with radix 13, stencil tiles of 26 cells
    ((52, 52, 52), 2, 4, 7, 2, dict(mode="adv", asymptotic=ASYM)),
    ((52, 52, 52), 3, 5, 7, 2, dict(asymptotic=ASYM)),
])
def test_shape_modules_bitwise(shape, C, K, R, steps, over):
    """Shapes outside the built-in set run from shape modules built on demand
    (liblenia_shape_<W>x<H>x<D>.so), bit-identical to the checker like the built-in ones."""
    import os
    from lenia_amd.engine import shape_supported
    from oracle import oracle as O
    assert shape_supported(*shape) == 2, f"{shape}: shape module not built (lenia_amd.build.TEST_SHAPES)"
    w = WL.world(shape, C, K, R, seed=sum(shape) + steps + C, **over)
    fields = ["A", "H", "S"] + (["MU"] if w["cfg"]["mode"] != "adv" else []) + \
        (["V"] if w["cfg"]["mode"] == "adv" or w["cfg"].get("mu_vel") else [])
    O.set_threads(min(8, os.cpu_count() or 1))
    try:
        r = run_pair(w, steps, fields=fields, keep_velocity=True)
    finally:
        O.set_threads(1)
    for f in fields:
        bad, err = mismatch(*r[f])
        assert bad == 0, f"{shape}: {f} differs in {bad} values (max abs {err:.3e})"


@pytest.mark.parametrize("shape,C,K,R", [((16, 8, 16), 3, 5, 3), ((64, 64, 64), 2, 4, 6)])
def test_oscillators_shapes_bitwise(shape, C, K, R):
    w = WL.world(shape, C, K, R, seed=sum(shape) + 7, asymptotic=ASYM, mu_vel=True, **OSC, osc_vel=1)
    r = run_pair(w, 3, fields=["A", "GA", "V"])
    for f in ("A", "GA", "V"):
        bad, err = mismatch(*r[f])
        assert bad == 0, f"{shape}: {f} differs in {bad} values (max abs {err:.3e})"


@pytest.mark.parametrize("over", [{}, {"asymptotic": ASYM}])
def test_config0_100_steps(over):
    """32^3, C=3, K=9, RT, Sobel, torus, growth table: 100 steps, bit-identical
    (so the max-abs error per channel is 0 <= 1e-4)."""
    w = WL.plain32(seed=1, **over)
    r = run_pair(w, 100, fields=["A"], keep_velocity=False)
    gpu, ref = r["A"]
    bad, err = mismatch(gpu, ref)
    per_channel = np.abs(gpu.reshape(3, -1).astype(np.float64) - ref.reshape(3, -1)).max(axis=1)
    assert bad == 0 and (per_channel <= 1e-4).all(), (bad, per_channel)


def test_physics_mode_vs_fp64_one_step():
    w = WL.plain32(seed=9, quirks=0, growth="none")
    r = run_pair(w, 1, fields=["A"], keep_velocity=False, prec=64)
    gpu, ref = r["A"]
    assert np.abs(gpu.astype(np.float64) - ref).reshape(3, -1).max(axis=1).max() <= 1e-4


def test_physics_mode_mass_conservation():
    from lenia_amd.engine import Engine
    w = WL.plain32(seed=4, quirks=0, growth="none")
    eng = Engine(keep_velocity=False)
    wid = eng.add_world(w["shape"], 3, w["kernels"], w["cfg"])
    eng.commit()
    eng.set(wid, "A", w["A"])
    m0 = eng.mass()[0]
    assert np.isclose(m0, w["A"].astype(np.float64).sum(), rtol=1e-12)
    eng.step(100)
    m1 = eng.mass()[0]
    eng.close()
    assert abs(m1 - m0) / m0 <= 1e-5, (m0, m1)


def test_graph_and_eager_identical():
    from lenia_amd.engine import Engine
    w = WL.plain32(seed=12, asymptotic=ASYM)
    outs = []
    for graph in (True, False):
        eng = Engine(graph=graph)
        wid = eng.add_world(w["shape"], 3, w["kernels"], w["cfg"])
        eng.commit()
        eng.set(wid, "A", w["A"])
        for n in (3, 4, 1, 2):  # odd and even chunks, graph replays between eager steps
            eng.step(n)
            eng.synchronize()
            outs.append(eng.get(wid, "A"))
        eng.close()
    half = len(outs) // 2
    for a, b in zip(outs[:half], outs[half:]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_batched_mixed_shapes_match_single_worlds():
    """A batch of mixed shapes and per-world SRT/RT equals each world run alone, and
    spot-checked worlds equal the CPU checker."""
    from lenia_amd.engine import Engine
    from oracle import oracle as O
    ws = WL.mixed(n_worlds=16, base_seed=300, asymptotic=ASYM)
    eng = Engine()
    for w in ws:
        eng.add_world(w["shape"], 3, w["kernels"], w["cfg"])
    eng.commit()
    for i, w in enumerate(ws):
        eng.set(i, "A", w["A"])
    eng.step(6)
    eng.synchronize()
    batch = [eng.get(i, "A") for i in range(len(ws))]
    eng.close()
    for i in (0, 1, 2, 3, 5, 10):
        w = ws[i]
        ow = O.World(w["shape"], 3, w["kernels"], w["cfg"], 32)
        ow.set(O.F_A, w["A"])
        ow.step(6)
        bad, err = mismatch(batch[i], ow.get(O.F_A))
        assert bad == 0, (i, w["shape"], w["cfg"]["mode"], bad, err)


def test_plain_and_full_gradient_sweeps_in_one_class():
    """A shape class whose worlds are all plain (Sobel, SRT or RT, no MU_VEL, V not live, no
    wall, a power-of-two theta_A) runs the gradient sweep compiled without the other paths;
    one world needing any of them sends its whole class through the full sweep.  Each
    batch below holds a plain class of 24^3 and a class of 16^3 that mixes plain worlds
    with one world of each kind that is not; every world equals the checker."""
    from lenia_amd.engine import Engine
    from oracle import oracle as O
    plain = [WL.world((24, 24, 24), 3, 9, 6, seed=700 + i, mode=m) for i, m in enumerate(("rt", "srt", "rt"))]
    extra = (dict(border="wall", sobel_border="wall"), dict(mu_vel=True), dict(theta_a=3.0),
             dict(gradient="central"), dict(mode="adv"))
    for j, over in enumerate(extra):
        mixed = [WL.world((16, 16, 16), 3, 9, 5, seed=720 + 10 * j + i, mode=m) for i, m in enumerate(("rt", "srt"))]
        mixed.append(WL.world((16, 16, 16), 3, 9, 5, seed=729 + 10 * j, **over))
        ws = plain + mixed
        eng = Engine()
        for w in ws:
            eng.add_world(w["shape"], 3, w["kernels"], w["cfg"])
        eng.commit()
        for i, w in enumerate(ws):
            eng.set(i, "A", w["A"])
        eng.step(4)
        eng.synchronize()
        got = [eng.get(i, "A") for i in range(len(ws))]
        eng.close()
        for i, w in enumerate(ws):
            ow = O.World(w["shape"], 3, w["kernels"], w["cfg"], 32)
            ow.set(O.F_A, w["A"])
            ow.step(4)
            bad, err = mismatch(got[i], ow.get(O.F_A))
            assert bad == 0, (over, i, w["shape"], w["cfg"]["mode"], bad, err)


def test_mixed_batch_4096_hundred_steps_one_launch():
    """4096 worlds of the four mixed shapes with per-world SRT/RT, advanced 100 steps by
    one call (one captured-graph launch for all shape classes); worlds of every shape and
    both modes spot-checked against the CPU checker bit for bit, and a graph of 100
    steps equals 100 single eager steps on a small mixed batch."""
    from lenia_amd.engine import Engine
    from oracle import oracle as O
    ws = WL.mixed(n_worlds=4096, base_seed=5000, asymptotic=ASYM)
    eng = Engine()
    for w in ws:
        eng.add_world(w["shape"], 3, w["kernels"], w["cfg"])
    eng.commit()
    for i, w in enumerate(ws):
        eng.set(i, "A", w["A"])
    eng.prepare_steps(100)
    eng.step(100)
    eng.synchronize()
    picks = []
    for s in range(4):  # the first SRT and the first RT world of each shape
        for mode in ("srt", "rt"):
            picks.append(next(i for i in range(s, 4096, 4) if ws[i]["cfg"]["mode"] == mode))
    picks.append(4095)
    got = {i: eng.get(i, "A") for i in picks}
    eng.close()
    O.set_threads(8)
    try:
        for i in picks:
            w = ws[i]
            ow = O.World(w["shape"], 3, w["kernels"], w["cfg"], 32)
            ow.set(O.F_A, w["A"])
            ow.step(100)
            bad, err = mismatch(got[i], ow.get(O.F_A))
            assert bad == 0, (i, w["shape"], w["cfg"]["mode"], bad, err)
    finally:
        O.set_threads(1)
    small = WL.mixed(n_worlds=8, base_seed=5100, asymptotic=ASYM)
    outs = []
    for graph in (True, False):
        eng = Engine(graph=graph)
        for w in small:
            eng.add_world(w["shape"], 3, w["kernels"], w["cfg"])
        eng.commit()
        for i, w in enumerate(small):
            eng.set(i, "A", w["A"])
        if graph:
            eng.step(100)
        else:
            for _ in range(100):
                eng.step(1)
        eng.synchronize()
        outs.append([eng.get(i, "A") for i in range(len(small))])
        eng.close()
    for a, b in zip(*outs):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_step_graph_cache_evicts_and_stays_exact():
    """More distinct step counts than graphs kept (least recently used evicted), odd
    and even counts (the A/S buffer parity flips), against single eager steps."""
    from lenia_amd import _lib as L
    from lenia_amd.engine import Engine
    ws = WL.mixed(n_worlds=4, base_seed=5200, asymptotic=ASYM)
    counts = [3, 1, 4, 2, 5, 7, 6, 8, 3, 1]
    outs = []
    for graph in (True, False):
        eng = Engine(graph=graph)
        for w in ws:
            eng.add_world(w["shape"], 3, w["kernels"], w["cfg"])
        eng.commit()
        for i, w in enumerate(ws):
            eng.set(i, "A", w["A"])
        snaps = []
        for n in counts:
            if graph:
                eng.step(n)
            else:
                for _ in range(n):
                    eng.step(1)
            eng.synchronize()
            snaps.append([eng.get(i, "A") for i in range(len(ws))] + [eng.get(i, "S") for i in range(len(ws))])
        if graph:
            with pytest.raises(L.LeniaHipError, match="nsteps"):
                eng.prepare_steps(0)
            with pytest.raises(L.LeniaHipError, match="nsteps"):
                eng.prepare_steps(257)
        eng.close()
        outs.append(snaps)
    for a_snap, b_snap in zip(*outs):
        for a, b in zip(a_snap, b_snap):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_large_class_runs_as_two_halves():
    """A class of at least 2048 worlds runs as two halves (two branches of the step
    graph): worlds from both halves, mixed SRT and RT, equal the checker after a graph
    launch of several steps, and the graph equals eager stepping bit for bit."""
    from lenia_amd.engine import Engine
    from oracle import oracle as O
    ws = [WL.world((8, 8, 8), 3, 9, 3, seed=900 + i, mode="srt" if i % 3 == 0 else "rt") for i in range(2052)]
    outs = []
    for graph in (True, False):
        eng = Engine(graph=graph)
        for w in ws:
            eng.add_world(w["shape"], 3, w["kernels"], w["cfg"])
        eng.commit()
        for i, w in enumerate(ws):
            eng.set(i, "A", w["A"])
        eng.step(5)
        eng.synchronize()
        outs.append({i: eng.get(i, "A") for i in (0, 1, 1025, 1026, 2050, 2051)})
        eng.close()
    for i, a in outs[0].items():
        assert np.array_equal(a.view(np.uint32), outs[1][i].view(np.uint32)), i
        w = ws[i]
        ow = O.World(w["shape"], 3, w["kernels"], w["cfg"], 32)
        ow.set(O.F_A, w["A"])
        ow.step(5)
        bad, err = mismatch(a, ow.get(O.F_A))
        assert bad == 0, (i, w["cfg"]["mode"], bad, err)


def test_srt_mu_needs_keep_mu():
    """SRT worlds move inside the gradient sweep and do not store mu unless asked."""
    from lenia_amd import _lib as L
    from lenia_amd.engine import Engine
    w = WL.plain32(seed=3, mode="srt")
    eng = Engine()
    wid = eng.add_world(w["shape"], 3, w["kernels"], w["cfg"])
    eng.commit()
    eng.set(wid, "A", w["A"])
    eng.step(1)
    with pytest.raises(L.LeniaHipError, match="LH_OPT_KEEP_MU"):
        eng.get(wid, "MU")
    eng.close()


def test_large_batch_one_thread_per_column_spectral_path():
    """128 worlds of 32^3 (69,632 spectral columns) take spec_z's one-thread-per-column
    schedule; smaller batches take the per-(column, kernel) one.  Spot-checked worlds of
    the large batch equal the CPU checker, and equal themselves run alone."""
    from lenia_amd.engine import Engine
    from oracle import oracle as O
    ws = [WL.plain32(seed=700 + i, asymptotic=ASYM) for i in range(128)]
    eng = Engine()
    for w in ws:
        eng.add_world(w["shape"], 3, w["kernels"], w["cfg"])
    eng.commit()
    for i, w in enumerate(ws):
        eng.set(i, "A", w["A"])
    eng.step(3)
    eng.synchronize()
    batch = {i: eng.get(i, "A") for i in (0, 63, 127)}
    eng.close()
    for i, a in batch.items():
        w = ws[i]
        ow = O.World(w["shape"], 3, w["kernels"], w["cfg"], 32)
        ow.set(O.F_A, w["A"])
        ow.step(3)
        bad, err = mismatch(a, ow.get(O.F_A))
        assert bad == 0, (i, bad, err)
        alone = run_pair(w, 3)["A"][0]
        assert np.array_equal(alone.view(np.uint32), a.view(np.uint32)), i


def test_set_field_midrun_and_state_roundtrip():
    from lenia_amd.engine import Engine
    from oracle import oracle as O
    w = WL.plain32(seed=30, asymptotic=ASYM, mu_vel=True)
    eng = Engine(keep_velocity=True)
    wid = eng.add_world(w["shape"], 3, w["kernels"], w["cfg"])
    eng.commit()
    eng.set(wid, "A", w["A"])
    eng.step(3)
    A3, V3, GA3, S3 = (eng.get(wid, f) for f in ("A", "V", "GA", "S"))
    ow = O.World(w["shape"], 3, w["kernels"], w["cfg"], 32)
    for f, a in ((O.F_A, A3), (O.F_V, V3), (O.F_GA, GA3), (O.F_S, S3)):
        ow.set(f, a)
    eng.set(wid, "A", A3)  # re-set mid-run: spectra are rebuilt from the new state
    eng.step(4)
    ow.step(4)
    for f, of in (("A", O.F_A), ("V", O.F_V), ("GA", O.F_GA), ("S", O.F_S)):
        bad, err = mismatch(eng.get(wid, f), ow.get(of))
        assert bad == 0, (f, bad, err)
    eng.close()


def test_mc_case_index_bitwise():
    from lenia_amd.engine import Engine
    from oracle import oracle as O
    w = WL.plain32(seed=40, asymptotic=ASYM)
    eng = Engine(keep_usum=True)
    wid = eng.add_world(w["shape"], 3, w["kernels"], w["cfg"])
    eng.commit()
    eng.set(wid, "A", w["A"])
    eng.step(5)
    eng.synchronize()
    for field in ("A", "H", "USUM"):
        f = eng.get(wid, field).reshape(3, 32, 32, 32)
        for c in range(3):
            for face_only in (True, False):
                gpu = eng.mc_case(wid, field, c, 0.5, face_only)
                ref = O.mc_case(f[c], 0.5, face_only, 32)
                assert np.array_equal(gpu, ref), (field, c, face_only)
    eng.close()


def test_usum_matches_checker():
    w = WL.plain32(seed=41, asymptotic=ASYM)
    r = run_pair(w, 2, fields=["USUM"], keep_usum=True)
    assert mismatch(*r["USUM"])[0] == 0


def test_rejects_bad_input():
    from lenia_amd import _lib as L
    from lenia_amd.engine import Engine
    eng = Engine()
    w = WL.plain32(seed=1)
    with pytest.raises(L.LeniaHipError, match="not built: sides must be divisible by 4"):
        eng.add_world((30, 30, 30), 3, w["kernels"], w["cfg"])
    with pytest.raises(L.LeniaHipError, match="not built into this library and has no shape module"):
        eng.add_world((60, 8, 60), 3, w["kernels"], w["cfg"])
    bad = [dict(k) for k in w["kernels"]]
    bad[0]["c0"] = 5
    with pytest.raises(L.LeniaHipError, match="channel out of range"):
        eng.add_world((32, 32, 32), 3, bad, w["cfg"])
    with pytest.raises(L.LeniaHipError, match="sigma"):
        eng.add_world((32, 32, 32), 3, w["kernels"], dict(w["cfg"], dd=0.5))
    with pytest.raises(L.LeniaHipError):
        eng.step(1)  # before commit
    wid = eng.add_world((32, 32, 32), 3, w["kernels"], w["cfg"])
    eng.commit()
    with pytest.raises(ValueError):
        eng.set(wid, "A", np.zeros(10, np.float32))
    with pytest.raises(L.LeniaHipError, match="U_SUM"):
        eng.get(wid, "USUM")
    eng.close()


def test_autotuned_and_default_schedules_identical():
    """The autotuner only picks among schedules with the same per-cell arithmetic: a
    mixed batch steps bit-identically with and without it."""
    from lenia_amd.engine import Engine
    ws = WL.mixed(n_worlds=8, base_seed=500, asymptotic=ASYM)
    outs = []
    for tune in (True, False):
        eng = Engine(autotune=tune)
        for w in ws:
            eng.add_world(w["shape"], 3, w["kernels"], w["cfg"])
        eng.commit()
        for i, w in enumerate(ws):
            eng.set(i, "A", w["A"])
        eng.step(5)
        outs.append([eng.get(i, "A") for i in range(len(ws))])
        eng.close()
    for a, b in zip(*outs):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("zs", [256, 16, 8, 4, 2, 1])
def test_every_schedule_candidate_identical(zs):
    """Every schedule the autotuner can pick (gradient and transport sweeps at any segment
    length -- whole depth, halves, quarters, 8, 4, 2, 1 planes -- and the tiled transport)
    steps a mixed batch bit-identically to the launchers' default schedule."""
    from lenia_amd.engine import Engine
    ws = WL.mixed(n_worlds=8, base_seed=510, asymptotic=ASYM)

    def run(schedule):
        eng = Engine(autotune=False, schedule=schedule)
        for w in ws:
            eng.add_world(w["shape"], 3, w["kernels"], w["cfg"])
        eng.commit()
        for i, w in enumerate(ws):
            eng.set(i, "A", w["A"])
        eng.step(3)
        out = [eng.get(i, "A") for i in range(len(ws))]
        eng.close()
        return out

    ref = run(None)
    for schedule in ((zs, zs, False, 1), (zs, 0, True, 2)):
        print(f"schedule {schedule}", flush=True)
        for a, b in zip(ref, run(schedule)):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), schedule
