"""Throughput of the Flow Lenia step on one MI355X.

Headline: world-steps/s of a batch of 32^3 worlds (C=3, K=9, RT reintegration, Sobel,
periodic border, reference growth table), synthetic seeded worlds resident in HBM,
next to the same engine driving a single world.  One JSON line on stdout.

  python bench.py [--steps K] [--warmup W] [--worlds B] [--no-cpu] [--no-profile] [--no-single]
                  [--no-mixed] [--no-precision] [--no-proposal] [--no-checkpoint] [--no-sampler]
                  [--no-mesh] [--no-big] [--no-traffic]

"roofline": the whole step's compulsory bytes (8 C N + 4 K Noct per world-step) over the
dominant kernel's average launch time (HIP events on the engine's stream), against the
8 TB/s HBM peak; "kernel_local" restates it with that kernel's own compulsory bytes.  A
batch of at least 2048 worlds of one shape runs as two halves (two launches of each
phase per step, each stepping half the worlds), so a launch's bytes are half the step's.
"traffic": HBM bytes per launch of that kernel from two rocprofv3 --pmc child runs of
this script (FETCH_SIZE, WRITE_SIZE), started before this process touches the GPU and
scaled by what the counters read for a plain copy of known size.
"cpu_baseline": the fp32 CPU restatement under oracle/ (gcc -O3 -march=native, built on
this host), one batch world per host core; "cpu_single_world": one world stepped by 2
threads (the reference's SIM_THREADS) and by all cores in use.  "parity": the 16 checked
GPU worlds against that restatement, bit for bit; "fp64_error": the same worlds against
the FP64 restatement after 1 and warmup+steps steps, in reference and physics mode.

Other legs: "mixed_batch" (equal shares of 24^3, 32^3, 36^3, 44x24x44, per-world SRT or
RT); "big_96" (one 96^3 world, C=4, K=16, R=12, central differences + blur, advection,
zero border); "precision_128" (one 128^3 world with the fp32 FFT convolution and with
DFT-as-GEMM on the matrix cores, with the matrix-core roofline of its dominant pass);
"proposal" (scoring 10^6 sampled parameter sets in one call, training throughput);
"checkpoint" (save and restore of world images); "random_worlds" (the generate_random
range sampler and the Gaussian seeding); "snapshot_readback" (a snapshot read back while
later steps run); "mesh" (marching-cubes meshes of views 1 and 3 for a 32^3 and a 128^3
world beside the CPU restatement of the mesh thread, parity checked bit for bit).

A step advances every world by one full simulation step (spectral convolution, growth,
gradients, force, mu, reintegration, end-of-step transform).  The timed region replays
the engine's captured hipGraph; per-kernel durations come from a second pass that
brackets each launch with HIP events on the engine's stream.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import multiprocessing as mp
import os
import re
import shutil
import signal
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

from lenia_amd import workloads as WL  # noqa: E402
from lenia_amd.engine import Engine  # noqa: E402
from lenia_amd import roofline as RF  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
FP32_PEAK_TFS = 157.3  # MI355X FP32 vector, published
BF16_PEAK_TFS = 2500.0  # MI355X dense bf16/fp16 matrix, published
# rocprofv3 kernel names of each phase the engine's profile reports
PHASE_KERNELS = {"fwd_xy": r"k_fwd_xy\b", "spec_z": r"k_spec_z(_split)?\b", "inv_growth": r"k_inv_growth\b",
                 "growth_fused": r"k_growth_fused\b",
                 "local": r"k_local_(sweep|tiled)\b", "transport": r"k_transport_(sweep|tiled)\b",
                 "inv_blur": r"k_inv_blur\b"}
COPY_MIB = 256  # calibration copy of the traffic probe


def traffic_probe(n_worlds):
    """Child of measure_traffic, run under rocprofv3 --pmc: the bench workload stepped
    twice, one plain copy of COPY_MIB MiB as a marker and calibration, two more steps."""
    worlds = [WL.plain32(seed=100 + i) for i in range(n_worlds)]
    eng = Engine(graph=False)
    for w in worlds:
        eng.add_world(w["shape"], w["C"], w["kernels"], w["cfg"])
    eng.commit()
    for i, w in enumerate(worlds):
        eng.set(i, "A", w["A"])
    eng.step(2)
    eng.synchronize()
    eng.calibration_copy(COPY_MIB << 20)
    eng.step(2)
    eng.synchronize()
    eng.close()


def _counter_rows(outdir):
    """(dispatch id, kernel name, counter, value) from rocprofv3's counter csv."""
    rows = []
    for path in glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True):
        with open(path, newline="") as fh:
            for r in csv.DictReader(fh):
                keys = {k.lower(): k for k in r}
                did = r.get(keys.get("dispatch_id", ""), r.get(keys.get("correlation_id", ""), "0"))
                rows.append((int(did), r[keys["kernel_name"]], r[keys["counter_name"]],
                             float(r[keys["counter_value"]])))
    return sorted(rows)


def measure_traffic(n_worlds, timeout_s=150):
    """HBM bytes per launch of each step phase, from two rocprofv3 --pmc runs of
    traffic_probe (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each scaled by what the
    counter reads for the plain copy of known size.  Run before this process touches
    the GPU.  Returns (phase -> bytes per launch, notes) or (None, reason)."""
    rp = shutil.which("rocprofv3")
    if not rp:
        return None, "rocprofv3 not found"
    per = {}
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        out = tempfile.mkdtemp(prefix="lh_traffic_", dir="/tmp")
        cmd = [rp, "--pmc", ctr, "--output-format", "csv", "-d", out, "-o", "run", "--",
               sys.executable, os.path.abspath(__file__), "--traffic-probe", "--worlds", str(n_worlds)]
        env = dict(os.environ, TMPDIR="/tmp")
        p = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE,
                             start_new_session=True)
        try:
            _, err = p.communicate(timeout=timeout_s)
        except subprocess.TimeoutExpired:
            os.killpg(p.pid, signal.SIGKILL)
            p.communicate()
            shutil.rmtree(out, ignore_errors=True)
            return None, f"rocprofv3 --pmc {ctr} exceeded {timeout_s}s"
        rows = _counter_rows(out)
        shutil.rmtree(out, ignore_errors=True)
        if not rows:
            tail = err.decode(errors="replace").strip().splitlines()[-1:] if err else []
            return None, f"rocprofv3 --pmc {ctr}: no counter rows (rc {p.returncode}) {' '.join(tail)}"
        per[ctr] = rows
    result, notes = {}, {}
    for ctr, rows in per.items():
        copy = [r for r in rows if re.search(r"k_copy\b", r[1]) and r[2] == ctr]
        if not copy:
            return None, "calibration copy not found in the counter rows"
        cid = copy[-1][0]
        true_kib = COPY_MIB * 1024  # the copy reads and writes COPY_MIB MiB
        scale = true_kib / copy[-1][3]
        notes[ctr + "_scale"] = scale
        after = [r for r in rows if r[0] > cid and r[2] == ctr]
        for phase, pat in PHASE_KERNELS.items():
            v = sum(r[3] for r in after if re.search(pat, r[1]))
            if v:
                result[phase] = result.get(phase, 0.0) + v * scale * 1024 / 2  # two steps
    return result, notes


def build_engine(worlds, profile=False):
    eng = Engine(graph=not profile, profile=profile)
    for w in worlds:
        eng.add_world(w["shape"], w["C"], w["kernels"], w["cfg"])
    eng.commit()
    for i, w in enumerate(worlds):
        eng.set(i, "A", w["A"])
    return eng


def timed(eng, steps, warmup):
    eng.step(warmup)
    # the timed steps are one graph launch; capture it before the clock starts
    if 1 <= steps <= 256:
        eng.prepare_steps(steps)
    eng.synchronize()
    t0 = time.perf_counter()
    eng.step(steps)
    eng.synchronize()
    return time.perf_counter() - t0


def _cpu_worker(args):
    seed, check_steps, budget_s = args
    from oracle import oracle as O
    w = WL.plain32(seed=seed)
    ow = O.World(w["shape"], w["C"], w["kernels"], w["cfg"], 32)
    ow.set(O.F_A, w["A"])
    t0 = time.perf_counter()
    ow.step(check_steps)
    snap = ow.get(O.F_A).astype(np.float32)
    n = check_steps
    while time.perf_counter() - t0 < budget_s:
        ow.step(2)
        n += 2
    return seed, snap, n, time.perf_counter() - t0


def _fp64_worker(args):
    """A of one world after 1 and n steps of the FP64 checker."""
    seed, physics, n = args
    from oracle import oracle as O
    w = fp64_world(seed, physics)
    ow = O.World(w["shape"], w["C"], w["kernels"], w["cfg"], 64)
    ow.set(O.F_A, w["A"])
    ow.step(1)
    a1 = ow.get(O.F_A)
    ow.step(n - 1)
    return seed, physics, a1, ow.get(O.F_A)


def fp64_world(seed, physics):
    """A checked world of the batch; physics=True: the same draws with the reference
    quirks off (analytic growth bell, mass-conserving flow driven by growth)."""
    return WL.plain32(seed=seed, quirks=0, growth="none") if physics else WL.plain32(seed=seed)


def cpu_baseline(seeds, check_steps, budget_s=12.0):
    """The fp32 CPU restatement (gcc -O3 -march=native, built on this host) on the host
    cores: world i of the GPU batch on core i, each stepped for ~budget_s.  Each worker
    also returns its world's A after the bench's warmup+timed step count, which the GPU
    result is compared with bitwise."""
    cores = max(1, min(16, os.cpu_count() or 1, len(seeds)))
    with mp.get_context("fork").Pool(cores) as pool:
        res = pool.map(_cpu_worker, [(seeds[i], check_steps, budget_s) for i in range(cores)])
    total_steps = sum(r[2] for r in res)
    wall = max(r[3] for r in res)
    base = dict(value=total_steps / wall, unit="world-steps/s", cores=cores, kind="port",
                sample=f"{cores} of the batch's 32^3 C=3 K=9 RT worlds, one per host core, "
                       f"~{budget_s:.0f}s each (fp32 restatement, gcc -O3 -march=native, one thread per world)")
    return base, {r[0]: r[1] for r in res}


def cpu_single_world(threads_list, budget_s=5.0):
    """Config 0 on the CPU restatement: one 32^3 C=3 K=9 world stepped by 2 threads (the
    reference's SIM_THREADS) and by all host cores in use, ~budget_s each."""
    from oracle import oracle as O
    out = {}
    w = WL.plain32(seed=100)
    for th in threads_list:
        O.set_threads(th)
        ow = O.World(w["shape"], w["C"], w["kernels"], w["cfg"], 32)
        ow.set(O.F_A, w["A"])
        ow.step(1)
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < budget_s:
            ow.step(2)
            n += 2
        dt = time.perf_counter() - t0
        out[f"threads_{th}"] = {"value": n / dt, "unit": "world-steps/s", "threads": th,
                                "ms_per_step": dt * 1e3 / n}
    O.set_threads(1)
    return out


def fp64_error_leg(seeds, n_steps):
    """Per-channel max-abs error of A against the FP64 checker after 1 and n_steps steps,
    for the batch's checked worlds in the reference mode (growth table) and in physics
    mode (analytic bell, mass-conserving flow), with the physics mode's FP64 mass drift."""
    gpu = {}
    for physics in (False, True):
        ws = [fp64_world(s, physics) for s in seeds]
        eng = build_engine(ws)
        m0 = eng.mass()
        eng.step(1)
        a1 = [eng.get(i, "A") for i in range(len(ws))]
        eng.step(n_steps - 1)
        an = [eng.get(i, "A") for i in range(len(ws))]
        m1 = eng.mass()
        eng.close()
        gpu[physics] = (a1, an, float(np.max(np.abs(m1 - m0) / m0)))
    cores = max(1, min(16, os.cpu_count() or 1))
    jobs = [(s, ph, n_steps) for ph in (False, True) for s in seeds]
    with mp.get_context("fork").Pool(cores) as pool:
        res = pool.map(_fp64_worker, jobs)
    out = {}
    for physics in (False, True):
        e1 = np.zeros(3)
        en = np.zeros(3)
        for k, s in enumerate(seeds):
            r = next(r for r in res if r[0] == s and r[1] == physics)
            g1 = gpu[physics][0][k].reshape(3, -1).astype(np.float64)
            gn = gpu[physics][1][k].reshape(3, -1).astype(np.float64)
            e1 = np.maximum(e1, np.abs(g1 - r[2].reshape(3, -1)).max(axis=1))
            en = np.maximum(en, np.abs(gn - r[3].reshape(3, -1)).max(axis=1))
        key = "physics" if physics else "reference"
        out[key] = {"worlds": len(seeds), "steps": [1, n_steps],
                    "max_abs_per_channel": [e1.tolist(), en.tolist()],
                    "worst_after_1": float(e1.max()), f"worst_after_{n_steps}": float(en.max())}
        if physics:
            out[key]["fp64_mass_drift_rel"] = gpu[physics][2]
    out["note"] = ("reference mode reads the growth from an 8-bit table, so a 1-ulp difference of u at a "
                   "bin edge moves H by a whole bin; both modes are chaotic, so differences grow with steps")
    return out


def precision_leg(steps, warmup):
    """One 128^3 world (C=3, K=9, R=16, analytic growth bell, mass-conserving flow), its
    growth convolution on the fp32 FFT path and as DFT-as-GEMM on the matrix cores."""
    w = WL.world((128, 128, 128), 3, 9, 16, seed=128, quirks=0, growth="none")
    paths = {}
    for conv in ("fft", "bf16x3", "fp16x3"):
        eng = Engine()
        eng.add_world(w["shape"], 3, w["kernels"], w["cfg"], conv=conv)
        eng.commit()
        eng.set(0, "A", w["A"])
        dt = timed(eng, steps, warmup)
        eng.close()
        paths[conv] = {"value": steps / dt, "unit": "world-steps/s", "ms_per_step": dt * 1e3 / steps}
    peng = Engine(graph=False, profile=True)
    peng.add_world(w["shape"], 3, w["kernels"], w["cfg"], conv="bf16x3")
    peng.commit()
    peng.set(0, "A", w["A"])
    peng.step(warmup)
    peng.synchronize()
    peng.profile_reset()
    peng.step(steps)
    peng.synchronize()
    prof = peng.profile()
    peng.close()
    gemm = {k: v for k, v in prof.items() if k.startswith("dft_")}
    dom = max(gemm, key=lambda k: gemm[k][1])
    n, ms = gemm[dom]
    avg_s = ms / n * 1e-3
    flops = RF.gemm_pass_flops(dom, 128) * RF.gemm_pass_fields(dom, 3, 9)
    achieved = flops / avg_s / 1e12
    return {"workload": "1 world 128^3, C=3, K=9, R=16, analytic growth, physics-mode flow",
            "paths": paths,
            "gemm_over_fft": paths["bf16x3"]["value"] / paths["fft"]["value"],
            "roofline": {"bound": "mfma", "kernel": dom, "achieved": achieved, "peak": BF16_PEAK_TFS,
                         "unit": "TFLOP/s", "frac": achieved / BF16_PEAK_TFS, "traffic": None,
                         "flops_per_launch": flops, "avg_launch_us": avg_s * 1e6},
            "conv_flops_per_step": RF.gemm_step_flops(128, 3, 9),
            "kernels": {k: {"launches": c, "avg_us": t * 1e3 / c} for k, (c, t) in prof.items()}}


def big_leg(steps, warmup):
    """One 96^3 world (C=4, K=16, R=12, central differences with the blur pass, advection
    with velocity damping, zero border): the FFT, gradient and advection kernels at
    HBM-resident size, with the whole step's compulsory bytes over its dominant kernel."""
    w = WL.big96(seed=0)
    eng = build_engine([w])
    dt = timed(eng, steps, warmup)
    mass = float(eng.mass()[0])
    eng.close()
    peng = build_engine([w], profile=True)
    peng.step(warmup)
    peng.synchronize()
    peng.profile_reset()
    peng.step(steps)
    peng.synchronize()
    prof = peng.profile()
    peng.close()
    dom = max(prof, key=lambda k: prof[k][1])
    n, ms = prof[dom]
    avg_s = ms / n * 1e-3
    step_bytes = RF.step_bytes((96, 96, 96), 4, 16, v_live=True)
    achieved = step_bytes / avg_s / 1e9
    return {"workload": "1 world 96^3, C=4, K=16, R=12, central differences + blur, advection (V, damping 0.9), "
                        "zero border, growth table",
            "value": steps / dt, "unit": "world-steps/s", "ms_per_step": dt * 1e3 / steps,
            "mass_finite": bool(np.isfinite(mass)),
            "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                         "algorithmic_bytes_per_step": step_bytes, "avg_launch_us": avg_s * 1e6},
            "step_roofline": {"achieved_GBs": step_bytes / dt * steps / 1e9,
                              "frac": step_bytes / dt * steps / 1e9 / HBM_PEAK_GBS},
            "kernels": {k: {"launches": c, "avg_us": t * 1e3 / c} for k, (c, t) in prof.items()}}


def proposal_leg():
    """The proposal MLP: one call scoring 10^6 parameter sets sampled on the GPU, and
    Adam steps of batch 256 on resident synthetic rows (throughput only)."""
    from lenia_amd import mlp as M
    from lenia_amd import proposal as P
    net = M.MLP(max_batch=65536)
    net.set_params(M.init_params(0))
    net.score_sampled(1, 65536)
    t0 = time.perf_counter()
    net.score_sampled(2, 1_000_000)
    t_score = time.perf_counter() - t0
    x = P.features(P.sample(3, np.arange(16384)))
    y = np.stack([(x[:, 54] > 0.5), x[:, 55], x[:, 58]], axis=1).astype(np.float32)
    net.set_data(x, y)
    net.train(1, 256, 1e-3, seed=0)
    t0 = time.perf_counter()
    net.train(4, 256, 1e-3, seed=1)
    t_train = time.perf_counter() - t0
    net.close()
    achieved = 1e6 * P.FLOPS_PER_SET / t_score / 1e12
    return {"workload": "MLP 81-256-256-256-3 (GELU, bf16 on the matrix cores, fp32 master weights)",
            "score_1e6_ms": t_score * 1e3, "score_sets_per_s": 1e6 / t_score,
            "train_samples_per_s": 4 * 16384 / t_train,
            "roofline": {"bound": "mfma", "kernel": "score_sampled (sampler + 4 GEMM layers)",
                         "achieved": achieved, "peak": BF16_PEAK_TFS, "unit": "TFLOP/s",
                         "frac": achieved / BF16_PEAK_TFS, "traffic": None,
                         "flops_per_set": P.FLOPS_PER_SET}}


def checkpoint_leg(eng, n):
    """Checkpoint images of the first n worlds of a batch engine: save all, restore all
    (host buffers in pageable memory; PCIe-bound), and a check that the restored state
    reads back bit for bit."""
    eng.synchronize()
    t0 = time.perf_counter()
    images = [eng.checkpoint(i) for i in range(n)]
    t_save = time.perf_counter() - t0
    before = eng.get(n - 1, "A")
    t0 = time.perf_counter()
    for i, img in enumerate(images):
        eng.restore(i, img)
    t_load = time.perf_counter() - t0
    total = sum(b.nbytes for b in images)
    return {"worlds": n, "image_bytes": images[0].nbytes, "save_GBs": total / t_save / 1e9,
            "load_GBs": total / t_load / 1e9, "save_ms_per_world": t_save * 1e3 / n,
            "load_ms_per_world": t_load * 1e3 / n,
            "restored_equal": bool(np.array_equal(eng.get(n - 1, "A").view(np.uint32), before.view(np.uint32)))}


def readback_leg(eng, n, steps, step_s):
    """A of n worlds read back from a snapshot while `steps` further steps run: the host
    reads finish without waiting for those steps, which keep their speed."""
    eng.synchronize()
    eng.snapshot(0, 0, n, ("A",))
    t0 = time.perf_counter()
    eng.step(steps)
    for i in range(n):
        eng.snapshot_read(0, i, "A")
    t_read = time.perf_counter() - t0
    eng.synchronize()
    t_all = time.perf_counter() - t0
    nbytes = n * 3 * 32 ** 3 * 4
    return {"worlds": n, "bytes": nbytes, "read_GBs": nbytes / t_read / 1e9, "read_ms": t_read * 1e3,
            "steps_ms_with_reads": t_all * 1e3, "steps_ms_alone": step_s * steps * 1e3,
            "reads_done_before_steps": t_read < t_all}


def sampler_leg(eng, n_worlds):
    """generate_random-range parameter sets sampled on the GPU (10^6 in one call, into
    host lh_kernel buffers: 80 MB) and every world of the batch engine seeded with its set's Gaussian blobs
    (A written, GA, S, V cleared, on the GPU)."""
    from lenia_amd.engine import sample_random_raw
    sample_random_raw(5, 0, 1000, 3)
    t0 = time.perf_counter()
    _, counts = sample_random_raw(5, 0, 1_000_000, 3)
    t_sample = time.perf_counter() - t0
    eng.seed_random(0, n_worlds, 5, 0)
    t0 = time.perf_counter()
    eng.seed_random(0, n_worlds, 5, 0)
    t_seed = time.perf_counter() - t0
    written = n_worlds * 32 ** 3 * (3 + 3 + 1 + 9) * 4
    return {"workload": "generate_random ranges, 3 channels; seeding 32^3 x 3 worlds",
            "sample_sets_per_s": 1e6 / t_sample, "sample_ms_1e6": t_sample * 1e3,
            "mean_kernels_per_set": float(counts.mean()),
            "seed_worlds_per_s": n_worlds / t_seed, "seed_GBs_written": written / t_seed / 1e9}


def mesh_leg(steps=4, reps=20):
    """Marching-cubes meshes of the renderer's views (lh_mesh: vertices, indices and
    normals on the GPU, copied into host buffers) for the reference's 32^3 x 3 world and
    a 128^3 x 3 world, beside the single-threaded CPU restatement of the mesh thread on
    the same fields (parity checked bit for bit)."""
    from oracle import oracle as O
    from lenia_amd.engine import mc_table
    out = {}
    tab = mc_table()
    for n in (32, 128):
        w = WL.world((n, n, n), 3, 9, 8, 300 + n)
        eng = Engine(keep_usum=True)
        wid = eng.add_world(w["shape"], 3, w["kernels"], w["cfg"])
        eng.commit()
        eng.set(wid, "A", w["A"])
        eng.step(steps)
        eng.synchronize()
        fields = [eng.get(wid, f).reshape(3, n, n, n) for f in ("A", "USUM", "H")]
        for view in (1, 3):
            eng.mesh(wid, view=view)
            t0 = time.perf_counter()
            for _ in range(reps):
                v, i = eng.mesh(wid, view=view)
            t_gpu = (time.perf_counter() - t0) / reps
            t0 = time.perf_counter()
            rv, ri = O.mesh(fields[:view], tab)
            t_cpu = time.perf_counter() - t0
            out[f"{n}^3_view{view}"] = {
                "vertices": len(v), "indices": len(i), "ms_per_mesh": t_gpu * 1e3,
                "meshes_per_s": 1.0 / t_gpu,
                "cpu_baseline": {"ms_per_mesh": t_cpu * 1e3, "cores": 1, "kind": "port"},
                "bitwise_equal": bool(np.array_equal(i, ri) and v.tobytes() == rv.tobytes())}
        eng.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--worlds", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-single", action="store_true", help="skip the single-world leg")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the 128^3 leg (fp32 FFT vs DFT-as-GEMM bf16x3 / fp16x3 convolution)")
    ap.add_argument("--no-proposal", action="store_true",
                    help="skip the proposal-network leg (scoring 10^6 sampled sets, training throughput)")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-shape leg (24^3, 32^3, 36^3, 44x24x44, per-world SRT/RT)")
    ap.add_argument("--no-checkpoint", action="store_true",
                    help="skip timing checkpoint images (save and restore of 256 worlds)")
    ap.add_argument("--no-sampler", action="store_true",
                    help="skip timing the random-world sampler (10^6 sets) and seeding (every world)")
    ap.add_argument("--no-mesh", action="store_true",
                    help="skip timing the marching-cubes meshes (32^3 and 128^3, views 1 and 3)")
    ap.add_argument("--no-big", action="store_true",
                    help="skip the 96^3 leg (C=4, K=16, R=12, central differences + blur, advection, zero border)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 --pmc child runs that measure HBM traffic per kernel")
    ap.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run one plain device-to-device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: worlds are independent and would be sharded "
              "world_id mod G across GPUs; multi-GPU runs are not part of this build.", file=sys.stderr)
        sys.exit(2)
    if args.traffic_probe:
        traffic_probe(args.worlds)
        return
    # the counter runs are children started before this process initialises the GPU
    traffic, traffic_notes = (None, "skipped") if args.no_traffic or args.no_profile else measure_traffic(args.worlds)
    if not args.no_cpu:
        from oracle import oracle as O
        O.use_native()  # the checker tuned for this host; the same results as the generic build

    worlds = [WL.plain32(seed=100 + i) for i in range(args.worlds)]
    seeds = [100 + i for i in range(min(args.worlds, 16))]

    eng = build_engine(worlds)
    dt = timed(eng, args.steps, args.warmup)
    value = args.worlds * args.steps / dt
    ms_per_step = dt * 1e3 / args.steps
    gpu_final = {seeds[i]: eng.get(i, "A") for i in range(len(seeds))}
    ckpt = None if args.no_checkpoint else checkpoint_leg(eng, min(args.worlds, 256))
    sampler = None if args.no_sampler else sampler_leg(eng, args.worlds)
    readback = None if args.no_checkpoint else readback_leg(eng, min(args.worlds, 256), args.steps, dt / args.steps)
    eng.close()

    single_rate, sdt = float("nan"), float("nan")
    if not args.no_single:
        single = build_engine(worlds[:1])
        sdt = timed(single, max(args.steps, 100), args.warmup)
        single_rate = max(args.steps, 100) / sdt
        single.close()

    line = {
        "metric": "world-steps/s at 32^3x3ch batched vs single, err vs FP64 <=1e-4",
        "value": value, "unit": "world-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"batched {args.worlds} x 32^3 worlds, C=3, K=9, R=8, RT, Sobel, "
                               f"torus, growth LUT, dt=0.2, sigma=0.65", "worlds": args.worlds,
                   "shape": [32, 32, 32], "channels": 3, "kernels": 9},
        "single_world": {"value": single_rate, "unit": "world-steps/s",
                         "ms_per_step": sdt * 1e3 / max(args.steps, 100)},
        "batched_over_single": value / single_rate,
    }

    if ckpt is not None:
        line["checkpoint"] = ckpt
    if sampler is not None:
        line["random_worlds"] = sampler
    if readback is not None:
        line["snapshot_readback"] = readback

    if not args.no_mixed:
        mixed = WL.mixed(args.worlds, base_seed=1000)
        meng = build_engine(mixed)
        mdt = timed(meng, args.steps, args.warmup)
        mass = meng.mass()
        meng.close()
        line["mixed_batch"] = {
            "workload": f"{args.worlds} worlds, equal shares of 24^3, 32^3, 36^3, 44x24x44, C=3, K=9, "
                        "per-world seeds and SRT/RT choice, all timed steps of every shape class one captured-graph launch",
            "value": args.worlds * args.steps / mdt, "unit": "world-steps/s",
            "ms_per_step": mdt * 1e3 / args.steps,
            "srt_worlds": sum(w["cfg"]["mode"] == "srt" for w in mixed),
            "mass_finite": bool(np.isfinite(mass).all())}
        del mixed

    if not args.no_proposal:
        line["proposal"] = proposal_leg()

    if not args.no_mesh:
        line["mesh"] = mesh_leg()

    if not args.no_precision:
        line["precision_128"] = precision_leg(args.steps, args.warmup)

    if not args.no_profile:
        peng = build_engine(worlds, profile=True)
        peng.step(args.warmup)
        peng.synchronize()
        peng.profile_reset()
        peng.step(args.steps)
        peng.synchronize()
        prof = peng.profile()
        peng.close()
        kernels = {k: {"launches": n, "avg_us": ms * 1e3 / n, "us_per_step": ms * 1e3 / args.steps}
                   for k, (n, ms) in prof.items()}
        dom = max(prof, key=lambda k: prof[k][1])
        n, ms = prof[dom]
        avg_s = ms / n * 1e-3
        # a batch of at least 2048 worlds runs as two halves: each phase launches once per
        # half, each launch over half the worlds
        lps = n / args.steps
        # the whole step's compulsory bytes (8 C N + 4 K Noct per world-step) over the
        # dominant kernel's time; per launch: the share of the worlds the launch steps
        step_bytes = RF.step_bytes((32, 32, 32), 3, 9) * args.worlds
        launch_bytes = step_bytes / lps
        achieved = launch_bytes / avg_s / 1e9
        kb = RF.kernel_bytes(dom, (32, 32, 32), 3, 9, mode="rt", v_live=False) * args.worlds / lps
        line["roofline"] = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                            "traffic": traffic[dom] / lps if traffic and dom in traffic else None,
                            "traffic_source": ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE child runs of this bench, "
                                               "scaled by the counters of a plain copy"
                                               if traffic else traffic_notes),
                            "algorithmic_bytes_per_launch": launch_bytes, "avg_launch_us": avg_s * 1e6,
                            "launches_per_step": lps,
                            "kernel_local": {"bytes_per_launch": kb, "achieved": kb / avg_s / 1e9,
                                             "frac": kb / avg_s / 1e9 / HBM_PEAK_GBS}}
        if traffic:
            per_step = {k: prof[k][0] / args.steps if k in prof else 1.0 for k in traffic}
            line["traffic_per_launch"] = {k: {"bytes": v / per_step[k], "alg_bytes": RF.kernel_bytes(
                k, (32, 32, 32), 3, 9, mode="rt", v_live=False) * args.worlds / per_step[k] if k in RF.KNOWN else None}
                for k, v in traffic.items()}
            line["traffic_per_launch"]["calibration"] = traffic_notes
        step_flops = RF.step_flops((32, 32, 32), 3, 9) * args.worlds
        line["step_valu"] = {"canonical_flops_per_step": step_flops,
                             "achieved_TFs": step_flops / (ms_per_step * 1e-3) / 1e12,
                             "frac_of_published_fp32": step_flops / (ms_per_step * 1e-3) / 1e12 / FP32_PEAK_TFS}
        line["step_roofline"] = {"algorithmic_bytes_per_step": step_bytes,
                                 "achieved_GBs": step_bytes / (ms_per_step * 1e-3) / 1e9,
                                 "frac": step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS}
        line["kernels"] = kernels

    if not args.no_big:
        line["big_96"] = big_leg(args.steps, args.warmup)

    if not args.no_cpu:
        base, cpu_final = cpu_baseline(seeds, args.warmup + args.steps)
        line["cpu_baseline"] = base
        line["cpu_single_world"] = cpu_single_world([2, max(1, min(16, os.cpu_count() or 1))])
        line["fp64_error"] = fp64_error_leg(seeds, args.warmup + args.steps)
        bad = sum(int(np.count_nonzero(gpu_final[s].view(np.uint32) != a.view(np.uint32)))
                  for s, a in cpu_final.items())
        line["parity"] = {"worlds_compared": len(cpu_final), "steps": args.warmup + args.steps,
                          "mismatching_values": bad, "bitwise_equal": bad == 0}
    if args.copy_mib:
        ceng = Engine()
        ceng.calibration_copy(args.copy_mib << 20)
        ceng.close()
        line["copy_calibration_bytes"] = 2 * (args.copy_mib << 20)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
