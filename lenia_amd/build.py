"""Builds liblenia_hip.so (HIP, gfx950) in-tree with hipcc.

Flags that matter for results:
  -ffp-contract=off        no mul+add fusion: every fp32 operation rounds as written,
                           which keeps the step bit-identical to its CPU restatement;
  -packed-fp32-ops off     no v_pk_* fp32 instructions (unreliable on this pool).
Division and sqrt stay IEEE-correct (HIP's default).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liblenia_hip.so")
SOURCES = ["lh_kernels.hip", "lh_kernels_big.hip", "lh_dftgemm.hip", "lh_mlp.hip", "lh_sample.hip", "lh_mesh.hip",
           "lh_engine.hip"]
ARCH = os.environ.get("LENIA_HIP_ARCH", "gfx950")


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def flags() -> list:
    return [f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", os.path.join(REPO, "include"), "-Wno-unused-result"]


# Shapes outside the library's built-in set that the GPU tests step: built as shape
# modules alongside the library.
TEST_SHAPES = ((28, 28, 28), (20, 12, 20), (52, 52, 52))
SHAPE_DIR = os.path.join(HERE, "shapes")


def shape_module_path(W: int, H: int, D: int, out_dir: str = SHAPE_DIR) -> str:
    return os.path.join(out_dir, f"liblenia_shape_{W}x{H}x{D}.so")


def _shape_module_cmd(W: int, H: int, D: int, out: str) -> list:
    for n in (W, H, D):
        if n % 4 or not 4 <= n <= 128:
            raise ValueError(f"shape {W}x{H}x{D}: every side must be divisible by 4 and within 4..128")
    return [hipcc()] + flags() + [f"-DLH_MODULE_W={W}", f"-DLH_MODULE_H={H}", f"-DLH_MODULE_D={D}", "-shared",
                                  os.path.join(CSRC, "lh_shape_module.hip"), "-o", out]


def build_shape(W: int, H: int, D: int, out_dir: str = SHAPE_DIR, verbose: bool = True) -> str:
    """Compiles the phase kernels of one world shape into a module the engine loads when a
    world of that shape is added (sides divisible by 4, 4..128)."""
    os.makedirs(out_dir, exist_ok=True)
    out = shape_module_path(W, H, D, out_dir)
    t0 = time.time()
    subprocess.run(_shape_module_cmd(W, H, D, out + ".tmp"), check=True)
    os.replace(out + ".tmp", out)
    if verbose:
        print(f"built {out} in {time.time() - t0:.1f}s")
    return out


def build(verbose: bool = True) -> str:
    objs = []
    t0 = time.time()
    procs = []
    os.makedirs(SHAPE_DIR, exist_ok=True)
    for (W, H, D) in TEST_SHAPES:
        out = shape_module_path(W, H, D)
        cmd = _shape_module_cmd(W, H, D, out + ".tmp")
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    for src in SOURCES:
        obj = os.path.join(CSRC, "_build", src.replace(".hip", ".o"))
        os.makedirs(os.path.dirname(obj), exist_ok=True)
        cmd = [hipcc()] + flags() + ["-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out.decode(errors="replace"))
            raise RuntimeError("hipcc failed: " + " ".join(cmd))
    for (W, H, D) in TEST_SHAPES:
        out = shape_module_path(W, H, D)
        os.replace(out + ".tmp", out)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_shim()
    if verbose:
        print(f"built {LIB} in {time.time() - t0:.1f}s")
    return LIB


SHIM = os.path.join(HERE, "liblenia_ref_shim.so")
HARNESS = os.path.join(HERE, "ref_harness")


def build_shim() -> None:
    """The reference-signature shim (host C++ over the C-ABI) and its 2-thread harness."""
    inc = ["-I", os.path.join(REPO, "include")]
    common = ["-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-Wall"] + inc
    subprocess.run(["g++"] + common + ["-shared", "-o", SHIM + ".tmp",
                                       os.path.join(HERE, "shim", "lenia_ref_shim.cpp"),
                                       "-L", HERE, "-llenia_hip", "-Wl,-rpath,$ORIGIN"], check=True)
    os.replace(SHIM + ".tmp", SHIM)
    subprocess.run(["g++"] + common + ["-o", HARNESS + ".tmp", os.path.join(HERE, "shim", "ref_harness.cpp"),
                                       "-L", HERE, "-llenia_ref_shim", "-llenia_hip", "-lpthread",
                                       "-Wl,-rpath,$ORIGIN"], check=True)
    os.replace(HARNESS + ".tmp", HARNESS)


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="Build liblenia_hip.so, or a shape module (--shape WxHxD).")
    ap.add_argument("--shape", action="append", default=[],
                    help="build the module of a world shape not built into the library, e.g. 28x28x28")
    ap.add_argument("--out-dir", default=SHAPE_DIR, help="where shape modules go (the engine also searches "
                                                          "$LENIA_SHAPE_DIR)")
    args = ap.parse_args()
    if args.shape:
        for sh in args.shape:
            W, H, D = (int(v) for v in sh.lower().split("x"))
            build_shape(W, H, D, args.out_dir)
    else:
        build()
