"""Seeded synthetic worlds for the Flow Lenia step.

A counter-based generator (splitmix64) materialises every random draw, so the C++
harness, the HIP engine and the CPU checker all consume identical arrays.  Draws follow
the workloads the engine is measured on:

* ``plain32``   one 32^3 world, C=3, K=9 (c0=k//3, c1=k%3), R=8, 1-3 rings, periodic
                border, Sobel gradients, RT reintegration, dt=0.2, sigma=0.65.
* ``big96``     one 96^3 world, C=4, K=16, R=12, zero border, central differences with
                the blur pass, advection with damping 0.9.
* ``mixed``     many worlds of shapes 24^3 / 32^3 / 36^3 / 44x24x44, params as plain32,
                per-world SRT or RT.
* ``precision`` one 128^3 world, C=3, K=9, R=16, analytic growth.

Kernel parameter ranges follow the reference's randomiser where the workload does not
fix them (lenia.cpp:370-429).
"""
from __future__ import annotations

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n consecutive outputs of splitmix64 started at `seed` (uint64)."""
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


class Stream:
    """Sequential uniform draws; floats are exact multiples of 2^-24 in [0, 1)."""

    def __init__(self, seed: int):
        self.seed = int(seed)
        self.pos = 0

    def u32(self, n: int) -> np.ndarray:
        z = splitmix64(self.seed + self.pos * 0x9E3779B97F4A7C15, n)
        self.pos += n
        return (z >> np.uint64(40)).astype(np.float32) * np.float32(2.0 ** -24)

    def uniform(self, lo: float, hi: float, n: int = 1) -> np.ndarray:
        return (np.float32(lo) + self.u32(n) * np.float32(hi - lo)).astype(np.float32)

    def integers(self, lo: int, hi: int, n: int = 1) -> np.ndarray:
        """Integers in [lo, hi]."""
        return (lo + np.floor(self.u32(n).astype(np.float64) * (hi - lo + 1))).astype(np.int64)


def base_cfg(**over) -> dict:
    cfg = dict(mode="rt", border="torus", sobel_border="torus", gradient="sobel", growth="rt",
               mu_vel=False, adv_vel=True, blur=False, dt=0.2, sigma=0.65, dd=1.0, s_norm=0.0,
               theta_a=2.0, gta=0.1, sta=0.1, v_damp=0.9, asymptotic=[0, 0, 0, 0],
               soft_clip=[0, 0, 0, 0], quirks=0x7FF)
    cfg.update(over)
    return cfg


def draw_kernels(st: Stream, C: int, K: int, radius: int, m_rng=(0.1, 0.4), s_rng=(0.02, 0.1),
                 h_rng=(0.2, 1.0)) -> list:
    ks = []
    for k in range(K):
        beta = int(st.integers(1, 3)[0])
        b = [float(v) for v in st.uniform(0.0, 1.0, beta)]
        ks.append(dict(c0=(k // C) % C, c1=k % C, radius=radius, beta_size=beta, b=b, r_a=1.0,
                       m=float(st.uniform(*m_rng)[0]), s=float(st.uniform(*s_rng)[0]),
                       h=float(st.uniform(*h_rng)[0])))
    return ks


def box_init(st: Stream, shape, C: int) -> np.ndarray:
    """U(0,1) noise in a centred box of half the extents, zero elsewhere: [C][D][H][W]."""
    W, H, D = shape
    A = np.zeros((C, D, H, W), dtype=np.float32)
    z0, y0, x0 = D // 4, H // 4, W // 4
    n = (D // 2) * (H // 2) * (W // 2)
    for c in range(C):
        A[c, z0:z0 + D // 2, y0:y0 + H // 2, x0:x0 + W // 2] = st.u32(n).reshape(D // 2, H // 2, W // 2)
    return A


def world(shape=(32, 32, 32), C=3, K=9, radius=8, seed=0, **cfg_over) -> dict:
    st = Stream(seed)
    kernels = draw_kernels(st, C, K, radius)
    A = box_init(st, shape, C)
    return dict(shape=tuple(shape), C=C, kernels=kernels, cfg=base_cfg(**cfg_over), A=A)


def plain32(seed=0, **over) -> dict:
    return world((32, 32, 32), 3, 9, 8, seed, **over)


def big96(seed=0, **over) -> dict:
    o = dict(mode="adv", border="none", sobel_border="none", gradient="central", blur=True,
             adv_vel=True, v_damp=0.9)
    o.update(over)
    return world((96, 96, 96), 4, 16, 12, seed, **o)


def precision128(seed=0, **over) -> dict:
    o = dict(quirks=0x7FF & ~0x40)
    o.update(over)
    return world((128, 128, 128), 3, 9, 16, seed, **o)


MIXED_SHAPES = ((24, 24, 24), (32, 32, 32), (36, 36, 36), (44, 24, 44))


def mixed(n_worlds=4096, base_seed=1000, **over) -> list:
    """Equal shares of the four shapes; per-world SRT or RT with p=1/2."""
    out = []
    for i in range(n_worlds):
        shape = MIXED_SHAPES[i % 4]
        st = Stream(base_seed + i)
        mode = "srt" if st.u32(1)[0] < 0.5 else "rt"
        kernels = draw_kernels(st, 3, 9, 8)
        A = box_init(st, shape, 3)
        cfg = base_cfg(mode=mode, **over)
        out.append(dict(shape=shape, C=3, kernels=kernels, cfg=cfg, A=A))
    return out
