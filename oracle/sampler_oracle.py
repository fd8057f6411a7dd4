"""CPU restatement of the batched random-world sampler (TEST INFRASTRUCTURE ONLY).

Only tests/ may import this module; the product (lenia_amd/) never does.

generate_random (lenia.cpp:370-429) and seed_grid_gaussian (lenia.cpp:305-334) with the
reference's rand() replaced by splitmix64 draws u32(seed, set, slot) =
mix64(seed*G + set*256 + slot + G) >> 32 (the convention of lenia_amd/csrc/lh_sample.hip):
neighbourhood 1 + u%8, kernel count u%10 + 1, per kernel (slots 2 + 12 i + ...): rings
u%2 + 1, peaks (float)(u%100)/100, m = (float)(100 + u%500)/1000, s = m/10 +
(float)(u%200)/1000, c0/c1 = u%channels (0 for kernel 0 or one channel), h = r_a = 1;
five blobs (slots 128 + 3 b + axis) with centres u%W, u%H, u%D, each cell of every
channel summing exp(-d^2/(2 r^2)) in blob order, d the plain (unwrapped) distance and
the exponential the checker's canonical exp.  float32 arithmetic throughout, as the
reference's float expressions.  Pinned by the splitmix64 reference outputs, the value
ranges (tests/test_sampler_oracle.py) and a direct per-cell evaluation.
"""
from __future__ import annotations

import numpy as np

from . import oracle as O

GOLD = 0x9E3779B97F4A7C15
MASK = 0xFFFFFFFFFFFFFFFF
BLOBS = 5
f32 = np.float32


def u32(seed: int, sets, slot: int) -> np.ndarray:
    sets = np.asarray(sets, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64((seed * GOLD) & MASK) + sets * np.uint64(256) + np.uint64(slot) + np.uint64(GOLD)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(32)).astype(np.uint64)


def sample_random(seed: int, first: int, n: int, channels: int) -> list:
    """Kernel dicts of sets first .. first+n-1 (generate_random lenia.cpp:372-412)."""
    out = []
    for sset in range(first, first + n):
        nb = 1 + int(u32(seed, sset, 0) % 8)
        kc = int(u32(seed, sset, 1) % 10) + 1
        ks = []
        for i in range(kc):
            base = 2 + 12 * i
            beta = int(u32(seed, sset, base) % 2) + 1
            b = [float(f32(int(u32(seed, sset, base + 1 + e) % 100)) / f32(100)) for e in range(beta)]
            m = f32(100 + int(u32(seed, sset, base + 3) % 500)) / f32(1000)
            s = m / f32(10) + f32(int(u32(seed, sset, base + 4) % 200)) / f32(1000)
            if channels == 1 or i == 0:
                c0 = c1 = 0
            else:
                c0 = int(u32(seed, sset, base + 5) % channels)
                c1 = int(u32(seed, sset, base + 6) % channels)
            ks.append(dict(c0=c0, c1=c1, radius=nb, beta_size=beta, b=b, r_a=1.0, m=float(m), s=float(s), h=1.0))
        out.append(ks)
    return out


def seed_state(seed: int, sset: int, shape, C: int) -> np.ndarray:
    """A [C][N] of set sset (seed_grid_gaussian lenia.cpp:305-334, r = neighbourhood)."""
    W, H, D = shape
    r = 1 + int(u32(seed, sset, 0) % 8)
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    den = f32(2 * r * r)
    a = np.zeros((D, H, W), dtype=np.float32)
    for b in range(BLOBS):
        cx = int(u32(seed, sset, 128 + 3 * b) % W)
        cy = int(u32(seed, sset, 129 + 3 * b) % H)
        cz = int(u32(seed, sset, 130 + 3 * b) % D)
        dist = ((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2).astype(np.float32)
        q = -dist / den
        uq, inv = np.unique(q, return_inverse=True)
        v = O.canon_exp(uq.astype(np.float64)).astype(np.float32)[inv].reshape(q.shape)
        a = a + v
    return np.tile(a.reshape(-1), C)
