"""Algorithmic (compulsory) byte and flop counts of the Flow Lenia step.

The one place these constants live; bench.py and the docs quote them.

Per world of W x H x D cells (N), C channels, K kernels, half spectrum
Nh = (W/2+1) H D complex bins and real kernel octant Noct = (W/2+1)(H/2+1)(D/2+1):

* whole step (compulsory):  8 C N (read + write A) + 4 K Noct (kernel octant)
                            [+ 24 C N when the velocity is live (MU_VEL or ADV)]
  32^3, C=3, K=9: 963,300 B per world-step.
* per kernel: every field the kernel must read or write, each exactly once.
"""
from __future__ import annotations

F32, C64 = 4, 8


def sizes(shape):
    W, H, D = shape
    N = W * H * D
    Nh = (W // 2 + 1) * H * D
    Noct = (W // 2 + 1) * (H // 2 + 1) * (D // 2 + 1)
    return N, Nh, Noct


def step_bytes(shape, C, K, v_live=False) -> int:
    N, _, Noct = sizes(shape)
    return 8 * C * N + 4 * K * Noct + (24 * C * N if v_live else 0)


def step_flops(shape, C, K) -> float:
    """Canonical flop count per world-step (real FFTs at 2.5 N log2 N each, growth,
    separable gradients, alpha/F, mu/V, separable RT gather, TRM)."""
    import math
    N, Nh, _ = sizes(shape)
    return (2.5 * N * math.log2(N) * (C + K) + 2 * K * Nh + 8 * K * N + 27 * (C + 1) * N
            + 12 * C * N + 18 * C * N + 153 * C * N + 4 * C * N)


def kernel_bytes(kernel: str, shape, C, K, mode="rt", v_live=False, growth_rt=True) -> int:
    N, Nh, Noct = sizes(shape)
    if kernel in ("fwd_xy",):
        return C * N * F32 + C * Nh * C64
    if kernel == "spec_z":
        return C * Nh * C64 + K * Noct * F32 + K * Nh * C64
    if kernel == "inv_growth":
        return K * Nh * C64 + K * 256 * F32 + C * N * F32 + C * N * F32 * (2 if growth_rt else 1)
    if kernel == "local":
        v = 6 * C * N * F32 if v_live else 0
        return N * F32 + 2 * C * N * F32 + 3 * C * N * F32 + v
    if kernel == "transport":
        # MU (3C) + source field (C) + GA r/w (2C) + A write (C) + S write
        return (3 * C + C + 2 * C + C) * N * F32 + N * F32
    raise KeyError(kernel)
