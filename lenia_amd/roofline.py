"""Algorithmic (compulsory) byte and flop counts of the Flow Lenia step.

The one place these constants live; bench.py and the docs quote them.

Per world of W x H x D cells (N), C channels, K kernels, half spectrum
Nh = (W/2+1) H D complex bins and real kernel octant Noct = (W/2+1)(H/2+1)(D/2+1):

* whole step (compulsory):  8 C N (read + write A) + 4 K Noct (kernel octant)
                            [+ 24 C N when the velocity is live (MU_VEL or ADV)]
  32^3, C=3, K=9: 963,300 B per world-step.
* per kernel: every field the kernel must read or write, each exactly once.
* DFT-as-GEMM convolution (cubic side n): matrix-core flops as issued, with P products
  per multiply-accumulate (3 for the hi+lo split, 1 unsplit):
    fx, ix (real n x n matrix on n^2 rows):          2 n^4 P           per channel / kernel
    fy, fz, iz, iy (complex n x n on n (n/2+1) lines): 8 n^3 (n/2+1) P per channel / kernel
  step: C (fx + fy + fz) + K (iz + iy + ix); 128^3, C=3, K=9, split: 97.8 GFLOP.
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
    if kernel == "growth_fused":
        # spec_z + inv_growth in one launch: the channels' half spectra in, growth out
        return C * Nh * C64 + K * Noct * F32 + K * 256 * F32 + C * N * F32 + C * N * F32 * (2 if growth_rt else 1)
    if kernel == "local":
        v = 6 * C * N * F32 if v_live else 0
        return N * F32 + 2 * C * N * F32 + 3 * C * N * F32 + v
    if kernel == "transport":
        # MU (3C) + source field (C) + GA r/w (2C) + A write (C) + S write
        return (3 * C + C + 2 * C + C) * N * F32 + N * F32
    raise KeyError(kernel)


KNOWN = ("fwd_xy", "spec_z", "inv_growth", "growth_fused", "local", "transport")


def gemm_pass_flops(kernel: str, n: int, split=True) -> float:
    P = 3 if split else 1
    if kernel in ("dft_fx", "dft_ix_growth"):
        return 2.0 * n ** 4 * P
    if kernel in ("dft_fy", "dft_fz", "dft_iz", "dft_iy"):
        return 8.0 * n ** 3 * (n // 2 + 1) * P
    raise KeyError(kernel)


def gemm_step_flops(n: int, C: int, K: int, split=True) -> float:
    f = lambda k: gemm_pass_flops(k, n, split)  # noqa: E731
    return C * (f("dft_fx") + f("dft_fy") + f("dft_fz")) + K * (f("dft_iz") + f("dft_iy") + f("dft_ix_growth"))


def gemm_pass_fields(kernel: str, C: int, K: int) -> int:
    """Fields (channels or kernels) one launch of a pass processes."""
    return C if kernel in ("dft_fx", "dft_fy", "dft_fz") else K
