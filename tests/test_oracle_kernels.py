"""Pins the checker's kernel construction, growth table and blur kernel.

References: init_kernel lenia.cpp:199-261 (ring profile, torus placement, sum
normalisation), bell render.h:674-677, init_bell_lut lenia.cpp:93-101 with bell_lookup
lenia.h:79-84, precompute_blur_kernel reintegration.cpp:13-78.
"""
import math

import numpy as np
import pytest

from lenia_amd import workloads as WL


def ring_value(k, r):
    """Closed form of one kernel cell at radius r, before normalisation."""
    d = r / k["radius"] * k["beta_size"] / k["r_a"]
    if d >= k["beta_size"]:
        return 0.0
    idx = min(int(math.floor(d)), k["beta_size"] - 1)
    fr = d - math.floor(d)
    return k["b"][idx] * math.exp(-(((fr - 0.5) / 0.15) ** 2) / 2)


@pytest.mark.parametrize("shape,radius", [((32, 32, 32), 8), ((24, 24, 24), 6), ((44, 24, 44), 8),
                                          ((12, 12, 12), 3)])
def test_kernel_field_closed_form(oracle, shape, radius):
    st = WL.Stream(11)
    k = WL.draw_kernels(st, 3, 1, radius)[0]
    W, H, D = shape
    f = oracle.kernel_field(shape, k, 64)
    assert np.isclose(f.sum(), 1.0, atol=1e-12)
    raw = np.zeros((D, H, W))
    for z in range(-D // 2, D // 2):
        for y in range(-H // 2, H // 2):
            for x in range(-W // 2, W // 2):
                raw[z % D, y % H, x % W] = ring_value(k, math.sqrt(x * x + y * y + z * z))
    assert np.allclose(f, raw / raw.sum(), atol=1e-12)
    f32 = oracle.kernel_field(shape, k, 32)
    assert np.allclose(f32, f, rtol=1e-5, atol=1e-9)


def test_kernel_even_and_real_spectrum(oracle):
    st = WL.Stream(5)
    k = WL.draw_kernels(st, 3, 1, 8)[0]
    f = oracle.kernel_field((32, 32, 32), k, 64)
    # torus-even along each axis: f(x) = f(-x mod n)
    assert np.array_equal(f, np.roll(f[::-1], 1, axis=0))
    assert np.array_equal(f, np.roll(f[:, ::-1], 1, axis=1))
    assert np.array_equal(f, np.roll(f[:, :, ::-1], 1, axis=2))
    full = np.fft.rfftn(f)
    assert np.abs(full.imag).max() < 1e-12
    oct_ = oracle.kernel_spectrum(f, 64)
    assert oct_.shape == (17, 17, 17)
    assert np.isclose(oct_[0, 0, 0], 1.0)                       # K(0) = sum = 1
    assert np.allclose(oct_, full.real[:17, :17, :], atol=1e-12)
    # mirrored bins equal the octant
    assert np.allclose(full.real[32 - 5, 32 - 7, 3], oct_[5, 7, 3], atol=1e-12)


@pytest.mark.parametrize("m,s", [(0.3, 0.15), (0.1, 0.02), (0.4, 0.1), (0.23, 0.125)])
def test_bell_table_closed_form(oracle, m, s):
    lut = oracle.bell_table(m, s, 32)
    im, is_ = int(np.float32(m) * np.float32(255)), int(np.float32(s) * np.float32(255))
    u = np.arange(256) / 255.0
    ref = np.exp(-(((u - im / 255.0) / (is_ / 255.0)) ** 2) / 2)
    assert np.allclose(lut, ref, rtol=2e-6, atol=1e-7)
    assert lut[im] == 1.0                                       # the bell peaks at its own bin


def test_blur_field(oracle):
    f = oracle.blur_field((16, 16, 16), 64)
    assert np.isclose(f.sum(), 1.0)
    g = np.exp(-np.array([0, 1, 2, 3, 4, 5, 6, 8, 12]) / 3.375)
    assert np.isclose(f[0, 0, 1] / f[0, 0, 0], g[1] / g[0], rtol=1e-6)
    assert np.isclose(f[2, 2, 2] / f[0, 0, 0], g[-1] / g[0], rtol=1e-6)
    assert f[3, 0, 0] == 0 and f[0, 0, 14] > 0                  # support |r| <= 2 on the torus


def test_canon_exp_accuracy(oracle):
    """The analytic-growth exp agrees with numpy.exp to ~1 ulp over the bell's range
    (x = -q^2/2 <= 0), is exactly 1 at 0, and its fp32 rounding matches float32(exp(x))."""
    rng = np.random.default_rng(3)
    x = np.concatenate([-rng.uniform(0, 1, 4000) ** 3 * 700, -rng.uniform(0, 104, 4000),
                        [0.0, -0.5, -1e-300, -np.log(2), -699.9, -700.5, -800.0]])
    got = oracle.canon_exp(x)
    ref = np.exp(x)
    assert got[x == 0.0][0] == 1.0
    assert (got[x < -700] == 0).all()
    ok = x >= -700
    rel = np.abs(got[ok] - ref[ok]) / ref[ok]
    assert rel.max() <= 4e-16, rel.max()
    f_ok = ok & (x > -104)
    assert np.count_nonzero(got[f_ok].astype(np.float32) != ref[f_ok].astype(np.float32)) == 0


def test_analytic_growth_is_reference_bell(oracle):
    """Growth without the LUT is bell(u, m, s) (render.h:674-677) per kernel: a world with
    one kernel whose H, read after a step, is h * bell(U) at U = 0 (empty world)."""
    k = dict(c0=0, c1=0, radius=2, beta_size=1, b=[1.0], r_a=1.0, m=0.15, s=0.03, h=0.7)
    w = oracle.World((8, 8, 8), 1, [k], WL.base_cfg(quirks=0, growth="none"), 32)
    w.set(oracle.F_A, np.zeros(512, np.float32))
    w.step(1)
    H = w.get(oracle.F_H)
    q = np.float32((np.float32(0.0) - np.float32(0.15)) / np.float32(0.03))
    g = np.float32(math.exp(-(float(q) * float(q)) / 2.0))
    assert np.all(H == np.float32(g * np.float32(0.7))), (H[0], g * 0.7)
