"""Pins the checker's FFTs (fft1 / rfft3 / irfft3) against numpy and closed forms.

The reference calls FFTW (fftwf_plan_dft_3d, lenia.cpp:193-194; executed at
lenia.cpp:259, 488, 1043); FFTW is not vendored and not installed here, so the
restatement's transforms are pinned by the mathematics instead.
"""
import numpy as np
import pytest

SIZES = [2, 3, 4, 5, 6, 7, 8, 9, 11, 12, 13, 16, 17, 24, 28, 31, 32, 36, 40, 44, 48, 52, 64, 96, 124, 128]
SHAPES = [(4, 2, 4), (8, 4, 8), (8, 8, 8), (12, 12, 12), (16, 8, 16), (24, 24, 24), (32, 32, 32),
          (36, 36, 36), (44, 24, 44), (48, 48, 48)]


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("prec,tol", [(32, 2e-6), (64, 1e-13)])
def test_fft1_matches_numpy(oracle, n, prec, tol):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    scale = np.sqrt(n) * 4
    assert np.abs(oracle.fft1(x, False, prec) - np.fft.fft(x)).max() <= tol * scale
    assert np.abs(oracle.fft1(x, True, prec) - np.fft.ifft(x) * n).max() <= tol * scale


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 7, 8, 9, 11, 12, 13, 28])
def test_fft1_brute_force_dft(oracle, n):
    rng = np.random.default_rng(100 + n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    k = np.arange(n)
    M = np.exp(-2j * np.pi * np.outer(k, k) / n)
    assert np.allclose(oracle.fft1(x, False, 64), M @ x, atol=1e-12)
    assert np.allclose(oracle.fft1(x, True, 64), M.conj() @ x, atol=1e-12)


@pytest.mark.parametrize("shape", SHAPES)
def test_rfft3_matches_numpy(oracle, shape):
    W, H, D = shape
    f = np.random.default_rng(W * H + D).random((D, H, W))
    for prec, tol in ((32, 3e-6), (64, 1e-13)):
        F = oracle.rfft3(f, prec)
        ref = np.fft.rfftn(f)
        assert F.shape == ref.shape
        assert np.abs(F - ref).max() <= tol * np.abs(ref).max()


@pytest.mark.parametrize("shape", SHAPES)
def test_irfft3_inverts(oracle, shape):
    W, H, D = shape
    f = np.random.default_rng(7 + W).random((D, H, W))
    for prec, tol in ((32, 2e-6), (64, 1e-13)):
        g = oracle.irfft3(np.fft.rfftn(f), W, prec) / f.size
        assert np.abs(g - f).max() <= tol * 8


def test_brute_force_3d_tiny(oracle):
    W, H, D = 4, 2, 4
    f = np.random.default_rng(1).random((D, H, W))
    out = np.zeros((D, H, W // 2 + 1), complex)
    for kz in range(D):
        for ky in range(H):
            for kx in range(W // 2 + 1):
                for z in range(D):
                    for y in range(H):
                        for x in range(W):
                            out[kz, ky, kx] += f[z, y, x] * np.exp(-2j * np.pi * (kx * x / W + ky * y / H + kz * z / D))
    assert np.allclose(oracle.rfft3(f, 64), out, atol=1e-12)


def test_delta_and_constant(oracle):
    W, H, D = 8, 8, 8
    f = np.zeros((D, H, W)); f[0, 0, 0] = 1.0
    assert np.allclose(oracle.rfft3(f, 32), 1.0, atol=0)            # delta -> all ones, exactly
    F = oracle.rfft3(np.ones((D, H, W)), 32)
    assert F[0, 0, 0] == W * H * D and np.abs(F.reshape(-1)[1:]).max() < 1e-5


def test_parseval_and_linearity(oracle):
    W, H, D = 12, 12, 12
    rng = np.random.default_rng(3)
    a, b = rng.random((D, H, W)), rng.random((D, H, W))
    Fa, Fb = oracle.rfft3(a, 64), oracle.rfft3(b, 64)
    assert np.allclose(oracle.rfft3(2 * a - 3 * b, 64), 2 * Fa - 3 * Fb, atol=1e-10)
    full = np.fft.fftn(a)
    assert np.isclose((np.abs(full) ** 2).sum() / a.size, (a ** 2).sum())
    # half spectrum holds every independent bin: weights 1 for kx=0 and W/2, else 2
    w = np.full(W // 2 + 1, 2.0); w[0] = w[-1] = 1.0
    assert np.isclose(((np.abs(Fa) ** 2) * w).sum() / a.size, (a ** 2).sum())
