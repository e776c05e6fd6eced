"""Pin the oracle's FFTs (FFTPACK conventions of fft.f) -- CPU only.

Checks: FFTPACK's own known answers (TFFTPK, fft.f:2897-3011), golden
vectors written by the reference's compiled fft.f (tests/golden/fftpack.npz),
numpy's FFT, and brute-force O(N^2) DFTs.  A wrong sign, a missing
normalisation or a mis-packed half-complex slot fails one of them.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "fftpack.npz")


def dft(x, sign):
    n = len(x)
    j = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(j, j) / n) @ x


@pytest.mark.parametrize("n", [8, 12, 16, 30, 64, 512])
def test_golden_fftpack(n):
    g = np.load(GOLD)
    tol = 1e-14 * n
    assert np.abs(O.cfft(g[f"c_in_{n}"], -1) - g[f"cfftf_{n}"]).max() < tol
    assert np.abs(O.cfft(g[f"c_in_{n}"], +1) - g[f"cfftb_{n}"]).max() < tol
    assert np.abs(O.rfft_pk(g[f"r_in_{n}"], -1) - g[f"rfftf_{n}"]).max() < tol
    assert np.abs(O.rfft_pk(g[f"r_in_{n}"], +1) - g[f"rfftb_{n}"]).max() < tol


@pytest.mark.parametrize("n", [8, 16, 64, 256])
def test_tfftpk_known_answers(n):
    # rfftf(ones) = N delta_0 ; rfftb(delta_0) = ones (fft.f:2897-2945)
    r = O.rfft_pk(np.ones(n), -1)
    assert abs(r[0] - n) < 1e-12 and np.abs(r[1:]).max() < 1e-12
    d = np.zeros(n)
    d[0] = 1.0
    assert np.abs(O.rfft_pk(d, +1) - 1.0).max() < 1e-14
    c = O.cfft(np.ones(n, complex), -1)
    assert abs(c[0] - n) < 1e-12 and np.abs(c[1:]).max() < 1e-12
    dc = np.zeros(n, complex)
    dc[0] = 1.0
    assert np.abs(O.cfft(dc, +1) - 1.0).max() < 1e-14


@pytest.mark.parametrize("n", [4, 6, 8, 10, 12, 16, 24, 32, 128])
def test_cfft_brute_force(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.allclose(O.cfft(x, -1), dft(x, -1), atol=1e-12 * n)
    assert np.allclose(O.cfft(x, +1), dft(x, +1), atol=1e-12 * n)
    assert np.allclose(O.cfft(x, -1), np.fft.fft(x), atol=1e-12 * n)


@pytest.mark.parametrize("n", [4, 8, 12, 16, 64])
def test_rfft_packing_vs_numpy(n):
    rng = np.random.default_rng(n + 1)
    x = rng.standard_normal(n)
    X = np.fft.rfft(x)               # X_k = sum x_j exp(-i..)
    r = O.rfft_pk(x, -1)
    assert abs(r[0] - X[0].real) < 1e-12
    for k in range(1, n // 2):
        assert abs(r[2 * k - 1] - X[k].real) < 1e-12
        assert abs(r[2 * k] - X[k].imag) < 1e-12
    assert abs(r[n - 1] - X[n // 2].real) < 1e-12
    # backward of forward = N x (fft.f:2259-2266: unnormalised)
    assert np.allclose(O.rfft_pk(r, +1), n * x, atol=1e-11 * n)
