"""Pin the oracle's FFT: brute-force DFT (tiny), numpy.fft, Parseval, inversion."""
import numpy as np
import pytest

import oracle


def dft(x, inverse=False):
    n = len(x)
    j = np.arange(n)
    s = 1.0 if inverse else -1.0
    W = np.exp(s * 2j * np.pi * np.outer(j, j) / n)
    return W @ x


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 16, 31, 36, 40, 64, 96, 67, 127, 181,
                               217, 362])
def test_brute_force(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    np.testing.assert_allclose(oracle.fft(x), dft(x), atol=1e-10 * max(n, 1))
    np.testing.assert_allclose(oracle.fft(x, inverse=True), dft(x, inverse=True), atol=1e-10 * max(n, 1))


@pytest.mark.parametrize("n", [128, 144, 160, 192, 256, 243, 1000])
def test_numpy_and_parseval(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    X = oracle.fft(x)
    np.testing.assert_allclose(X, np.fft.fft(x), rtol=0, atol=1e-9 * np.sqrt(n))
    assert abs(np.sum(np.abs(X) ** 2) / n - np.sum(np.abs(x) ** 2)) < 1e-9 * np.sum(np.abs(x) ** 2)
    np.testing.assert_allclose(oracle.fft(X, inverse=True) / n, x, atol=1e-12 * n)


def test_float_path_accuracy():
    rng = np.random.default_rng(1)
    x = rng.standard_normal(256) + 1j * rng.standard_normal(256)
    X = oracle.fft(x, dtype=np.float32)
    err = np.linalg.norm(X - np.fft.fft(x)) / np.linalg.norm(np.fft.fft(x))
    assert err < 1e-6
