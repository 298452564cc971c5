"""Pins the oracle's FFT (oracle/spimage_oracle.c: or_fft) without the reference.

The reference's transform is FFTW3 c2c over (z, y, x), unnormalised, forward
sign -1 (reference src/fft.c:79,95,163,185).  FFTW is absent here, so the
oracle's own transform is checked against
  * a brute-force O(N^2) DFT written from the definition,
  * numpy's pocketfft (an independent library),
  * Parseval, the impulse response and the shift theorem,
and a planted sign error must fail the brute-force comparison.
"""
import numpy as np
import pytest


def dft_nd(x, sign):
    """Brute-force DFT along every axis (matrix products in float64)."""
    out = x.astype(np.complex128)
    for ax in range(out.ndim):
        n = out.shape[ax]
        k = np.arange(n)
        W = np.exp(sign * 2j * np.pi * np.outer(k, k) / n)
        out = np.moveaxis(np.tensordot(W, np.moveaxis(out, ax, 0), axes=(1, 0)), 0, ax)
    return out


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


SMALL = [(8, 8), (6, 10), (9, 15), (4, 6, 5), (16, 12), (7, 7), (3, 5, 25)]


@pytest.mark.parametrize("shape", SMALL)
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft_matches_bruteforce_dft(o64, o32, shape, sign):
    rng = np.random.default_rng(hash(shape) & 0xFFFF)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    ref = dft_nd(x, sign)
    assert rel(o64.fft(x, sign), ref) < 1e-13
    assert rel(o32.fft(x.astype(np.complex64), sign).astype(np.complex128), dft_nd(x.astype(np.complex64), sign)) < 1e-6


@pytest.mark.parametrize("shape", [(64, 64), (96, 80), (16, 16, 16), (384, 48), (50, 30, 18)])
def test_fft_matches_numpy(o64, shape):
    rng = np.random.default_rng(7)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    assert rel(o64.fft(x, -1), np.fft.fftn(x)) < 1e-13
    assert rel(o64.fft(x, +1), np.fft.ifftn(x) * x.size) < 1e-13


def test_fft_parseval_impulse_shift(o64):
    rng = np.random.default_rng(3)
    x = rng.standard_normal((32, 24)) + 1j * rng.standard_normal((32, 24))
    X = o64.fft(x, -1)
    assert abs(np.sum(np.abs(X) ** 2) / x.size - np.sum(np.abs(x) ** 2)) < 1e-9 * np.sum(np.abs(x) ** 2)
    d = np.zeros((16, 16), complex)
    d[0, 0] = 1
    assert np.allclose(o64.fft(d, -1), 1.0)
    # shift by (a, b): X -> X * exp(-2 pi i (k a / ny + l b / nx))
    a, b = 3, 5
    xs = np.roll(np.roll(x, a, axis=0), b, axis=1)
    ky = np.arange(32)[:, None]
    kx = np.arange(24)[None, :]
    ph = np.exp(-2j * np.pi * (ky * a / 32 + kx * b / 24))
    assert rel(o64.fft(xs, -1), X * ph) < 1e-12


def test_planted_sign_error_is_caught(o64):
    """A transform with the wrong sign (the plausible mistake) fails the pin."""
    rng = np.random.default_rng(11)
    x = rng.standard_normal((8, 6)) + 1j * rng.standard_normal((8, 6))
    wrong = o64.fft(x, +1)  # planted: inverse sign where forward is meant
    assert rel(wrong, dft_nd(x, -1)) > 0.5
