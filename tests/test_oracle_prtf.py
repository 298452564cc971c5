"""Pins the PRTF oracle (oracle/prtf_oracle.py).

  * cross-correlation and convolution against their brute-force sums
    c(t) = sum_y a(y+t) conj(b(y)) and sum_y a(y) b(t-y) on small arrays;
  * superimpose: a cyclically shifted copy comes back onto the original with
    the opposite shift; a copy reflected through the origin (x -> n-1-x) is
    recognised by the convolution peak and comes back exactly (even sizes);
  * the reflection keeps the middle plane of an odd number of planes, as the
    reference's in-place swap loop does;
  * phase match undoes a constant phase;
  * prtf: 1 wherever |F| > 0 for copies that alignment makes identical
    (shifted, reflected, phase-shifted), 0 where a transform vanishes, about
    1/sqrt(n) for independent random phases;
  * radial shell averages of constant and radially symmetric images.
"""
import numpy as np
import pytest

from oracle import prtf_oracle as P


def _brute_corr(a, b):
    out = np.zeros(a.shape, complex)
    for t in np.ndindex(*a.shape):
        out[t] = np.sum(np.roll(a, tuple(-s for s in t), axis=tuple(range(a.ndim))) * np.conj(b))
    return out


def _brute_conv(a, b):
    out = np.zeros(a.shape, complex)
    for t in np.ndindex(*a.shape):
        # b(t - y) over y: flip b through index 0, then shift by t
        bf = np.roll(np.flip(b, axis=tuple(range(b.ndim))), tuple(s + 1 for s in t), axis=tuple(range(b.ndim)))
        out[t] = np.sum(a * bf)
    return out


@pytest.mark.parametrize("shape", [(6, 8), (3, 4, 5)])
def test_correlation_and_convolution_sums(shape):
    rng = np.random.default_rng(1)
    a = rng.random(shape) + 1j * rng.random(shape)
    b = rng.random(shape) + 1j * rng.random(shape)
    assert np.allclose(P.cross_correlate(a, b), _brute_corr(a, b))
    assert np.allclose(P.convolute(a, b), _brute_conv(a, b))


def _object(shape, seed):
    rng = np.random.default_rng(seed)
    obj = np.zeros(shape)
    sl = tuple(slice(n // 4, n // 4 + max(2, n // 3)) for n in shape)
    obj[sl] = rng.random(obj[sl].shape) + 0.2
    return obj.astype(complex)


@pytest.mark.parametrize("shape,shift", [((32, 40), (5, -7)), ((8, 10, 12), (3, 2, -4))])
def test_superimpose_shift(shape, shift):
    a = _object(shape, 2)
    b = np.roll(a, shift, axis=tuple(range(a.ndim)))
    mb = np.roll((np.abs(a) > 0).astype(int), shift, axis=tuple(range(a.ndim)))
    out, mask, s, refl = P.superimpose(a, b, mb, enantio=True)
    assert not refl
    assert np.allclose(out, a)
    assert np.array_equal(mask, (np.abs(a) > 0).astype(int))
    want = tuple((-v) % n for v, n in zip(shift, shape))
    got = (s[1], s[0]) if len(shape) == 2 else (s[2], s[1], s[0])
    assert got == want


@pytest.mark.parametrize("shape", [(32, 40), (8, 10, 12)])
def test_superimpose_enantiomorph(shape):
    a = _object(shape, 3)
    b = np.flip(a, axis=tuple(range(a.ndim)))  # x -> n-1-x on every axis
    b = np.roll(b, (2,) * a.ndim, axis=tuple(range(a.ndim)))
    out, _, _, refl = P.superimpose(a, b, np.ones(shape, int), enantio=True)
    assert refl
    assert np.allclose(out, a)
    out2, _, _, refl2 = P.superimpose(a, b, np.ones(shape, int), enantio=False)
    assert not refl2 and not np.allclose(out2, a)


def test_reflect_keeps_middle_plane_of_odd_depth():
    img = np.arange(3 * 4 * 5).reshape(3, 4, 5).astype(complex)
    r = P.reflect_origo(img)
    assert np.array_equal(r[0], img[2, ::-1, ::-1])
    assert np.array_equal(r[2], img[0, ::-1, ::-1])
    assert np.array_equal(r[1], img[1])
    r2 = P.reflect_origo(np.arange(4 * 6).reshape(4, 6).astype(complex))
    assert np.array_equal(r2, np.arange(24).reshape(4, 6)[::-1, ::-1])


def test_phase_match():
    a = _object((24, 24), 4) * np.exp(1j * np.linspace(0, 1, 576)).reshape(24, 24)
    b, phi = P.phase_match(a, a * np.exp(0.7j))
    assert np.allclose(b, a)
    assert abs(phi + 0.7) < 1e-12


@pytest.mark.parametrize("shape", [(32, 32), (8, 10, 12)])
def test_prtf_of_aligned_copies_is_one(shape):
    a = _object(shape, 5)
    rng = np.random.default_rng(6)
    imgs = [a]
    for k in range(5):
        b = a * np.exp(1j * rng.uniform(-3, 3))
        if k % 2:
            b = np.flip(b, axis=tuple(range(a.ndim)))
        b = np.roll(b, tuple(int(v) for v in rng.integers(-5, 6, a.ndim)), axis=tuple(range(a.ndim)))
        imgs.append(b)
    imgs = np.array(imgs)
    sup = (np.abs(imgs) > 0).astype(int)
    out = P.prtf(imgs, sup, full_out=True)
    f = np.fft.fftshift(np.abs(np.fft.fftn(a)))
    assert np.allclose(out["prtf"][f > 1e-9], 1.0)
    assert np.allclose(out["super_image"], 6 * a)
    assert [d[1] for d in out["decisions"]] == [False, True, False, True, False]
    for i in range(6):
        assert np.allclose(out["images"][i], a)


def test_prtf_zero_where_a_transform_vanishes():
    a = np.zeros((16, 16), complex)
    a[0, 0] = 1
    a[0, 8] = 1  # F(k) = 1 + exp(-i pi kx): zero on odd kx
    out = P.prtf(np.array([a, a]), np.ones((2, 16, 16), int), translate_=False)
    p = np.fft.ifftshift(out["prtf"])
    assert np.all(p[:, 1::2] == 0) and np.allclose(p[:, ::2], 1.0)


def test_prtf_of_random_phases():
    rng = np.random.default_rng(7)
    n, shape = 64, (32, 32)
    imgs = np.fft.ifftn(np.exp(2j * np.pi * rng.random((n,) + shape)), axes=(1, 2))
    out = P.prtf(imgs, np.ones((n,) + shape, int), translate_=False)
    m = out["prtf"].mean()
    assert 0.5 / np.sqrt(n) < m < 2 / np.sqrt(n)


def test_radial():
    assert np.allclose(P.radial(np.full((16, 16), 3.0)), 3.0)
    y, x = np.mgrid[:16, :16]
    r = np.sqrt((x - 8) ** 2 + (y - 8) ** 2)
    prof = P.radial(np.floor(r))
    assert np.allclose(prof[:6], np.arange(6), atol=0.5)
    assert len(P.radial(np.ones((8, 8, 8)))) == 4
