"""GPU PRTF (libspimage_amd.prtf, spx_prtf) against the PRTF oracle, which
restates the reference's prtf() (src/_spimage_prtf.py:17-101) in float64.

Inputs are noisy copies of one object with random cyclic shifts, random
reflections through the origin and random constant phases, so that every
alignment decision (shift, reflection) is clear-cut; the decisions must be
the same, which the aligned reconstructions and their masks show (masks
exactly, images to float rounding).  Tolerances: aligned images and the
running sum within 1e-5 relative L2; the PRTF within 1e-4 absolute (it is
an average of unit phasors, whose direction is ill-conditioned only where
|F| is tiny).
"""
import numpy as np
import pytest

from _helpers import rel

pytestmark = pytest.mark.gpu


def _copies(shape, n, seed, reflect=True, noise=0.01):
    rng = np.random.default_rng(seed)
    obj = np.zeros(shape, complex)
    sl = tuple(slice(n_ // 4, n_ // 4 + n_ // 3) for n_ in shape)
    blk = obj[sl].shape
    obj[sl] = (rng.random(blk) + 0.3) * np.exp(0.5j * rng.random(blk))
    sup = (np.abs(obj) > 0).astype(np.int32)
    imgs, masks = [], []
    for k in range(n):
        b, m = obj.copy(), sup.copy()
        if reflect and k % 3 == 2:
            b = np.flip(b)
        b = b * np.exp(1j * rng.uniform(-3, 3)) + noise * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
        s = tuple(int(v) for v in rng.integers(-4, 5, len(shape)))
        imgs.append(np.roll(b, s, axis=tuple(range(len(shape)))))
        masks.append(np.roll(m, s, axis=tuple(range(len(shape)))))
    return np.array(imgs).astype(np.complex64), np.array(masks)


@pytest.mark.parametrize("shape", [(64, 64), (48, 80), (16, 16, 16), (15, 16, 12)])
@pytest.mark.parametrize("enantio", [True, False])
def test_prtf_matches_oracle(shape, enantio):
    from libspimage_amd import prtf as gp
    from oracle import prtf_oracle as P
    imgs, masks = _copies(shape, 7, seed=len(shape) + 3 * enantio)
    g = gp.prtf(imgs, masks, enantio=enantio, full_out=True)
    o = P.prtf(imgs, masks, enantio=enantio, full_out=True)
    # the reflected copies take the convolution branch exactly when allowed
    assert any(d[1] for d in o["decisions"]) == enantio
    assert np.array_equal(g["masks"], o["masks"])
    for i in range(len(imgs)):
        assert rel(g["images"][i], o["images"][i]) < 1e-5, i
    assert rel(g["super_image"], o["super_image"]) < 1e-5
    assert np.abs(g["prtf"] - o["prtf"]).max() < 1e-4
    assert np.array_equal(g["super_mask"], o["super_mask"])
    assert np.allclose(g["prtf_r"], o["prtf_r"], atol=1e-4)


@pytest.mark.parametrize("n", [8, 9, 19])
def test_prtf_over_chunk_boundaries(n):
    """Reconstructions travel to and from the device in chunks of 8 beside
    the alignment loop: counts on, just past and well past a chunk boundary
    give the oracle's result."""
    from libspimage_amd import prtf as gp
    from oracle import prtf_oracle as P
    imgs, masks = _copies((48, 80), n, seed=40 + n)
    masks[n - 1] = 0  # a reconstruction without a mask in the last chunk
    g = gp.prtf(imgs, masks, full_out=True)
    o = P.prtf(imgs, masks, full_out=True)
    assert np.array_equal(g["masks"], o["masks"])
    for i in range(n):
        assert rel(g["images"][i], o["images"][i]) < 1e-5, i
    assert rel(g["super_image"], o["super_image"]) < 1e-5
    assert np.abs(g["prtf"] - o["prtf"]).max() < 1e-4
    assert np.array_equal(g["super_mask"], o["super_mask"])


def test_prtf_without_translation():
    from libspimage_amd import prtf as gp
    from oracle import prtf_oracle as P
    imgs, masks = _copies((64, 64), 5, seed=9, reflect=False)
    g = gp.prtf(imgs, masks, translate=False, full_out=True)
    o = P.prtf(imgs, masks, translate_=False, full_out=True)
    assert np.array_equal(g["images"], imgs)
    assert np.abs(g["prtf"] - o["prtf"]).max() < 1e-4
    assert rel(g["super_image"], o["super_image"]) < 1e-5


def test_prtf_of_aligned_copies_is_one():
    """Noise-free shifted, reflected, phase-shifted copies of a real object
    align exactly: the PRTF is 1 wherever the object's transform is not 0."""
    from libspimage_amd import prtf as gp
    rng = np.random.default_rng(4)
    shape = (64, 64)
    obj = np.zeros(shape)
    obj[10:30, 12:28] = rng.random((20, 16)) + 0.2
    imgs = []
    for k in range(6):
        b = obj * np.exp(1j * rng.uniform(-3, 3))
        if k % 2:
            b = np.flip(b)
        imgs.append(np.roll(b, tuple(int(v) for v in rng.integers(-6, 7, 2)), axis=(0, 1)))
    g = gp.prtf(np.array(imgs), None)
    f = np.fft.fftshift(np.abs(np.fft.fftn(obj)))
    assert np.abs(g["prtf"][f > 1e-3 * f.max()] - 1).max() < 1e-4


def test_prtf_rejects_bad_input():
    from libspimage_amd import prtf as gp
    from libspimage_amd import spimage as sp
    with pytest.raises(ValueError):
        gp.prtf(np.zeros((4, 8), np.complex64), None)
    assert sp.spx_prtf(None, None, 8, 8, 1, 2, 1, 1, None, None, None, None, None) == -1
    imgs = np.zeros((2, 7, 11), np.complex64)  # 11 has no Stockham plan
    with pytest.raises(RuntimeError):
        gp.prtf(imgs, None)
