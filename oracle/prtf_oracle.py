"""ORACLE -- test infrastructure, not part of the product.

A numpy (float64) restatement of the reference's phase-retrieval transfer
function over many reconstructions, ``prtf()`` of src/_spimage_prtf.py:17-101,
with the libspimage routines it calls:

  superimpose      sp_image_superimpose (src/image_util.c:2973-3005): the
                   integer shift maximising |cross-correlation| of the
                   magnitudes (sp_image_cross_correlate, :1313-1370), or,
                   with SpEnantiomorph, of the convolution
                   (sp_image_convolute, :1474-1553) when that is larger, in
                   which case b is reflected through the origin first and
                   the shift is the convolution's peak + 1;
  argmax           sp_c3matrix_max (include/spimage/linear_alg.h:1974-1988):
                   the first index of the largest |value|;
  reflect          sp_image_reflect(.., SP_ORIGO) (src/image_util.c:161-175,
                   240-306): the image only, not the mask; in 3-D with an odd
                   number of planes the middle plane stays as it is (the
                   reference swaps each of its pairs twice);
  translate        sp_image_translate(.., SP_TRANSLATE_WRAP_AROUND)
                   (src/image_util.c:3188-3215): image and mask, cyclic;
  phase_match      sp_image_phase_match(a, b, 2) (src/image_util.c:3160-3186):
                   b times exp(i arg sum |a|^2 exp(i(arg a - arg b)));
  radial           _radial (src/_spimage_utils.py:42-68) for ``prtf_r``.

Only tests/ and bench.py's CPU-baseline leg import this module, and only as
the checker or as the timed CPU baseline.  Its pins are in
tests/test_oracle_prtf.py.
"""
from __future__ import annotations

import numpy as np


def argmax_abs(c: np.ndarray) -> int:
    """First raster index of the largest magnitude."""
    return int(np.argmax(np.abs(c).ravel()))


def index_to_xyz(i: int, shape) -> tuple[int, int, int]:
    """Raster index -> (x, y, z) for an array of shape (y, x) or (z, y, x)."""
    if len(shape) == 2:
        y, x = np.unravel_index(i, shape)
        return int(x), int(y), 0
    z, y, x = np.unravel_index(i, shape)
    return int(x), int(y), int(z)


def cross_correlate(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """IFFT(FFT(a) conj(FFT(b))) / N (same size, no padding)."""
    return np.fft.ifftn(np.fft.fftn(a) * np.conj(np.fft.fftn(b)))


def convolute(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """IFFT(FFT(a) FFT(b)) / N."""
    return np.fft.ifftn(np.fft.fftn(a) * np.fft.fftn(b))


def reflect_origo(img: np.ndarray) -> np.ndarray:
    if img.ndim == 2:
        return img[::-1, ::-1].copy()
    out = img[::-1, ::-1, ::-1].copy()
    nz = img.shape[0]
    if nz % 2 == 1:
        out[nz // 2] = img[nz // 2]
    return out


def translate(img: np.ndarray, x: int, y: int, z: int) -> np.ndarray:
    """Pixel p moves to p + (x, y, z), wrapping around."""
    if img.ndim == 2:
        return np.roll(img, (y, x), axis=(0, 1))
    return np.roll(img, (z, y, x), axis=(0, 1, 2))


def superimpose(a: np.ndarray, b: np.ndarray, mask_b: np.ndarray, enantio: bool):
    """b (and its mask) moved onto a; returns (b, mask, (x, y, z), reflected)."""
    aa = np.abs(a).astype(np.complex128)
    bb = np.abs(b).astype(np.complex128)
    direct = cross_correlate(aa, bb)
    i = argmax_abs(direct)
    best = np.abs(direct.ravel()[i])
    x, y, z = index_to_xyz(i, a.shape)
    reflected = False
    if enantio:
        conv = convolute(aa, bb)
        i2 = argmax_abs(conv)
        if np.abs(conv.ravel()[i2]) > best:
            x2, y2, z2 = index_to_xyz(i2, a.shape)
            x, y, z = x2 + 1, y2 + 1, z2 + 1
            b = reflect_origo(b)
            reflected = True
    return translate(b, x, y, z), translate(mask_b, x, y, z), (x, y, z), reflected


def phase_match(a: np.ndarray, b: np.ndarray) -> tuple[np.ndarray, float]:
    """Weighted (|a|^2) constant-phase match of b to a."""
    ang = np.angle(a) - np.angle(b)
    avg = np.sum(np.abs(a) ** 2 * np.exp(1j * ang))
    phi = float(np.angle(avg))
    return np.abs(b) * np.exp(1j * (np.angle(b) + phi)), phi


def prtf(images_rs, supports, translate_=True, enantio=True, full_out=False):
    """The reference's prtf() with the same output keys; `decisions` adds the
    (shift, reflected, phase) chosen for each image after the first."""
    images = np.asarray(images_rs).astype(np.complex128)
    masks = np.asarray(supports).astype(np.int64)
    n = images.shape[0]
    f0 = np.fft.fftn(images[0])
    sum_fs = f0.copy()
    pos = np.abs(sum_fs) > 0
    sum_fs[pos] /= np.abs(sum_fs[pos])
    zeros = (~pos).astype(np.int64)
    avg = images[0].copy()
    avg_msk = np.zeros(images.shape[1:])
    out_images = np.zeros_like(images)
    out_masks = np.zeros(images.shape, bool)
    out_images[0] = images[0]
    out_masks[0] = masks[0] != 0
    decisions = []
    for i in range(1, n):
        img, msk = images[i], masks[i]
        shift, refl, phi = (0, 0, 0), False, 0.0
        if translate_:
            img, msk, shift, refl = superimpose(avg, img, msk, enantio)
            img, phi = phase_match(avg, img)
        decisions.append((shift, refl, phi))
        avg = avg + img
        if msk.sum() > 0:
            avg_msk[msk == 0] += 1
        out_images[i] = img
        out_masks[i] = msk != 0
        f = np.fft.fftn(img)
        pos = np.abs(f) > 0
        f[pos] /= np.abs(f[pos])
        zeros += ~pos
        sum_fs += f
    p = np.abs(sum_fs / n)
    p[zeros > 0] = 0
    out = {"prtf": np.fft.fftshift(p), "super_image": avg, "decisions": decisions}
    if full_out:
        out["prtf_r"] = radial(out["prtf"])
        out["super_mask"] = avg_msk > 0
        out["images"] = out_images
        out["masks"] = out_masks
    return out


def radial(image, f=np.mean, shell_thickness=1.0):
    """Shell averages about shape[0] // 2 (assumes equal sizes along all axes)."""
    n_dim = image.ndim
    im_dim = image.shape[0]
    c0 = image.shape[0] // 2
    num_shells = int((im_dim - c0) / shell_thickness)
    r_out = np.arange(shell_thickness, (num_shells + 1) * shell_thickness, shell_thickness)
    r_in = r_out - 1
    c = np.array([c0, c0, c0])
    if n_dim == 2:
        image = image[:, :, None]
        c = np.array([c0, c0, 0])
    elif n_dim == 1:
        image = image[:, None, None]
        c = np.array([c0, 0, 0])
    x, y, z = np.meshgrid(np.arange(image.shape[0]), np.arange(image.shape[1]), np.arange(image.shape[2]),
                          indexing="ij")
    r2 = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2
    vals = []
    for ri, ro in zip(r_in, r_out):
        shell = (r2 >= ri ** 2) & (r2 < ro ** 2)
        vals.append(f(np.squeeze(image[shell])))
    vals = np.array(vals)
    return vals[np.isfinite(vals)]
