"""Phase-retrieval transfer function over many reconstructions, on the GPU.

``prtf(images_rs, supports, translate=True, enantio=True, full_out=False)``
has the signature and output dictionary of the reference's ``spimage.prtf``
(src/_spimage_prtf.py:17-101); the alignment, the running sums and the
transforms run on the device through ``spx_prtf`` (include/sp_prtf.h).
The shell average ``prtf_r`` of ``full_out`` is the reference's ``radial``
(src/_spimage_utils.py:42-68), computed here with numpy on the result.
"""
from __future__ import annotations

import numpy as np

from . import spimage as sp


def prtf(images_rs, supports, translate=True, enantio=True, full_out=False):
    """`images_rs`: (N, y, x) or (N, z, y, x) complex reconstructions;
    `supports`: their masks, same shape.  Returns {"prtf", "super_image"} and,
    with full_out, "prtf_r", "super_mask", "images", "masks"."""
    imgs = np.ascontiguousarray(images_rs, dtype=np.complex64)
    if imgs.ndim not in (3, 4):
        raise ValueError("images_rs must be (N, y, x) or (N, z, y, x)")
    n = imgs.shape[0]
    shape = imgs.shape[1:]
    if imgs.ndim == 3:
        nz, ny, nx = 1, shape[0], shape[1]
    else:
        nz, ny, nx = shape
    masks = np.ascontiguousarray(supports, dtype=np.int32) if supports is not None else None
    if masks is not None and masks.shape != imgs.shape:
        raise ValueError("supports must have the shape of images_rs")
    N = nx * ny * nz
    p = np.empty(N, np.float32)
    sup = np.empty(shape, np.complex64)
    aligned = np.empty(imgs.shape, np.complex64) if full_out else None
    amasks = np.empty(imgs.shape, np.int32) if full_out else None
    count = np.empty(shape, np.float32) if full_out else None

    def fp(a):
        return None if a is None else a.ctypes.data_as(sp.PF)

    def ip(a):
        return None if a is None else a.ctypes.data_as(sp.PI)

    rc = sp.spx_prtf(fp(imgs), ip(masks), nx, ny, nz, n, 1 if translate else 0, 1 if enantio else 0, fp(p), fp(sup),
                     fp(aligned), ip(amasks), fp(count))
    if rc != 0:
        raise RuntimeError(f"spx_prtf failed ({rc}): {sp.last_error()}")
    out = {"prtf": np.fft.fftshift(p.reshape(shape)), "super_image": sup}
    if full_out:
        out["prtf_r"] = radial(out["prtf"])
        out["super_mask"] = count > 0
        out["images"] = aligned
        out["masks"] = amasks != 0
    return out


def radial(image, f=np.mean, shell_thickness=1.0):
    """Shell averages about index shape[0] // 2 on every axis, shells
    [r - 1, r) for r = 1, 2, ...; non-finite averages are dropped
    (src/_spimage_utils.py:42-68, which assumes equal sizes along all axes)."""
    n_dim = image.ndim
    c0 = image.shape[0] // 2
    num_shells = int((image.shape[0] - c0) / shell_thickness)
    r_out = np.arange(shell_thickness, (num_shells + 1) * shell_thickness, shell_thickness)
    img = image
    c = [c0, c0, c0]
    if n_dim == 2:
        img = image[:, :, None]
        c = [c0, c0, 0]
    elif n_dim == 1:
        img = image[:, None, None]
        c = [c0, 0, 0]
    axes = np.ogrid[tuple(slice(0, n) for n in img.shape)]
    r2 = sum((g - ci) ** 2 for g, ci in zip(axes, c))
    if f is np.mean and shell_thickness == 1.0:
        # shell r - 1 <= sqrt(r2) < r is floor(sqrt(r2)) = r - 1 (exact for
        # these integer r2): one pass with bincount instead of one mask per shell
        k = np.floor(np.sqrt(r2.astype(np.float64))).astype(np.int64).ravel()
        keep = k < len(r_out)
        cnt = np.bincount(k[keep], minlength=len(r_out))
        tot = np.bincount(k[keep], weights=img.ravel()[keep].astype(np.float64), minlength=len(r_out))
        with np.errstate(invalid="ignore", divide="ignore"):
            vals = tot / cnt
    else:
        vals = np.array([f(np.squeeze(img[(r2 >= (ro - 1) ** 2) & (r2 < ro ** 2)])) for ro in r_out])
    return vals[np.isfinite(vals)]
