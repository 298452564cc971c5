"""Seeded synthetic diffraction workloads for the phasing engine.

An object made of random Gaussian blobs inside a region, its far-field
intensity |FFT(object)|^2 scaled to a photon budget and Poisson-sampled, the
square root taken as measured amplitude, a central beamstop (pixels closer
than `beamstop` to an array corner in wrap-around order, the reference's
``sp_image_dist(..., SP_TO_CORNER)``, src/image_util.c:989-1009) masked out,
and a support: the region dilated by a few pixels.  Arrays follow the
reference's FFT layout: the object sits at the array origin
(``ifftshift``), zero frequency in the corners.

Everything here is numpy on the host and deterministic for a given seed.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Pattern:
    amplitudes: np.ndarray   # float32, measured amplitudes (0 under the beamstop)
    mask: np.ndarray         # int32, 1 = measured
    support: np.ndarray      # float32, 1 inside the initial support
    obj: np.ndarray          # complex64 ground truth (origin-centred)
    start: np.ndarray        # complex64 start model: random phases -> IFFT / N
    seed: int


def _fftn(x, inverse=False):
    """numpy's FFT; volumes above 4 Mi elements use scipy's multi-threaded one."""
    if x.size > (1 << 22):
        import os

        import scipy.fft
        f = scipy.fft.ifftn if inverse else scipy.fft.fftn
        return f(x, workers=min(16, os.cpu_count() or 1))
    return np.fft.ifftn(x) if inverse else np.fft.fftn(x)


def _grid(shape):
    axes = [np.arange(n) - n // 2 for n in shape]
    return np.meshgrid(*axes, indexing="ij")


def corner_distance(shape) -> np.ndarray:
    """Distance to the nearest corner, x fastest (src/image_util.c:989-1009)."""
    d2 = np.zeros(shape)
    for ax, n in enumerate(shape):
        c = np.arange(n, dtype=np.float64)
        c = np.minimum(c, n - 1 - c)
        sh = [1] * len(shape)
        sh[ax] = n
        d2 = d2 + (c.reshape(sh)) ** 2
    return np.sqrt(d2)


def make_pattern(shape, n_gauss=20, sigma=(2.0, 6.0), radius=40.0, photons=1e5, beamstop=4.0,
                 dilate=3.0, seed=0, complex_blobs=0, irregular=False, window=None) -> Pattern:
    """Build one seeded pattern.

    `radius` is a scalar (disc / sphere) or a tuple of semi-axes (ellipse /
    ellipsoid) in pixels.  `complex_blobs` adds that many uniform-random-phase
    blobs.  `irregular` replaces the disc by a low-pass random field
    thresholded around it.  `window` (in units of the blob's sigma) limits
    each Gaussian to a box around its centre, for large volumes."""
    shape = tuple(int(s) for s in shape)
    rng = np.random.default_rng(seed)
    g = _grid(shape)
    semi = np.broadcast_to(np.asarray(radius, dtype=np.float64), (len(shape),))
    rr = np.sqrt(sum((gi / s) ** 2 for gi, s in zip(g, semi)))
    region = rr <= 1.0
    if irregular:
        noise = rng.standard_normal(shape)
        freqs = np.meshgrid(*[np.fft.fftfreq(n) for n in shape], indexing="ij")
        kk = np.sqrt(sum(f * f for f in freqs))
        field = np.real(_fftn(_fftn(noise) * np.exp(-(kk / 0.01) ** 2), inverse=True))
        field /= np.abs(field).max() + 1e-30
        region = (rr + 0.35 * field) <= 1.0
    obj = np.zeros(shape, dtype=np.complex128)
    nblob = n_gauss + complex_blobs
    for b in range(nblob):
        # centre uniform inside the region's bounding ellipse
        while True:
            c = rng.uniform(-1, 1, len(shape)) * semi
            if np.sqrt(np.sum((c / semi) ** 2)) <= 1.0:
                break
        s = rng.uniform(*sigma)
        a = rng.uniform(0.5, 1.0)
        if window is None:
            box = tuple(slice(None) for _ in shape)
        else:
            box = tuple(slice(max(0, int(np.floor(ci - window * s)) + n // 2),
                              min(n, int(np.ceil(ci + window * s)) + n // 2 + 1)) for ci, n in zip(c, shape))
        r2 = sum((gi[box] - ci) ** 2 for gi, ci in zip(g, c))
        blob = a * np.exp(-r2 / (2 * s * s))
        if b >= n_gauss:
            blob = blob * np.exp(2j * np.pi * rng.uniform())
        obj[box] += blob
    obj[~region] = 0
    obj = np.fft.ifftshift(obj)
    region = np.fft.ifftshift(region)
    intens = np.abs(_fftn(obj)) ** 2
    intens *= photons / intens.sum()
    counts = rng.poisson(intens).astype(np.float64)
    amp = np.sqrt(counts).astype(np.float32)
    mask = (corner_distance(shape) >= beamstop).astype(np.int32)
    amp[mask == 0] = 0
    # support: region dilated by `dilate` pixels (distance transform by FFT
    # convolution with a ball)
    ball = np.fft.ifftshift(np.sqrt(sum(gi.astype(np.float64) ** 2 for gi in g)) <= dilate)
    conv = np.real(_fftn(_fftn(region.astype(np.float64)) * _fftn(ball), inverse=True))
    support = (conv > 0.5).astype(np.float32)
    # random-phase start: the reference's SpModelRandomPhases recipe
    # (src/phasing.c:626-646) with this generator
    phase = rng.uniform(0, 2 * np.pi, shape)
    start = (_fftn(amp * np.exp(1j * phase), inverse=True)).astype(np.complex64)
    return Pattern(amp, mask, support, obj.astype(np.complex64), start, seed)


def disc_256(seed=0) -> Pattern:
    """Plain 2-D HIO case, 256 x 256: 20 Gaussians (sigma 2-6 px) in a disc of
    radius 40 px, 1e5 photons, beamstop 4 px, support = disc + 3 px."""
    return make_pattern((256, 256), n_gauss=20, sigma=(2, 6), radius=40, photons=1e5, beamstop=4, dilate=3,
                        seed=seed)


def blob_1024(seed=1000) -> Pattern:
    """Single 2-D reconstruction, 1024 x 1024: 80 Gaussians plus complex
    random-phase blobs in an irregular blob of radius ~180 px, 1e7 photons,
    beamstop 8 px."""
    return make_pattern((1024, 1024), n_gauss=80, sigma=(2, 6), radius=180, photons=1e7, beamstop=8, dilate=3,
                        seed=seed, complex_blobs=20, irregular=True)


def ellipsoid_256(seed=3000) -> Pattern:
    """3-D reconstruction, 256^3: 200 Gaussians (sigma 2-8 voxels, evaluated
    within 5 sigma) in an ellipsoid with semi-axes 60/50/40, 1e8 photons,
    spherical beamstop 5."""
    return make_pattern((256, 256, 256), n_gauss=200, sigma=(2, 8), radius=(40, 50, 60), photons=1e8,
                        beamstop=5, dilate=3, seed=seed, window=5.0)


def disc_scaled(n: int, seed: int) -> Pattern:
    """disc_256 scaled to an n x n array (disc radius and blob count by n/256)."""
    f = n / 256.0
    return make_pattern((n, n), n_gauss=max(4, int(round(20 * f))), sigma=(2, 6), radius=40 * f,
                        photons=1e5 * f * f, beamstop=4, dilate=3, seed=seed)


def smooth_sizes(lo: int = 128, hi: int = 1024) -> list[int]:
    """FFT-smooth lengths 2^a 3^b 5^c in [lo, hi]."""
    out = set()
    p2 = 1
    while p2 <= hi:
        p3 = p2
        while p3 <= hi:
            p5 = p3
            while p5 <= hi:
                if p5 >= lo:
                    out.add(p5)
                p5 *= 5
            p3 *= 3
        p2 *= 2
    return sorted(out)


def serving_pattern(shape, seed: int) -> Pattern:
    """Light pattern for a serving request of any shape: 10 Gaussians in a
    disc / ball of radius n/6.4 (40 px at 256), 1e5 photons per 256^2 of
    area, beamstop 4, support = region + 3 px."""
    n = min(shape)
    f = n / 256.0
    return make_pattern(shape, n_gauss=10, sigma=(2, 6), radius=40 * f, photons=1e5 * f * f * (n if len(shape) == 3 else 1),
                        beamstop=4, dilate=3, seed=seed)


def serving_trace(rate: float = 200.0, seconds: float = 3.0, seed: int = 4000, sizes_2d=None, sizes_3d=(64, 80, 96, 128),
                  frac_3d: float = 0.15, its=(100, 1000), quantum: int = 50):
    """Poisson request trace: (arrival, shape, algorithm, iterations, seed)
    tuples.  85 % 2-D shapes n x n with n drawn from `sizes_2d` (default the
    FFT-smooth lengths in [128, 1024]), 15 % 3-D n^3 from `sizes_3d`;
    algorithm uniform over ER / HIO / RAAR; iterations uniform over the
    multiples of `quantum` in `its`."""
    rng = np.random.default_rng(seed)
    sizes_2d = list(sizes_2d) if sizes_2d is not None else smooth_sizes(128, 1024)
    t, out, k = 0.0, [], 0
    iters = np.arange(its[0], its[1] + 1, quantum)
    while True:
        t += rng.exponential(1.0 / rate)
        if t > seconds:
            break
        if rng.uniform() < frac_3d:
            n = int(rng.choice(sizes_3d))
            shape = (n, n, n)
        else:
            n = int(rng.choice(sizes_2d))
            shape = (n, n)
        alg = ("er", "hio", "raar")[int(rng.integers(3))]
        out.append((t, shape, alg, int(rng.choice(iters)), seed * 1000 + k))
        k += 1
    return out
