"""Pins the oracle's blur, shrinkwrap updates and parameter schedules.

  * blur (reference src/image_filter.c:80-91): equals the spectrum times the
    closed-form transfer function exp(-2 pi^2 (r/nx)^2 sigma^2) with r the
    wrap-around frequency distance (src/image_util.c:1012-1032), keeps the
    integral and the position of the maximum (the reference's
    tests/ImageTests.c:194-243 assert exactly these two);
  * area update (src/support_update.c:136-151): threshold = k-th largest
    |blur| with k = (int)(N * area), checked by a brute-force sort, and the
    mask is |blur| > threshold;
  * threshold update (src/support_update.c:167-179): max |blur| * rel;
  * Bezier schedule (include/spimage/map.h:103-141): end points, midpoint,
    monotonicity, constant outside the keys, and agreement with the closed
    (Cardano) solution of the same curve.
"""
import numpy as np
import pytest

from oracle.oracle import ER, SUP_AREA, SUP_THRESHOLD


def transfer(shape, sigma):
    axes = []
    for n in shape:
        i = np.arange(n, dtype=np.float64)
        axes.append(np.minimum(i, n - i))
    grids = np.meshgrid(*axes, indexing="ij")
    r = np.sqrt(sum(g * g for g in grids)) / shape[-1]
    return np.exp(-2 * np.pi ** 2 * r ** 2 * sigma ** 2)


@pytest.mark.parametrize("shape", [(32, 32), (24, 40), (8, 10, 12)])
@pytest.mark.parametrize("sigma", [0.7, 3.0])
def test_blur_closed_form(o64, shape, sigma):
    rng = np.random.default_rng(1)
    x = rng.random(shape) + 0j
    out = o64.blur(x, sigma)
    ref = np.fft.ifftn(np.fft.fftn(x) * transfer(shape, sigma))
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-12


@pytest.mark.parametrize("shape", [(64, 64), (16, 16, 16)])
def test_blur_keeps_integral_and_maximum(o32, shape):
    x = np.zeros(shape, np.complex64)
    peak = tuple(s // 3 for s in shape)
    x[peak] = 10.0
    x[tuple(s // 3 + 1 for s in shape)] = 3.0
    out = o32.blur(x, 2.0)
    assert abs(out.real.sum() - x.real.sum()) < 1e-4 * abs(x.real.sum())
    assert np.unravel_index(np.argmax(np.abs(out)), shape) == peak


def test_area_update_matches_sort(o32):
    rng = np.random.default_rng(5)
    shape = (64, 64)
    ph = o32.phaser(shape)
    g = (rng.random(shape) + 1j * rng.random(shape)).astype(np.complex64)
    ph.set_inputs(np.ones(shape, np.float32), np.ones(shape, bool), np.zeros(shape, bool), g)
    t = ph.support_area_update(3.0, 0.08)
    b = np.abs(o32.blur(np.abs(g).astype(np.complex64), 3.0))
    k = int(np.float32(b.size) * np.float32(0.08))
    assert t == np.sort(b.ravel())[::-1][k]
    assert np.array_equal((ph.flags & 1).astype(bool), b > t)
    assert (ph.flags & 2).all()  # measured bits untouched


def test_threshold_update(o32):
    rng = np.random.default_rng(6)
    shape = (32, 48)
    ph = o32.phaser(shape)
    g = (rng.random(shape) + 1j * rng.random(shape)).astype(np.complex64)
    ph.set_inputs(np.ones(shape, np.float32), np.ones(shape, bool), np.zeros(shape, bool), g)
    t = ph.support_threshold_update(1.5, 0.2)
    b = np.abs(o32.blur(np.abs(g).astype(np.complex64), 1.5))
    assert t == np.float32(b.max()) * np.float32(0.2)
    assert np.array_equal((ph.flags & 1).astype(bool), b > t)


def cardano(p0, y0, p3, y3, x):
    u = (x - p0) / (p3 - p0)
    q = 0.5 - u
    d = np.sqrt(q * q / 4 + 1 / 64)
    t = np.cbrt(-q / 2 + d) + np.cbrt(-q / 2 - d) + 0.5
    w = 1 - t
    return y0 * w ** 3 + 3 * y0 * t * w * w + 3 * y3 * t * t * w + y3 * t ** 3


def test_bezier_schedule(o64):
    pairs = [(0, 3.0), (300, 0.7)]
    assert o64.bezier(pairs, -5) == 3.0
    assert o64.bezier(pairs, 0) == 3.0
    assert o64.bezier(pairs, 300) == 0.7
    assert o64.bezier(pairs, 1000) == 0.7
    assert abs(o64.bezier(pairs, 150) - 1.85) < 1e-12
    xs = np.arange(0, 301)
    ys = np.array([o64.bezier(pairs, x) for x in xs])
    assert np.all(np.diff(ys) <= 0)
    for x in (1, 17, 99, 250, 299):
        assert abs(o64.bezier(pairs, x) - cardano(0, 3.0, 300, 0.7, x)) < 1e-12
    three = [(0, 0.5), (100, 0.9), (400, 0.2)]
    assert abs(o64.bezier(three, 100) - 0.9) < 1e-15
    assert abs(o64.bezier(three, 250) - 0.55) < 1e-12
    assert o64.bezier([(0, 0.9)], 123) == 0.9


def test_scheduled_area_updates_consume_iterations(o32):
    """The area support runs every `period` iterations and consumes one."""
    rng = np.random.default_rng(8)
    shape = (32, 32)
    ph = o32.phaser(shape)
    amp = rng.random(shape).astype(np.float32)
    g = (rng.random(shape) + 1j * rng.random(shape)).astype(np.complex64)
    ph.set_inputs(amp, np.ones(shape, bool), np.ones(shape, bool), g)
    s = o32.sup(SUP_AREA, period=10, blur=[(0, 3.0), (40, 0.7)], values=[(0, 0.1)])
    ph.iterate(o32.alg(ER), 40, s)
    assert ph.support_log() == [9, 19, 29, 39]  # the 40th iteration is the 4th update
    assert abs(ph.support_fraction() - 0.1) < 0.02
    s2 = o32.sup(SUP_THRESHOLD, period=10, blur=[(0, 1.0)], values=[(0, 0.3)])
    ph2 = o32.phaser(shape)
    ph2.set_inputs(amp, np.ones(shape, bool), np.ones(shape, bool), g)
    ph2.iterate(o32.alg(ER), 21, s2)
    assert ph2.support_log() == [9, 19]
