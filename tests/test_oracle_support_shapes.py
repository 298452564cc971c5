"""Pins the oracle's closing, centre-image and template support updates.

  * closing (reference src/support_update.c:213-280): equals scipy.ndimage's
    binary dilation with the 6-neighbour cross repeated `size` times (faces
    add nothing: border value 0) followed by as many erosions (faces remove
    nothing: border value 1); the other flag bits stay;
  * centre image (src/support_update.c:282-389): the support comes back as a
    cyclic roll of itself whose float64 centre of mass, in the reference's
    shifted frame, lies within one pixel of the origin; a support straddling
    the middle row (2-D) or a middle plane (3-D) first moves by half the
    image, which the hand-built cases below need to end up centred; the
    other flag bits stay;
  * template (src/support_update.c:47-85,181-207): |blur| of the template,
    its descending sort, the template's area fraction and the threshold at
    index (int)(N * area * fraction) against numpy and a brute-force sort;
  * the scheduler runs a support array's algorithms in order at each update.
"""
import numpy as np
import pytest
from scipy import ndimage

from oracle.oracle import ER, SUP_CENTRE, SUP_CLOSE, SUP_STATIC, SUP_TEMPLATE

CROSS = ndimage.generate_binary_structure(3, 1)


def _phaser(o, support, measured=None, seed=0):
    shape = support.shape
    rng = np.random.default_rng(seed)
    ph = o.phaser(shape)
    meas = rng.random(shape) > 0.3 if measured is None else measured
    g = (rng.random(shape) + 1j * rng.random(shape)).astype(np.complex64)
    ph.set_inputs(np.ones(shape, np.float32), meas, support, g)
    return ph, meas


def _as3(a):
    return a[None] if a.ndim == 2 else a


@pytest.mark.parametrize("shape", [(40, 56), (12, 10, 14)])
@pytest.mark.parametrize("size", [1, 2, 3])
def test_close_matches_scipy(o32, shape, size):
    sup = _holey(shape, seed=size + len(shape))
    ph, meas = _phaser(o32, sup)
    ph.support_close_update(size)
    s3 = _as3(sup)
    d = ndimage.binary_dilation(s3, CROSS, iterations=size, border_value=0)
    want = ndimage.binary_erosion(d, CROSS, iterations=size, border_value=1).reshape(shape)
    got = (ph.flags & 1).astype(bool)
    assert 0 < want.sum() < want.size and not np.array_equal(want, sup)
    assert np.array_equal(got, want)
    assert np.array_equal((ph.flags & 2) != 0, meas)
    # a plausible slip, the 26-neighbour box, gives a different closing here
    box = np.ones((3, 3, 3), bool)
    other = ndimage.binary_erosion(ndimage.binary_dilation(s3, box, iterations=size, border_value=0), box,
                                   iterations=size, border_value=1).reshape(shape)
    assert not np.array_equal(other, want)


def _holey(shape, seed):
    """A few blobs with random holes and a sprinkle of isolated pixels."""
    rng = np.random.default_rng(seed)
    sup = np.zeros(shape, bool)
    for _ in range(3):
        c = tuple(rng.integers(0, n) for n in shape)
        sup |= _blob(shape, c, min(shape) / 6)
    sup &= rng.random(shape) > 0.3
    sup |= rng.random(shape) < 0.02
    return sup


def _shifted_com(sup):
    """Centre of mass in the frame shifted by (n-1)//2, then shifted back (float64)."""
    idx = np.nonzero(sup)
    com = []
    for ax, n in enumerate(sup.shape):
        h = (n - 1) // 2
        com.append(((idx[ax] + h) % n).mean() - h)
    return np.array(com)


def _roll_of(a, b):
    """The shift s with np.roll(a, s) == b, or None."""
    for s in np.ndindex(*a.shape):
        if np.array_equal(np.roll(a, s, axis=tuple(range(a.ndim))), b):
            return s
    return None


def _blob(shape, centre, radius):
    grids = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    r2 = sum((g - c) ** 2 for g, c in zip(grids, centre))
    return r2 <= radius ** 2


@pytest.mark.parametrize("shape,centre,radius", [((32, 40), (9, 12), 2.5), ((33, 27), (20, 5), 2.5),
                                                 ((32, 40), (15, 12), 2.5), ((32, 40), (8, 19), 2.5),
                                                 ((12, 14, 16), (3, 4, 5), 1.5), ((12, 14, 16), (3, 6, 5), 1.5),
                                                 ((12, 14, 16), (5, 4, 7), 1.5)])
def test_centre_image(o32, shape, centre, radius):
    # the blobs either keep clear of the middle rows / planes ((n-1)//2) or
    # straddle them symmetrically; the latter cross the border of the shifted
    # frame and need the half-image move first
    sup = _blob(shape, centre, radius)
    ph, meas = _phaser(o32, sup)
    ph.support_centre_image()
    got = (ph.flags & 1).astype(bool)
    assert np.array_equal((ph.flags & 2) != 0, meas)
    assert got.sum() == sup.sum()
    assert _roll_of(sup, got) is not None
    com = _shifted_com(got)
    assert np.all(np.abs(com) < 1.0), com


def test_centre_image_edge_touch(o32):
    """A blob that only touches a middle plane with its edge still moves by
    half the image first (the reference tests for any support pixel on the
    plane), so it need not end up centred; it is still rolled intact."""
    shape = (12, 14, 16)
    sup = _blob(shape, (3, 4, 5), 2.5)  # z 1..5 touches z = 5
    ph, meas = _phaser(o32, sup)
    ph.support_centre_image()
    got = (ph.flags & 1).astype(bool)
    assert _roll_of(sup, got) is not None
    assert np.array_equal((ph.flags & 2) != 0, meas)


def test_centre_image_half_move_matters(o32):
    """Without the half-image move, a support straddling the middle row would
    be centred from its wrapped mean and end up far from the origin."""
    shape = (32, 40)
    sup = _blob(shape, (15, 12), 2.5)  # rows 13..17 straddle row 15
    naive = _shifted_com(sup)
    assert np.abs(naive).max() > 5  # the wrapped mean is off
    ph, _ = _phaser(o32, sup)
    ph.support_centre_image()
    assert np.all(np.abs(_shifted_com((ph.flags & 1).astype(bool))) < 1.0)


def _raster_float32_centre_shift(sup):
    """The translation the reference's 2-D centre-image update applies
    (src/support_update.c:304-389), restated in numpy: the middle-row and
    middle-column moves, then the centre of mass of the moved support with
    each coordinate summed sequentially in float32 in raster order
    (src/linear_alg.c:444-475; np.add.accumulate is a sequential float32
    loop), divided by the count in float32, truncated toward zero."""
    ny, nx = sup.shape
    hy, hx = (ny - 1) // 2, (nx - 1) // 2
    oy = hy if sup[hy, :].any() else 0
    ox = hx if sup[:, hx].any() else 0
    moved = np.roll(sup, (oy, ox), axis=(0, 1))
    ys, xs = np.nonzero(moved)  # raster order: y, then x
    sx = np.add.accumulate(((xs + hx) % nx).astype(np.float32))[-1]
    sy = np.add.accumulate(((ys + hy) % ny).astype(np.float32))[-1]
    cnt = np.float32(len(xs))
    cx = np.float32(sx / cnt) - np.float32(hx)
    cy = np.float32(sy / cnt) - np.float32(hy)
    exact_x = ((xs + hx) % nx).sum() / len(xs) - hx
    exact_y = ((ys + hy) % ny).sum() / len(ys) - hy
    return (oy + int(-cy), ox + int(-cx)), (oy + int(-exact_y), ox + int(-exact_x))


def test_centre_image_float32_raster_order(o32):
    """A support whose coordinate sums pass 2^24: summed in float32 in raster
    order (as the reference sums them) the centre of mass lands one pixel
    away from the exact one, and the oracle moves the support by the float32
    result."""
    shape = (1024, 1024)
    for y0, x0, h, w in ((94, 102, 677, 775), (180, 40, 551, 431)):
        sup = np.zeros(shape, bool)
        sup[y0:y0 + h, x0:x0 + w] = True
        (dy, dx), exact = _raster_float32_centre_shift(sup)
        assert (dy, dx) != exact  # the summation order decides this case
        ph, _ = _phaser(o32, sup)
        ph.support_centre_image()
        assert np.array_equal((ph.flags & 1).astype(bool), np.roll(sup, (dy, dx), axis=(0, 1)))


def test_centre_image_empty_support(o32):
    sup = np.zeros((16, 16), bool)
    ph, _ = _phaser(o32, sup)
    ph.support_centre_image()
    assert not (ph.flags & 1).any()


@pytest.mark.parametrize("shape", [(48, 64), (12, 10, 14)])
def test_template_prepare_and_update(o32, shape):
    tmpl = _blob(shape, tuple(n // 2 for n in shape), min(shape) / 4).astype(np.complex64)
    absv, srt, frac = o32.template(tmpl, 2.0)
    b = np.abs(o32.blur(tmpl, 2.0))
    assert np.array_equal(absv, b.astype(np.float32))
    assert np.array_equal(srt, np.sort(absv.ravel())[::-1])
    assert frac == np.float32(np.count_nonzero(tmpl)) / np.float32(tmpl.size)
    ph, meas = _phaser(o32, np.zeros(shape, bool))
    for area in (1.0, 0.8, 1.3):
        t = ph.support_template_update((absv, srt, frac), area)
        k = int(np.float32(tmpl.size) * np.float32(area) * np.float32(frac))
        assert t == srt[k]
        got = (ph.flags & 1).astype(bool)
        assert np.array_equal(got, absv > srt[k])
        assert np.array_equal((ph.flags & 2) != 0, meas)
        # about area * template-area pixels are kept
        assert abs(int(got.sum()) - k) <= np.count_nonzero(srt == srt[k])


def test_support_array_runs_in_order(o32):
    """An update runs every algorithm of the array in order and consumes one iteration."""
    rng = np.random.default_rng(3)
    shape = (24, 24)
    sup = rng.random(shape) < 0.3
    alg = o32.alg(ER)
    a, _ = _phaser(o32, sup, seed=4)
    a.iterate(alg, 12, [o32.sup(SUP_STATIC, 5), o32.sup(SUP_CLOSE, 5, close_size=1),
                        o32.sup(SUP_CENTRE, 5)])
    b, _ = _phaser(o32, sup, seed=4)
    for n in (4, 4):
        b.iterate(alg, n)
        b.support_close_update(1)
        b.support_centre_image()
        b.iteration = b.iteration + 1
    b.iterate(alg, 2)
    assert a.support_log() == [4, 9]
    assert a.iteration == b.iteration == 12
    assert np.array_equal(a.flags, b.flags)
    assert np.array_equal(a.g1, b.g1)


def test_template_in_scheduler(o32):
    shape = (32, 32)
    tmpl = _blob(shape, (16, 16), 6).astype(np.complex64)
    t = o32.template(tmpl, 1.5)
    ph, _ = _phaser(o32, np.ones(shape, bool))
    ph.iterate(o32.alg(ER), 10, o32.sup(SUP_TEMPLATE, 5, values=[(0, 1.0), (10, 0.5)], template=t))
    assert ph.support_log() == [4, 9]
    area = o32.bezier([(0, 1.0), (10, 0.5)], 9)
    k = int(np.float32(tmpl.size) * np.float32(area) * np.float32(t[2]))
    assert np.array_equal((ph.flags & 1).astype(bool), t[0] > t[1][k])
