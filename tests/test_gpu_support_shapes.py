"""GPU parity of the closing, centre-image and template support updates
(reference src/support_update.c:47-85,181-389) against the oracle, through
the C ABI.

Closing and centre image are bit operations on the support: the masks must be
identical.  The template update thresholds the blurred template: given the
same blurred values the mask must be identical; against the oracle's own blur
(float rounding apart) it must agree wherever the value lies further from the
threshold than the two blurs differ.
"""
import numpy as np
import pytest

from _helpers import GpuPhaser, rel

pytestmark = pytest.mark.gpu


def _blob(shape, centre, radius):
    grids = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    return sum((g - c) ** 2 for g, c in zip(grids, centre)) <= radius ** 2


def _holey(shape, seed):
    rng = np.random.default_rng(seed)
    sup = np.zeros(shape, bool)
    for _ in range(3):
        sup |= _blob(shape, tuple(rng.integers(0, n) for n in shape), min(shape) / 6)
    sup &= rng.random(shape) > 0.3
    sup |= rng.random(shape) < 0.02
    return sup


def _inputs(shape, seed):
    rng = np.random.default_rng(seed)
    amp = rng.random(shape).astype(np.float32)
    meas = rng.random(shape) > 0.1
    g = ((rng.random(shape) - 0.5) + 1j * (rng.random(shape) - 0.5)).astype(np.complex64)
    return amp, meas, g


def _update_once(sp, gph, sarr):
    assert sp.sp_support_array_update(sarr, gph.ph) == 0
    gph.ph.contents.iteration += 1  # the getters cache per iteration
    return gph.get("support").real.astype(bool)


def _oracle(o32, amp, meas, sup, g):
    oph = o32.phaser(sup.shape)
    oph.set_inputs(amp, meas, sup, g)
    return oph


@pytest.mark.parametrize("shape", [(64, 64), (128, 128), (20, 24, 30), (250, 96)])
@pytest.mark.parametrize("size", [1, 3])
def test_close(sp, o32, shape, size):
    sup = _holey(shape, seed=size * 7 + len(shape))
    amp, meas, g = _inputs(shape, 1)
    sarr = sp.sp_support_array_init(sp.sp_support_close_alloc(size), 1)
    with GpuPhaser(sp, "er", amp, meas, sup, g, sup_array=sarr) as gph:
        got = _update_once(sp, gph, sarr)
        model = gph.get("model")
    oph = _oracle(o32, amp, meas, sup, g)
    oph.support_close_update(size)
    want = (oph.flags & 1).astype(bool)
    assert not np.array_equal(want, sup)
    assert np.array_equal(got, want)
    assert np.array_equal(model, g)  # the iterate is untouched


CENTRE_CASES = [((32, 40), (9, 12), 2.5), ((32, 40), (15, 12), 2.5), ((32, 40), (8, 19), 2.5),
                ((30, 27), (20, 5), 2.5), ((12, 15, 16), (3, 4, 5), 1.5), ((12, 15, 16), (5, 4, 7), 1.5),
                ((12, 15, 16), (3, 4, 5), 2.5), ((24, 20, 16), (11, 9, 7), 4.0)]


@pytest.mark.parametrize("shape,centre,radius", CENTRE_CASES)
def test_centre_image(sp, o32, shape, centre, radius):
    sup = _blob(shape, centre, radius)
    amp, meas, g = _inputs(shape, 2)
    sarr = sp.sp_support_array_init(sp.sp_support_centre_image_alloc(), 1)
    with GpuPhaser(sp, "er", amp, meas, sup, g, sup_array=sarr) as gph:
        got = _update_once(sp, gph, sarr)
        model = gph.get("model")
    oph = _oracle(o32, amp, meas, sup, g)
    oph.support_centre_image()
    assert np.array_equal(got, (oph.flags & 1).astype(bool))
    assert got.sum() == sup.sum()
    assert np.array_equal(model, g)


@pytest.mark.parametrize("rect", [(94, 102, 677, 775), (180, 40, 551, 431), (100, 50, 600, 880)])
def test_centre_image_float_sums_past_2_24(sp, o32, rect):
    """Half of a 1024^2 image in the support: the centre-of-mass sums pass
    2^24, where float addition rounds and the order of the sum decides the
    result; the engine sums in the reference's raster order.  For the first
    two rectangles the float32 raster-order result is one pixel away from the
    exact centre of mass (tests/test_oracle_support_shapes.py pins this)."""
    shape = (1024, 1024)
    y0, x0, h, w = rect
    sup = np.zeros(shape, bool)
    sup[y0:y0 + h, x0:x0 + w] = True
    amp, meas, g = _inputs(shape, 3)
    sarr = sp.sp_support_array_init(sp.sp_support_centre_image_alloc(), 1)
    with GpuPhaser(sp, "er", amp, meas, sup, g, sup_array=sarr) as gph:
        got = _update_once(sp, gph, sarr)
    oph = _oracle(o32, amp, meas, sup, g)
    oph.support_centre_image()
    assert np.array_equal(got, (oph.flags & 1).astype(bool))


def test_centre_image_volume_sums_past_2_24(sp, o32):
    """A 128^3 volume with half of it in the support: the z and y sums pass
    2^24 and many 4096-voxel chunks are summed pixel by pixel."""
    shape = (128, 128, 128)
    sup = np.zeros(shape, bool)
    sup[10:90, 20:110, 5:100] = True
    sup[40:44, :, :] = False
    amp, meas, g = _inputs(shape, 4)
    sarr = sp.sp_support_array_init(sp.sp_support_centre_image_alloc(), 1)
    with GpuPhaser(sp, "er", amp, meas, sup, g, sup_array=sarr) as gph:
        got = _update_once(sp, gph, sarr)
    oph = _oracle(o32, amp, meas, sup, g)
    oph.support_centre_image()
    want = (oph.flags & 1).astype(bool)
    assert not np.array_equal(want, sup)
    assert np.array_equal(got, want)


def _cabs32(z):
    z = np.asarray(z, np.complex64)
    return np.sqrt(z.imag * z.imag + z.real * z.real)


@pytest.mark.parametrize("shape", [(128, 128), (96, 80), (16, 20, 24)])
def test_template(sp, o32, shape):
    tmpl = _blob(shape, tuple(n // 2 for n in shape), min(shape) / 4).astype(np.float32)
    amp, meas, g = _inputs(shape, 4)
    area = [(0, 1.2), (20, 0.8)]
    img = sp.image_from_array(tmpl)
    amap = sp.smap(area)
    salg = sp.sp_support_template_alloc(img, 2.0, amap)
    assert bool(salg)
    blured = sp.image_view(sp.sp_gaussian_blur(img, 2.0)).copy()  # the engine's blur of the template
    sarr = sp.sp_support_array_init(salg, 1)
    start = np.ones(shape, bool)
    with GpuPhaser(sp, "er", amp, meas, start, g, sup_array=sarr) as gph:
        gph.ph.contents.iteration = 13
        got = _update_once(sp, gph, sarr)
    a = o32.bezier(area, 13)
    # identical blurred values: identical mask
    absg = _cabs32(blured)
    frac = np.float32(np.count_nonzero(tmpl)) / np.float32(tmpl.size)
    srt = np.sort(absg.ravel())[::-1].copy()
    oph = _oracle(o32, amp, meas, start, g)
    t = oph.support_template_update((absg, srt, float(frac)), a)
    assert np.array_equal(got, (oph.flags & 1).astype(bool))
    # against the oracle's own blur of the template
    absv, srto, fraco = o32.template(tmpl.astype(np.complex64), 2.0)
    assert fraco == frac
    assert rel(absg, absv) < 1e-6
    oph2 = _oracle(o32, amp, meas, start, g)
    to = oph2.support_template_update((absv, srto, fraco), a)
    band = float(np.abs(absg - absv).max())
    clear = np.abs(absv - to) > max(1e-6 * to, band)
    assert np.array_equal(got[clear], (oph2.flags & 1).astype(bool)[clear])
    assert abs(t - to) <= band + 1e-6 * to


def test_template_and_close_in_scheduler(sp, o32):
    """A support array of template then closing, every 5 iterations of HIO:
    the same masks and, to float rounding, the same iterates as the oracle."""
    from oracle.oracle import SUP_CLOSE, SUP_TEMPLATE
    shape = (64, 64)
    tmpl = (_blob(shape, (32, 32), 12) & ~_blob(shape, (30, 34), 3)).astype(np.float32)
    amp, meas, g = _inputs(shape, 5)
    area = [(0, 1.0), (20, 0.9)]
    img = sp.image_from_array(tmpl)
    salg = sp.sp_support_template_alloc(img, 1.5, sp.smap(area))
    blured = sp.image_view(sp.sp_gaussian_blur(img, 1.5)).copy()
    sarr = sp.sp_support_array_init(salg, 5)
    sp.sp_support_array_append(sarr, sp.sp_support_close_alloc(2))
    start = _blob(shape, (32, 32), 14)
    with GpuPhaser(sp, "hio", amp, meas, start, g, beta=0.9, sup_array=sarr) as gph:
        assert gph.iterate(12) == 0
        gmask = gph.get("support").real.astype(bool)
        gmodel = gph.get("model")
    absg = _cabs32(blured)
    tdesc = (absg, np.sort(absg.ravel())[::-1].copy(), float(np.float32(np.count_nonzero(tmpl)) / tmpl.size))
    oph = _oracle(o32, amp, meas, start, g)
    oph.iterate(o32.alg(1, 0, beta=[(0, 0.9)]), 12,
                [o32.sup(SUP_TEMPLATE, 5, values=area, template=tdesc), o32.sup(SUP_CLOSE, 5, close_size=2)])
    assert oph.support_log() == [4, 9]
    assert np.array_equal(gmask, (oph.flags & 1).astype(bool))
    assert rel(gmodel, oph.g1) < 1e-4
