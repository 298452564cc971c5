"""The transform that allocates nothing (spx_fft_noalloc) against the entry
points it stands in for, sp_image_fft_fast / sp_image_ifft_fast, which
allocate a device staging buffer on every call.

Two calls in a row on each case, the second on a smaller array than the
first (the fixed device array keeps the first call's contents).  The first
call is at the largest accepted size, 2^21 elements.  Each result must be
within relative L2 1e-6 of the allocating entry point's (the two may pick
different transform plans, tuned against default, so they need not agree
bit for bit) and within 2e-6 of the FP64 oracle's transform.
"""
import numpy as np
import pytest

from _helpers import rel

pytestmark = pytest.mark.gpu

MAX_ELEMENTS = 1 << 21


def _allocating(sp, a, direction):
    img = sp.image_from_array(np.ascontiguousarray(a, np.complex64))
    out = sp.image_from_array(np.zeros(a.shape, np.complex64))
    (sp.sp_image_fft_fast if direction < 0 else sp.sp_image_ifft_fast)(img, out)
    r = sp.image_view(out).copy()
    sp.sp_image_free(out)
    sp.sp_image_free(img)
    return r


def _fixed(sp, a, direction):
    a = np.ascontiguousarray(a, np.complex64)
    out = np.full_like(a, np.nan)
    nz, ny, nx = (1,) * (3 - a.ndim) + a.shape
    rc = sp.spx_fft_noalloc(a.ctypes.data, out.ctypes.data, nx, ny, nz, direction)
    assert rc == 0, (rc, sp.last_error())
    return out


@pytest.mark.parametrize("cases", [
    [((1024, 2048), -1), ((96, 80), -1)],
    [((2048, 1024), +1), ((48, 60), +1)],
    [((128, 128, 128), -1), ((16, 24, 20), +1)],
])
def test_fixed_array_transform_matches_allocating_one(sp, o64, cases):
    rng = np.random.default_rng(len(cases[0][0]) * 7 + cases[0][1])
    for shape, direction in cases:
        assert np.prod(shape) <= MAX_ELEMENTS
        a = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
        got = _fixed(sp, a, direction)
        ref = _allocating(sp, a, direction)
        exact = o64.fft(a, direction)
        e_ref, e_exact = rel(got, ref), rel(got, exact)
        print(f"{shape} direction {direction:+d}: {e_ref:.2e} from the allocating transform, "
              f"{e_exact:.2e} from the FP64 oracle", flush=True)
        assert e_ref <= 1e-6
        assert e_exact <= 2e-6


def test_fixed_array_transform_rejects_too_large(sp):
    a = np.zeros(4, np.complex64)
    assert sp.spx_fft_noalloc(a.ctypes.data, a.ctypes.data, 2048, 1025, 1, -1) == -12
    assert "fixed device array" in sp.last_error()


def test_fixed_array_transform_rejects_bluestein_lengths(sp):
    # the chirp tables of lengths without a Stockham plan would be allocated
    a = np.zeros(47 * 64, np.complex64)
    assert sp.spx_fft_noalloc(a.ctypes.data, a.ctypes.data, 64, 47, 1, -1) == -11
    assert "not supported" in sp.last_error()
