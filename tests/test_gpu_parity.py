"""GPU parity of the HIP engine against the oracle, through the C ABI.

Tolerances (float32 arithmetic like the reference, DOUBLE_PRECISION=OFF):
  * one transform or one iteration: relative L2 <= 1e-5 against the oracle's
    FP32 restatement (whose FFT is exact to float rounding);
  * error metrics after 100 ER iterations: within 1e-3 relative;
  * support masks: bit-exact wherever |blur| lies more than 1e-6 (relative to
    the threshold) away from the threshold.
"""
import ctypes as C

import numpy as np
import pytest

from _helpers import ALG_IDS, GpuPhaser, random_case, rel

pytestmark = pytest.mark.gpu

ONE_ITER_TOL = 1e-5


# ---------------------------------------------------------------------------
# transforms
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(8, 8), (16, 16), (48, 64), (96, 80), (384, 384), (1024, 1024),
                                   (125, 60), (16, 16, 16), (64, 48, 80), (20, 36, 30),
                                   # mixed-radix plans (6, 9, 10, 12, 15 stages): 10.10.10, 9.12.9,
                                   # 10.9.10, 15.3.15, 9.9.9, 12.5.12, 15.15, 5.15.2.5, 10.10
                                   (1000, 972), (900, 675), (729, 720), (225, 750), (100, 90, 96),
                                   # long lines: two LDS buffers up to 2304, one (in-place stages) from 2400
                                   (2048, 1536), (4096, 512), (2400, 3072), (2, 4096, 4000),
                                   # beyond 4096: 8.8.4.4.8 / 16.8.4.16, 9.9.9.9, 15.5.5.15, 12.9.6.12
                                   (16, 8192), (8192, 12), (6561, 8), (8, 5625), (3, 7776, 2)])
def test_fft_matches_oracle(sp, o64, shape):
    rng = np.random.default_rng(sum(shape))
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    img = sp.image_from_array(x)
    f = sp.sp_image_fft(img)
    assert bool(f)
    fw = sp.image_view(f).copy()
    assert f.contents.shifted == 1 and f.contents.phased == 1
    assert list(f.contents.detector.contents.image_center)[:2] == [shape[-1] / 2.0, shape[-2] / 2.0]
    b = sp.sp_image_ifft(f)
    back = sp.image_view(b).copy()
    for i in (img, f, b):
        sp.sp_image_free(i)
    ref = o64.fft(x, -1)
    assert rel(fw, ref) < 2e-6
    assert rel(back / x.size, x) < 3e-6


def test_fft_device_entry(sp, o64):
    """spx_fft_device on device buffers, in place, through the plain-pointer ABI."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    shape = (64, 96)
    x = (np.random.default_rng(3).standard_normal(shape) * (1 + 1j)).astype(np.complex64)
    d = C.c_void_p()
    assert hip.hipMalloc(C.byref(d), C.c_size_t(x.nbytes)) == 0
    try:
        assert hip.hipMemcpy(d, x.ctypes.data_as(C.c_void_p), C.c_size_t(x.nbytes), 1) == 0
        assert sp.spx_fft_device(d, d, shape[1], shape[0], 1, -1) == 0
        out = np.empty_like(x)
        assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), d, C.c_size_t(x.nbytes), 2) == 0
    finally:
        hip.hipFree(d)
    assert rel(out, o64.fft(x, -1)) < 2e-6
    assert sp.spx_fft_device(None, None, 4099, 4, 1, -1) < 0  # a prime above 4096
    assert "not supported" in sp.last_error()


# image transforms of lengths FFTW plans: 7-smooth ones (Stockham passes with
# radix-7 stages), and through Bluestein's method primes, 121 = 11^2, mixed
# with plan lengths, 2047 (padding 4096) and 4093, 4095 (padding 8192, the
# largest)
@pytest.mark.parametrize("shape", [(1, 97), (1, 2047), (49, 64), (448, 30), (13, 11, 17), (7, 1000), (3, 121, 8),
                                   (30, 7, 2), (1, 4093), (3, 4095), (2053, 4)])
def test_fft_any_length_matches_oracle(sp, o64, shape):
    rng = np.random.default_rng(sum(shape) + 7)
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    img = sp.image_from_array(x)
    f = sp.sp_image_fft(img)
    assert bool(f), sp.last_error()
    fw = sp.image_view(f).copy()
    b = sp.sp_image_ifft(f)
    back = sp.image_view(b).copy()
    for i in (img, f, b):
        sp.sp_image_free(i)
    e_f, e_b = rel(fw, o64.fft(x, -1)), rel(back / x.size, x)
    print(f"{shape}: forward {e_f:.2e} from the FP64 oracle, round trip {e_b:.2e}")
    assert e_f < 3e-6
    assert e_b < 5e-6


def test_fft_any_length_device_entry(sp, o64):
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    shape = (21, 77)
    x = (np.random.default_rng(5).standard_normal(shape) * (1 - 2j)).astype(np.complex64)
    d, e = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(d), C.c_size_t(x.nbytes)) == 0
    assert hip.hipMalloc(C.byref(e), C.c_size_t(x.nbytes)) == 0
    try:
        assert hip.hipMemcpy(d, x.ctypes.data_as(C.c_void_p), C.c_size_t(x.nbytes), 1) == 0
        assert sp.spx_fft_device(d, e, shape[1], shape[0], 1, +1) == 0, sp.last_error()
        out = np.empty_like(x)
        assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), e, C.c_size_t(x.nbytes), 2) == 0
    finally:
        hip.hipFree(d)
        hip.hipFree(e)
    assert rel(out, o64.fft(x, +1)) < 3e-6


@pytest.mark.parametrize("shape", [(49, 50), (47, 50)])
def test_blur_any_length_matches_oracle(sp, o32, shape):
    rng = np.random.default_rng(9)
    x = rng.random(shape).astype(np.complex64)
    img = sp.image_from_array(x)
    out = sp.sp_gaussian_blur(img, 2.0)
    assert bool(out), sp.last_error()
    got = sp.image_view(out).copy()
    sp.sp_image_free(out)
    sp.sp_image_free(img)
    assert rel(got, o32.blur(x, 2.0)) < 3e-6


# ---------------------------------------------------------------------------
# one iteration of every algorithm under every constraint
# ---------------------------------------------------------------------------
CONSTRAINTS = [0, 1, 2, 2 | 8, 4, 4 | 8, 16]


def oracle_run(o, alg, amp, meas, sup, g0, n=1, beta=0.9, constraints=0, gammas=(0.0, 0.0), sup_obj=None):
    ph = o.phaser(g0.shape)
    ph.set_inputs(amp, meas, sup, g0)
    a = o.alg(ALG_IDS[alg], constraints, beta=[(0, beta)], gamma1=gammas[0], gamma2=gammas[1])
    assert ph.iterate(a, n, sup_obj) == 0
    return ph


@pytest.mark.parametrize("alg", ["er", "hio", "raar", "dm"])
@pytest.mark.parametrize("c", CONSTRAINTS)
@pytest.mark.parametrize("shape", [(64, 64), (128, 128), (16, 16, 16), (90, 100)])
def test_one_iteration(sp, o32, alg, c, shape):
    # 64 x 64 runs in the row-major layout, 128 x 128 in the 4-column tiled
    # layout of the Fourier-space arrays, 16^3 is 3-D
    obj, amp, meas, sup, g0 = random_case(shape, seed=c + 10 * len(shape))
    with GpuPhaser(sp, alg, amp, meas, sup, g0, beta=0.9, constraints=c) as gph:
        assert gph.iterate(1) == 0
        gam = gph.gammas() if alg == "dm" else (0.0, 0.0)
        oph = oracle_run(o32, alg, amp, meas, sup, g0, beta=0.9, constraints=c, gammas=gam)
        assert rel(gph.get("model"), oph.g1) < ONE_ITER_TOL
        assert rel(gph.get("old_model"), oph.g0) < ONE_ITER_TOL
        assert rel(gph.get("model_before_projection"), oph.gp) < ONE_ITER_TOL
        assert rel(gph.get("model_change"), oph.g1 - oph.g0) < ONE_ITER_TOL
        assert gph.ph.contents.iteration == 1


def test_zero_model_takes_non_finite_branch(sp, o32):
    """A zero start model makes every measured |G| zero: amp/0 is not finite,
    so the projection sets G = (amp, 0) (src/phasing.c:1092-1097)."""
    obj, amp, meas, sup, _ = random_case((32, 32), seed=1)
    g0 = np.zeros((32, 32), np.complex64)
    with GpuPhaser(sp, "er", amp, meas, sup, g0) as gph:
        gph.iterate(1)
        oph = oracle_run(o32, "er", amp, meas, sup, g0)
        assert np.isfinite(gph.get("model")).all()
        assert rel(gph.get("model"), oph.g1) < ONE_ITER_TOL


@pytest.mark.parametrize("shape", [(32, 32), (128, 128)])
def test_amplitude_margins(sp, o32, shape):
    obj, amp, meas, sup, g0 = random_case(shape, seed=2)
    lo, hi = amp * 0.8, amp * 1.2
    with GpuPhaser(sp, "hio", amp, meas, sup, g0, constraints=32) as gph:
        sp.sp_phaser_set_amplitudes_margins(gph.ph, gph._img(lo), gph._img(hi))
        gph.iterate(3)
        oph = o32.phaser(shape)
        oph.set_inputs(amp, meas, sup, g0)
        oph.set_margins(lo, hi)
        oph.iterate(o32.alg(1, 32, beta=[(0, 0.9)]), 3)
        assert rel(gph.get("model"), oph.g1) < 3 * ONE_ITER_TOL


@pytest.mark.parametrize("shape", [(32, 32), (128, 128)])
def test_phased_amplitude_objective(sp, o32, shape):
    obj, amp, meas, sup, g0 = random_case(shape, seed=3)
    phased = np.fft.fftn(obj).astype(np.complex64)
    with GpuPhaser(sp, "er", amp, meas, sup, g0) as gph:
        sp.sp_phaser_set_phased_amplitudes(gph.ph, gph._img(phased, meas))
        sp.sp_phaser_set_objective(gph.ph, sp.SpRecoverAmplitudes)
        gph.iterate(2)
        oph = o32.phaser(shape)
        oph.set_inputs(amp, meas, sup, g0)
        oph.phased_amp[...] = phased
        oph.objective = 1
        oph.iterate(o32.alg(4), 2)
        assert rel(gph.get("model"), oph.g1) < 2 * ONE_ITER_TOL


def test_tiled_layout_equals_row_major_layout(sp):
    """The 4-column tiled layout of the Fourier-space arrays changes where
    values live, not the arithmetic: with the same transform plans
    (SPX_FFT_E=8 while this test's lengths are first seen) both layouts give
    bit-identical models, Fourier errors and Fourier models, for HIO, RAAR,
    DM with margins, and ER with the phased-amplitude objective."""
    import os
    shape = (200, 640)  # lengths used by no other test, so tuned here first
    obj, amp, meas, sup, g0 = random_case(shape, seed=9)
    phased = np.fft.fftn(obj).astype(np.complex64)
    lo, hi = amp * 0.9, amp * 1.1
    saved = {k: os.environ.get(k) for k in ("SPX_FFT_E", "SPX_NO_T4", "SPX_T4")}
    os.environ["SPX_FFT_E"] = "8"
    try:
        for alg, c, extra in (("hio", 0, None), ("raar", 2, None), ("dm", 32, "margins"), ("er", 0, "phased")):
            out = []
            for tiled in ("0", "1"):
                os.environ["SPX_NO_T4"] = "0" if tiled == "1" else "1"
                os.environ["SPX_T4"] = tiled
                with GpuPhaser(sp, alg, amp, meas, sup, g0, beta=0.8, constraints=c) as gph:
                    if extra == "margins":
                        sp.sp_phaser_set_amplitudes_margins(gph.ph, gph._img(lo), gph._img(hi))
                    if extra == "phased":
                        sp.sp_phaser_set_phased_amplitudes(gph.ph, gph._img(phased, meas))
                        sp.sp_phaser_set_objective(gph.ph, sp.SpRecoverAmplitudes)
                    assert gph.iterate(4) == 0
                    out.append((gph.get("model"), gph.efourier(), gph.get("fmodel")))
            (m0, e0, f0), (m1, e1, f1) = out
            assert np.array_equal(m0, m1), alg
            assert e0 == e1, alg
            assert np.array_equal(f0, f1), alg
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def test_hio_iteration_2048(sp, o32):
    """2048 x 2048 (tiled layout with 4 lines of 2048 per work-group) against the oracle."""
    obj, amp, meas, sup, g0 = random_case((2048, 2048), seed=5, fill=0.1)
    with GpuPhaser(sp, "hio", amp, meas, sup, g0, beta=0.9) as gph:
        assert gph.iterate(1) == 0
        oph = oracle_run(o32, "hio", amp, meas, sup, g0, beta=0.9)
        assert rel(gph.get("model"), oph.g1) < ONE_ITER_TOL


def test_full_size_hio_iteration_1024(sp, o32):
    """One HIO iteration at 1024^2 on the seeded blob pattern."""
    from libspimage_amd import workloads
    p = workloads.blob_1024()
    meas = p.mask.astype(bool)
    sup = p.support.astype(bool)
    with GpuPhaser(sp, "hio", p.amplitudes, meas, sup, p.start, beta=0.9) as gph:
        gph.iterate(1)
        oph = oracle_run(o32, "hio", p.amplitudes, meas, sup, p.start, beta=0.9)
        assert rel(gph.get("model"), oph.g1) < ONE_ITER_TOL


# ---------------------------------------------------------------------------
# trajectories and metrics
# ---------------------------------------------------------------------------
def test_er_metric_trajectory_100(sp, o32):
    from libspimage_amd import workloads
    p = workloads.disc_scaled(128, seed=7)
    meas, sup = p.mask.astype(bool), p.support.astype(bool)
    with GpuPhaser(sp, "er", p.amplitudes, meas, sup, p.start) as gph:
        oph = o32.phaser((128, 128))
        oph.set_inputs(p.amplitudes, meas, sup, p.start)
        alg = o32.alg(4)
        for k in range(10):
            gph.iterate(10)
            oph.iterate(alg, 10)
            eg, fg = gph.efourier()
            eo, fo = oph.efourier()
            assert abs(eg - eo) <= 1e-3 * abs(eo), (k, eg, eo)
            assert abs(fg - fo) <= 1e-3 * abs(fo)
            er_g = sp.sp_phaser_ereal(gph.ph)
            assert abs(er_g - oph.ereal()) <= 1e-3 * abs(oph.ereal()) + 1e-7
        assert abs(sp.sp_phaser_support_fraction(gph.ph) - oph.support_fraction()) < 1e-7
        assert rel(gph.get("model"), oph.g1) < 1e-3


def test_fmodel_and_masks(sp, o64):
    obj, amp, meas, sup, g0 = random_case((48, 40), seed=4)
    with GpuPhaser(sp, "hio", amp, meas, sup, g0) as gph:
        gph.iterate(5)
        model = gph.get("model")
        fm = sp.sp_phaser_fmodel_with_mask(gph.ph)
        assert rel(sp.image_view(fm), o64.fft(model, -1)) < 2e-6
        assert np.array_equal(sp.mask_view(fm), meas.astype(np.int32))
        s = gph.get("support")
        assert np.array_equal(s.real.astype(bool), sup)
        ms = sp.sp_phaser_model_with_support(gph.ph)
        assert np.array_equal(sp.mask_view(ms), sup.astype(np.int32))
        a = sp.sp_phaser_amplitudes(gph.ph)
        assert np.array_equal(sp.image_view(a).real, amp)


# ---------------------------------------------------------------------------
# model initialisation (src/phasing.c:606-724) by its defining properties
# ---------------------------------------------------------------------------
def test_model_initialisations(sp, o64):
    obj, amp, meas, sup, g0 = random_case((32, 48), seed=5)
    with GpuPhaser(sp, "er", amp, meas, sup, None) as gph:
        sp.spx_phaser_set_seed(gph.ph, 1234)
        assert sp.sp_phaser_init_model(gph.ph, None, sp.SpModelRandomPhases) == 0
        m = gph.get("model")
        assert np.allclose(np.abs(o64.fft(m, -1)), amp, atol=2e-5 * amp.max())
        assert sp.sp_phaser_init_model(gph.ph, None, sp.SpModelZeroPhases) == 0
        F = o64.fft(gph.get("model"), -1)
        assert np.allclose(F, amp, atol=2e-5 * amp.max())
        assert sp.sp_phaser_init_model(gph.ph, None, sp.SpModelRandomValues) == 0
        m = gph.get("model").astype(np.complex128)
        # sum |g|^2 = sum amp^2 / N (Parseval; src/phasing.c:669-679 scales to the mean)
        target = np.sum(amp.astype(np.float64) ** 2) / amp.size
        assert abs(np.sum(np.abs(m) ** 2) - target) < 1e-4 * target
        assert sp.sp_phaser_init_model(gph.ph, None, sp.SpModelRandomPhases | sp.SpModelMaskedOutZeroed) == 0
        F = o64.fft(gph.get("model"), -1)
        assert np.abs(F[~meas]).max() < 1e-4 * amp.max()
        assert sp.sp_phaser_init_model(gph.ph, None, 0) == -3
        # a fresh model restarts g0 and gp from zero
        assert np.abs(gph.get("old_model")).max() == 0


def test_patterson_support(sp):
    obj, amp, meas, sup, g0 = random_case((64, 64), seed=6, beamstop=False)
    with GpuPhaser(sp, "er", amp, meas, None, g0) as gph:
        assert sp.sp_phaser_init_support(gph.ph, None, sp.SpSupportFromPatterson, 0.04) == 0
        s = gph.get("support").real.astype(bool)
        pat = np.abs(np.fft.ifftn(amp.astype(np.float64) ** 2) * amp.size)
        t = pat.max() * 0.04
        clear = np.abs(pat - t) > 1e-5 * t
        assert np.array_equal(s[clear], (pat > t)[clear])


# ---------------------------------------------------------------------------
# shrinkwrap and the support scheduler
# ---------------------------------------------------------------------------
def _support_mask_matches(gmask, omask, blur_abs, t, band=0.0):
    """Masks must agree wherever |blur| is further than 1e-6 t (and `band`)
    from the threshold t."""
    clear = np.abs(blur_abs - t) > max(1e-6 * max(t, 1e-30), band)
    return np.array_equal(gmask[clear], omask[clear]), int((~clear).sum())


@pytest.mark.parametrize("kind", ["area", "threshold"])
@pytest.mark.parametrize("shape", [(128, 128), (24, 32, 40)])
def test_shrinkwrap_in_scheduler(sp, o32, kind, shape):
    from libspimage_amd import workloads
    p = workloads.make_pattern(shape, n_gauss=8, radius=tuple(s / 6 for s in shape), photons=1e6, beamstop=2,
                               seed=11)
    meas, sup = p.mask.astype(bool), p.support.astype(bool)
    blur = [(0, 3.0), (40, 0.7)]
    vals = [(0, 0.08)] if kind == "area" else [(0, 0.2)]
    bmap, vmap = sp.smap(blur), sp.smap(vals)
    salg = sp.sp_support_area_alloc(bmap, vmap) if kind == "area" else sp.sp_support_threshold_alloc(bmap, vmap)
    sarr = sp.sp_support_array_init(salg, 10)
    with GpuPhaser(sp, "raar", p.amplitudes, meas, sup, p.start, beta=0.87, sup_array=sarr) as gph:
        oph = o32.phaser(shape)
        oph.set_inputs(p.amplitudes, meas, sup, p.start)
        from oracle.oracle import SUP_AREA, SUP_THRESHOLD
        osup = o32.sup(SUP_AREA if kind == "area" else SUP_THRESHOLD, period=10, blur=blur, values=vals)
        # 9 iterations, then one support update that consumes iteration 9
        assert gph.iterate(10) == 0
        oph.iterate(o32.alg(2, 0, beta=[(0, 0.87)]), 10, osup)
        assert oph.support_log() == [9]
        assert gph.ph.contents.iteration == 10
        gmask = gph.get("support").real.astype(bool)
        omask = (oph.flags & 1).astype(bool)
        # nine iterations of float32 rounding differences: 1e-4 (1e-5 per iteration)
        gmodel = gph.get("model")
        assert rel(gmodel, oph.g1) < 1e-4
        # the blurs each engine thresholded (its own iterate; the support
        # update consumed iteration 9 without changing it)
        import ctypes as C
        sig = o32.bezier(blur, 9)
        b = np.abs(o32.blur(np.abs(oph.g1).astype(np.complex64), sig))
        bg = np.abs(o32.blur(np.abs(gmodel).astype(np.complex64), sig))
        tg = C.c_float()
        assert sp.spx_phaser_last_threshold(gph.ph, C.byref(tg)) == 0
        # the GPU mask is its own blur thresholded (the identical-input bar)
        ok, amb = _support_mask_matches(gmask, bg > tg.value, bg, tg.value)
        assert ok, f"GPU mask differs from its own blur away from the threshold ({amb} ambiguous pixels)"
        # against the oracle: pixels further from the oracle's threshold than
        # the largest difference between the two blurs fall the same way
        ok, amb = _support_mask_matches(gmask, omask, b, oph_threshold(oph), band=float(np.abs(bg - b).max()))
        assert ok, f"support masks differ away from the threshold ({amb} ambiguous pixels)"


def oph_threshold(oph):
    return float(oph.o.lib.or_last_threshold(oph.h))


@pytest.mark.parametrize("kind", ["area", "threshold"])
@pytest.mark.parametrize("shape,sigma,val", [((256, 256), 3.0, 0.08), ((256, 256), 0.7, 0.08),
                                             ((1024, 1024), 3.0, 0.08), ((32, 32, 32), 2.0, 0.05),
                                             ((96, 80), 1.5, 0.3)])
def test_shrinkwrap_identical_input(sp, o32, kind, shape, sigma, val):
    """Support update on the same iterate: threshold within 1e-6 relative and
    the mask bit-exact away from it (src/support_update.c:136-179)."""
    from libspimage_amd import workloads
    p = workloads.make_pattern(shape, n_gauss=10, radius=tuple(s / 6 for s in shape), photons=1e6, beamstop=2,
                               seed=21)
    g = p.start * 3 + np.fft.ifftn(np.fft.fftn(p.obj)).astype(np.complex64)
    meas, sup = p.mask.astype(bool), p.support.astype(bool)
    rel_t = 0.2 if kind == "threshold" else val
    bmap, vmap = sp.smap([(0, sigma)]), sp.smap([(0, rel_t)])
    salg = sp.sp_support_area_alloc(bmap, vmap) if kind == "area" else sp.sp_support_threshold_alloc(bmap, vmap)
    sarr = sp.sp_support_array_init(salg, 1)
    with GpuPhaser(sp, "er", p.amplitudes, meas, sup, g, sup_array=sarr) as gph:
        assert sp.sp_support_array_update(sarr, gph.ph) == 0
        tg = C.c_float()
        assert sp.spx_phaser_last_threshold(gph.ph, C.byref(tg)) == 0
        gph.ph.contents.iteration += 1  # refresh the getters' cache
        gmask = gph.get("support").real.astype(bool)
    oph = o32.phaser(shape)
    oph.set_inputs(p.amplitudes, meas, sup, g)
    to = oph.support_area_update(sigma, val) if kind == "area" else oph.support_threshold_update(sigma, rel_t)
    assert abs(tg.value - to) <= 1e-6 * to + 1e-30, (tg.value, to)
    b = np.abs(o32.blur(np.abs(g).astype(np.complex64), sigma))
    ok, amb = _support_mask_matches(gmask, (oph.flags & 1).astype(bool), b, to)
    assert ok, f"masks differ away from the threshold ({amb} pixels within 1e-6)"
    if kind == "area":
        k = int(np.float32(b.size) * np.float32(val))
        assert abs(int(gmask.sum()) - k) <= amb + 1


def test_support_fraction_after_area_update(sp):
    obj, amp, meas, sup, g0 = random_case((64, 64), seed=12)
    salg = sp.sp_support_area_alloc(sp.smap([(0, 2.0)]), sp.smap([(0, 0.1)]))
    sarr = sp.sp_support_array_init(salg, 5)
    with GpuPhaser(sp, "er", amp, meas, sup, g0, sup_array=sarr) as gph:
        gph.iterate(5)
        f = sp.sp_phaser_support_fraction(gph.ph)
        # exactly k = (int)(N * 0.1) pixels lie above the k-th largest value
        # unless |blur| ties at the threshold
        assert f <= int(np.float32(4096) * np.float32(0.1)) / 4096 + 1e-7
        assert f >= 0.09


# ---------------------------------------------------------------------------
# error codes of the boundary (src/phasing.c:559-566,895-917)
# ---------------------------------------------------------------------------
def test_error_codes(sp):
    ph = sp.sp_phaser_alloc()
    er = sp.sp_phasing_er_alloc(0)
    assert sp.sp_phaser_init(None, er, None, sp.SpEngineHIP) == -1
    assert sp.sp_phaser_init(ph, None, None, sp.SpEngineHIP) == -2
    assert sp.sp_phaser_init(ph, er, None, sp.SpEngineCPU) == -9
    assert sp.sp_phaser_init(ph, er, None, sp.SpEngineHIP) == 0
    assert sp.sp_phaser_iterate(ph, 1) == -3  # no model
    amp = np.ones((16, 16), np.float32)
    img = sp.image_from_array(amp, np.ones((16, 16)))
    sp.sp_phaser_set_amplitudes(ph, img)
    assert sp.sp_phaser_init_model(ph, None, sp.SpModelRandomPhases) == 0
    assert sp.sp_phaser_iterate(ph, 1) == 0  # flags exist (set by the amplitudes)
    assert sp.sp_phaser_iterate(None, 1) == -1
    sp.sp_phaser_free(ph)
    sp.sp_image_free(img)


def test_reference_convergence_case(sp):
    """tests/PhasingTests.c:10-59 on the HIP engine: 4x4 object, 8x8 frame, HIO."""
    solved = False
    for attempt in range(10):
        rng = np.random.default_rng(200 + attempt)
        obj = np.zeros((8, 8))
        obj[:4, :4] = rng.random((4, 4))
        amp = np.abs(np.fft.fftn(obj)).astype(np.float32)
        with GpuPhaser(sp, "hio", amp, np.ones((8, 8), bool), obj != 0, None, beta=0.9) as gph:
            sp.spx_phaser_set_seed(gph.ph, 77 + attempt)
            assert sp.sp_phaser_init_model(gph.ph, None, sp.SpModelRandomPhases) == 0
            for _ in range(300):
                gph.iterate(1)
                ch = sp.image_view(sp.sp_phaser_model_change(gph.ph))
                if np.sum(np.abs(ch.astype(np.complex128)) ** 2) <= 1e-10:
                    solved = True
                    break
        if solved:
            break
    assert solved
