"""Phasing arrays whose axis lengths have no Stockham plan (products of 2, 3,
5 and 7 up to 8192): the engine runs such single reconstructions with
unfused iterations over Bluestein transforms (FFTW plans any length,
reference src/fft.c:79,95,163,185; the reference's CUDA engine takes any
size through cuFFT).  The same bars as the Stockham shapes: one iteration
within relative L2 1e-5 of the FP32 oracle, metrics within 1e-3 after 100
ER iterations, support masks bit-exact away from the threshold."""
import numpy as np
import pytest

from _helpers import ALG_IDS, GpuPhaser, random_case, rel

pytestmark = pytest.mark.gpu

SHAPES = [(47, 64), (22, 63), (11, 12, 16), (97, 11)]


def _oracle(o, alg, amp, meas, sup, g0, n=1, beta=0.9, constraints=0, gammas=(0.0, 0.0), sup_obj=None):
    ph = o.phaser(g0.shape)
    ph.set_inputs(amp, meas, sup, g0)
    a = o.alg(ALG_IDS[alg], constraints, beta=[(0, beta)], gamma1=gammas[0], gamma2=gammas[1])
    assert ph.iterate(a, n, sup_obj) == 0
    return ph


@pytest.mark.parametrize("alg", ["er", "hio", "raar", "dm"])
@pytest.mark.parametrize("c", [0, 2, 16])
@pytest.mark.parametrize("shape", SHAPES)
def test_one_iteration_any_length(sp, o32, alg, c, shape):
    obj, amp, meas, sup, g0 = random_case(shape, seed=c + 3 * len(shape))
    with GpuPhaser(sp, alg, amp, meas, sup, g0, beta=0.9, constraints=c) as gph:
        assert gph.iterate(1) == 0
        gam = gph.gammas() if alg == "dm" else (0.0, 0.0)
        oph = _oracle(o32, alg, amp, meas, sup, g0, beta=0.9, constraints=c, gammas=gam)
        assert rel(gph.get("model"), oph.g1) < 1e-5
        assert rel(gph.get("model_before_projection"), oph.gp) < 1e-5
        assert sp.spx_phaser_tiled(gph.ph) == 0


# lines beyond 4096: Stockham plans up to 8192 (fused passes), Bluestein up
# to 4096 (unfused iterations over 8192-point transforms)
@pytest.mark.parametrize("alg", ["hio", "dm"])
@pytest.mark.parametrize("shape", [(16, 8192), (6144, 8), (2, 4093)])
def test_one_iteration_long_lines(sp, o32, alg, shape):
    obj, amp, meas, sup, g0 = random_case(shape, seed=11)
    with GpuPhaser(sp, alg, amp, meas, sup, g0, beta=0.9) as gph:
        assert gph.iterate(1) == 0
        gam = gph.gammas() if alg == "dm" else (0.0, 0.0)
        oph = _oracle(o32, alg, amp, meas, sup, g0, beta=0.9, gammas=gam)
        e = rel(gph.get("model"), oph.g1)
        print(f"{shape} {alg}: relative L2 {e:.2e} from the FP32 oracle", flush=True)
        assert e < 1e-5


@pytest.mark.parametrize("shape", [(47, 64), (11, 12, 16)])
def test_er_metrics_any_length(sp, o32, shape):
    obj, amp, meas, sup, g0 = random_case(shape, seed=5)
    with GpuPhaser(sp, "er", amp, meas, sup, g0) as gph:
        oph = o32.phaser(shape)
        oph.set_inputs(amp, meas, sup, g0)
        worst = 0.0
        for _ in range(10):
            assert gph.iterate(10) == 0
            assert oph.iterate(o32.alg(4, 0), 10) == 0
            ef, fc = gph.efourier()
            oef, ofc = oph.efourier()
            er_, oer = gph.sp.sp_phaser_ereal(gph.ph), oph.ereal()
            worst = max(worst, abs(ef - oef) / oef, abs(fc - ofc) / ofc, abs(er_ - oer) / oer)
        assert worst <= 1e-3, worst
        assert rel(gph.get("fmodel"), o32.fft(oph.g1, -1)) < 1e-4


def test_area_shrinkwrap_any_length(sp, o32):
    import ctypes as C
    from test_gpu_parity import _support_mask_matches
    shape = (94, 70)
    from libspimage_amd import workloads
    p = workloads.make_pattern(shape, n_gauss=8, radius=(15, 12), photons=1e6, beamstop=2, seed=31)
    g = (p.start * 3 + p.obj).astype(np.complex64)
    meas, sup = p.mask.astype(bool), p.support.astype(bool)
    sarr = sp.sp_support_array_init(sp.sp_support_area_alloc(sp.smap([(0, 2.0)]), sp.smap([(0, 0.08)])), 1)
    with GpuPhaser(sp, "er", p.amplitudes, meas, sup, g, sup_array=sarr) as gph:
        assert sp.sp_support_array_update(sarr, gph.ph) == 0
        tg = C.c_float()
        assert sp.spx_phaser_last_threshold(gph.ph, C.byref(tg)) == 0
        gph.ph.contents.iteration += 1
        gmask = gph.get("support").real.astype(bool)
    oph = o32.phaser(shape)
    oph.set_inputs(p.amplitudes, meas, sup, g)
    to = oph.support_area_update(2.0, 0.08)
    assert abs(tg.value - to) <= 1e-5 * to, (tg.value, to)
    b = np.abs(o32.blur(np.abs(g).astype(np.complex64), 2.0))
    ok, amb = _support_mask_matches(gmask, (oph.flags & 1).astype(bool), b, to, band=1e-5 * to)
    assert ok, amb


def test_random_phase_start_any_length(sp):
    """SpModelRandomPhases: the start's transform has the measured moduli."""
    shape = (63, 53)
    obj, amp, meas, sup, _ = random_case(shape, seed=8)
    with GpuPhaser(sp, "hio", amp, meas, sup, None) as gph:
        assert sp.sp_phaser_init_model(gph.ph, None, sp.SpModelRandomPhases) == 0
        assert sp.sp_phaser_init_support(gph.ph, gph._img(sup.astype(np.float32)), 0, 0.0) == 0
        m = gph.get("model")
    F = np.fft.fftn(m.astype(np.complex128))
    assert np.allclose(np.abs(F)[meas], amp[meas], rtol=1e-4, atol=1e-4 * amp.max())


def test_batches_refuse_any_length(sp):
    # also with one reconstruction: the unfused path takes no per-slot tables
    for nrec in (1, 2):
        assert not bool(sp.lib.spx_batch_create(47, 64, 1, nrec))
        assert "batches need" in sp.last_error()
