"""Axis lengths with factors of 7 (2^a 3^b 5^c 7^d up to 8192): the Stockham
passes run them with radix-7 stages, so they take the fused iterations,
batches and the tuner like the 2-3-5 lengths (FFTW plans any length,
reference src/fft.c:79,95,163,185).  The same bars as the other Stockham
shapes: transforms within relative L2 2e-6 of the FP64 oracle, one iteration
within 1e-5 of the FP32 oracle, metrics within 1e-3 after 100 ER iterations,
a batch equal to its reconstructions run alone, bit for bit."""
import numpy as np
import pytest

from _helpers import ALG_IDS, GpuPhaser, random_case, rel
from test_gpu_batch import make_batch, patterns, run_single

pytestmark = pytest.mark.gpu


# plans (8 elements per thread at the edges): 7 alone; 7 7; 7 7 7; 8 7 8;
# 3 7 7 7 (edges differ); 10 7 2 10; 7 7 7 7; 7 16 7 7 (edge radix 7);
# 7 10 7 2 7; 8 4 4 7 8 (in one LDS buffer)
@pytest.mark.parametrize("shape", [(1, 7), (1, 49), (1, 343), (3, 448), (1, 1029), (2, 1400), (1, 2401), (1, 5488),
                                   (1, 6860), (1, 7168), (14, 21, 35), (98, 56)])
def test_fft_radix7_matches_oracle(sp, o64, shape):
    for L in shape:
        assert L == 1 or sp.spx_supported_length(L) == 1, L
    rng = np.random.default_rng(sum(shape) + 77)
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
    print(f"{shape}: forward {e_f:.2e} from the FP64 oracle, round trip {e_b:.2e}", flush=True)
    assert e_f < 2e-6
    assert e_b < 4e-6


def _oracle(o, alg, amp, meas, sup, g0, beta=0.9, constraints=0, gammas=(0.0, 0.0)):
    ph = o.phaser(g0.shape)
    ph.set_inputs(amp, meas, sup, g0)
    a = o.alg(ALG_IDS[alg], constraints, beta=[(0, beta)], gamma1=gammas[0], gamma2=gammas[1])
    assert ph.iterate(a, 1) == 0
    return ph


@pytest.mark.parametrize("alg", ["er", "hio", "raar", "dm"])
@pytest.mark.parametrize("c", [0, 2])
@pytest.mark.parametrize("shape", [(49, 64), (70, 63), (7, 12, 16), (448, 56), (14, 49, 10)])
def test_one_iteration_radix7(sp, o32, alg, c, shape):
    obj, amp, meas, sup, g0 = random_case(shape, seed=c + 5 * len(shape))
    with GpuPhaser(sp, alg, amp, meas, sup, g0, beta=0.9, constraints=c) as gph:
        assert gph.iterate(1) == 0
        gam = gph.gammas() if alg == "dm" else (0.0, 0.0)
        oph = _oracle(o32, alg, amp, meas, sup, g0, beta=0.9, constraints=c, gammas=gam)
        assert rel(gph.get("model"), oph.g1) < 1e-5
        assert rel(gph.get("model_before_projection"), oph.gp) < 1e-5


def test_er_metrics_radix7(sp, o32):
    shape = (98, 84)
    obj, amp, meas, sup, g0 = random_case(shape, seed=6)
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


@pytest.mark.parametrize("shape", [(56, 98), (14, 21, 28)])
def test_batch_radix7_equals_single_phasers(sp, shape):
    pats = patterns(shape, 3, seed=70)
    hio = sp.sp_phasing_hio_alloc(sp.smap([(0, 0.9)]), 0)
    b = make_batch(sp, pats, hio)
    b.iterate(9)
    got = b.model()
    ef, _ = b.efourier()
    b.close()
    for r, p in enumerate(pats):
        m, s, e = run_single(sp, p, "hio", 9)
        assert np.array_equal(got[r], m), r
        assert abs(ef[r] - e) <= 1e-6 * abs(e)


def test_fixed_array_transform_radix7(sp, o64):
    """spx_fft_noalloc takes the radix-7 plans up to 4096 points (its default
    plans may differ from the tuned ones: relative L2 1e-6 between them)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    shape = (49, 448)
    x = (np.random.default_rng(17).standard_normal(shape) * (1 + 1j)).astype(np.complex64)
    d, e = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(d), C.c_size_t(x.nbytes)) == 0
    assert hip.hipMalloc(C.byref(e), C.c_size_t(x.nbytes)) == 0
    try:
        assert hip.hipMemcpy(d, x.ctypes.data_as(C.c_void_p), C.c_size_t(x.nbytes), 1) == 0
        assert sp.spx_fft_device(d, e, shape[1], shape[0], 1, -1) == 0, sp.last_error()
        ref = np.empty_like(x)
        assert hip.hipMemcpy(ref.ctypes.data_as(C.c_void_p), e, C.c_size_t(x.nbytes), 2) == 0
    finally:
        hip.hipFree(d)
        hip.hipFree(e)
    out = np.empty_like(x)
    assert sp.spx_fft_noalloc(x.ctypes.data, out.ctypes.data, shape[1], shape[0], 1, -1) == 0, sp.last_error()
    assert rel(ref, o64.fft(x, -1)) < 2e-6
    assert rel(out, ref) <= 1e-6
