"""Pins the oracle's metrics, Fourier constraints, phased projection and model
change (oracle/spimage_oracle.c: or_efourier, or_ereal, or_support_fraction,
or_apply_fourier_constraints, or_phased_projection, or_model_change).

Reference semantics restated by the oracle: src/phasing.c:803-893 (metrics),
:1061-1068 (centrosymmetric constraint), :1103-1109 (phased-amplitude
projection), :533-554 (model change).  Every expected value here is a closed
form of a hand-built input, not the oracle's formula retyped:
  * a delta g1 = c at the origin transforms to the constant c, and a plane
    wave exp(2 pi i k0.x / n) to the single bin N at k0, so the sums over
    measured and unmeasured pixels of E_F and FcFo are exact;
  * each check also computes the value a plausible mistake would give
    (FcFo inverted as the CUDA engine has it, the unmeasured |F|^2 dropped
    from the E_F denominator, the outside and inside sums of E_R swapped)
    and asserts that the oracle is far from it.
"""
import numpy as np
import pytest

MEASURED, SUPPORT = 2, 1
CENTRO, REAL_OBJ, POS_REAL = 16, 1, 2


def _phaser(o, shape, amp, meas, sup, g1):
    p = o.phaser(shape)
    p.set_inputs(amp, meas, sup, g1)
    return p


@pytest.mark.parametrize("prec,tol", [("f32", 2e-6), ("f64", 1e-13)])
def test_efourier_delta_constant_transform(o32, o64, prec, tol):
    """g1 = c delta(0): F = c everywhere.  Measured pixels with amplitudes a_i,
    unmeasured pixels (beamstop) contribute |c|^2 to the denominator."""
    o = o32 if prec == "f32" else o64
    shape = (16, 16)
    c = 3.0 - 4.0j  # |c| = 5
    g1 = np.zeros(shape, np.complex128)
    g1[0, 0] = c
    amp = np.zeros(shape)
    amp.flat[:] = 1.0 + (np.arange(256) % 7)  # values 1..7
    meas = np.ones(shape, bool)
    meas[0, :5] = False  # 5 unmeasured pixels
    amp[~meas] = 0
    p = _phaser(o, shape, amp, meas, np.zeros(shape, bool), g1)
    ef, fcfo = p.efourier()

    a = amp[meas]
    nom = np.sum((5.0 - a) ** 2)
    den = np.sum(a ** 2) + 5 * 25.0
    want_ef = np.sqrt(nom / den)
    want_fcfo = a.sum() / (a.size * 5.0)
    assert abs(ef - want_ef) <= tol * want_ef
    assert abs(fcfo - want_fcfo) <= tol * want_fcfo

    # planted mistakes the closed form must reject
    wrong_den = np.sqrt(nom / np.sum(a ** 2))  # unmeasured |F|^2 dropped
    assert abs(ef - wrong_den) > 1e-3 * want_ef
    assert abs(fcfo - 1.0 / want_fcfo) > 1e-2  # Sum|F| / Sum amp (the CUDA engine's ratio)


@pytest.mark.parametrize("k0_measured", [True, False])
def test_efourier_plane_wave_single_bin(o64, k0_measured):
    """g1 = exp(2 pi i (ky y / ny + kx x / nx)): F = N at (ky, kx), 0 elsewhere."""
    shape = (8, 12)
    ny, nx = shape
    N = nx * ny
    ky, kx = 3, 5
    y, x = np.mgrid[0:ny, 0:nx]
    g1 = np.exp(2j * np.pi * (ky * y / ny + kx * x / nx))
    amp = np.full(shape, 2.0)
    meas = np.ones(shape, bool)
    meas[0, 0] = False
    meas[ky, kx] = k0_measured
    amp[~meas] = 0
    p = _phaser(o64, shape, amp, meas, np.zeros(shape, bool), g1)
    ef, fcfo = p.efourier()

    nmeas = int(meas.sum())
    if k0_measured:
        # the bin: (N - 2)^2; every other measured pixel: (0 - 2)^2
        nom = (N - 2.0) ** 2 + (nmeas - 1) * 4.0
        den = nmeas * 4.0
        want_fcfo = 2.0 * nmeas / N
    else:
        nom = nmeas * 4.0
        den = nmeas * 4.0 + N * N
    want_ef = np.sqrt(nom / den)
    assert ef == pytest.approx(want_ef, rel=1e-12)
    if k0_measured:
        assert fcfo == pytest.approx(want_fcfo, rel=1e-12)
    else:
        # no measured pixel carries any |F|: Sum|F| over them is 0 up to rounding
        assert not np.isfinite(fcfo) or fcfo > 1e10
        # dropping the unmeasured |F|^2 from the denominator would give 1
        assert abs(ef - 1.0) > 0.5


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_ereal_exact_split(o32, o64, prec):
    """E_R = sqrt(sum_outside |gp|^2 / sum_all |gp|^2) on a hand-built gp."""
    o = o32 if prec == "f32" else o64
    shape = (8, 8)
    p = _phaser(o, shape, np.zeros(shape), np.ones(shape, bool), np.zeros(shape, bool), np.zeros(shape))
    gp = np.zeros(shape, np.complex128)
    sup = np.zeros(shape, bool)
    sup[:4, :4] = True
    gp[:4, :4] = 3 + 4j            # 16 inside at |.|^2 = 25: 400
    gp[6, :] = 1 + 0j              # 8 outside at 1: 8
    gp[7, 7] = 0 + 2j              # 1 outside at 4: 4
    p.gp[...] = gp
    p.flags[...] = (sup * SUPPORT) | MEASURED
    want = np.sqrt(12.0 / 412.0)
    assert p.ereal() == pytest.approx(want, rel=1e-6 if prec == "f32" else 1e-14)
    # inside and outside swapped would give sqrt(400/412)
    assert abs(p.ereal() - np.sqrt(400.0 / 412.0)) > 0.5


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("count", [0, 1, 37, 4096])
def test_support_fraction_exact_count(o32, o64, prec, count):
    o = o32 if prec == "f32" else o64
    shape = (64, 64)
    rng = np.random.default_rng(count)
    sup = np.zeros(4096, bool)
    sup[rng.permutation(4096)[:count]] = True
    p = _phaser(o, shape, np.zeros(shape), np.ones(shape, bool), sup.reshape(shape), np.zeros(shape))
    # count / 4096 is exact in binary floating point
    assert p.support_fraction() == count / 4096.0


def test_support_fraction_ignores_measured_bit(o64):
    shape = (4, 4)
    p = _phaser(o64, shape, np.zeros(shape), np.ones(shape, bool), np.eye(4, dtype=bool), np.zeros(shape))
    assert p.support_fraction() == 0.25
    p.flags[...] = MEASURED  # measured everywhere, no support
    assert p.support_fraction() == 0.0


@pytest.mark.parametrize("constraints", [CENTRO, CENTRO | POS_REAL, REAL_OBJ, 0])
def test_fourier_constraints_centrosymmetric(o32, constraints):
    """Centrosymmetric objects: im = 0 on every pixel, re untouched
    (src/phasing.c:1063-1067); no other constraint bit acts in Fourier space."""
    import ctypes as C
    rng = np.random.default_rng(5)
    a = (rng.standard_normal(300) + 1j * rng.standard_normal(300)).astype(np.complex64)
    before = a.copy()
    o32.lib.or_apply_fourier_constraints(a.ctypes.data, C.c_long(a.size), constraints)
    assert np.array_equal(a.real, before.real)
    if constraints & CENTRO:
        assert np.all(a.imag == 0)
    else:
        assert np.array_equal(a, before)


def test_phased_projection_replaces_measured_only(o32):
    """SpRecoverAmplitudes objective: G = phased_amp where measured, the model
    elsewhere (src/phasing.c:1103-1109)."""
    import ctypes as C
    rng = np.random.default_rng(6)
    n = 500
    G = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    ph = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    flags = rng.integers(0, 4, n).astype(np.int32)  # all four support/measured combinations
    before = G.copy()
    want = np.where(flags & MEASURED, ph, G)
    o32.lib.or_phased_projection.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_long]
    o32.lib.or_phased_projection(G.ctypes.data, ph.ctypes.data, flags.ctypes.data, n)
    assert np.array_equal(G, want)
    # a projection keyed on the support bit instead would give another array
    assert not np.array_equal(np.where(flags & SUPPORT, ph, before), want)


def test_phased_projection_inside_iteration(o64):
    """One ER iteration with the phased-amplitude objective and a full support
    and no beamstop returns IFFT(phased) exactly: the projection replaced every
    Fourier value."""
    shape = (8, 8)
    rng = np.random.default_rng(7)
    ph = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    g = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    p = _phaser(o64, shape, np.ones(shape), np.ones(shape, bool), np.ones(shape, bool), g)
    p.phased_amp[...] = ph
    p.objective = 1
    assert p.iterate(o64.alg(4, 0), 1) == 0
    assert np.allclose(p.g1, np.fft.ifftn(ph), rtol=0, atol=1e-13)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_model_change_is_g1_minus_g0(o32, o64, prec):
    import ctypes as C
    o = o32 if prec == "f32" else o64
    shape = (6, 10)
    rng = np.random.default_rng(8)
    p = _phaser(o, shape, np.zeros(shape), np.ones(shape, bool), np.zeros(shape, bool), np.zeros(shape))
    g0 = (rng.integers(-8, 8, shape) + 1j * rng.integers(-8, 8, shape))
    g1 = (rng.integers(-8, 8, shape) + 1j * rng.integers(-8, 8, shape))
    p.g0[...] = g0
    p.g1[...] = g1
    out = np.empty(shape, o.np_cplx)
    o.lib.or_model_change(C.c_void_p(p.h), out.ctypes.data)
    assert np.array_equal(out, (g1 - g0).astype(o.np_cplx))  # small integers: exact
    assert not np.array_equal(out, (g0 - g1).astype(o.np_cplx))
