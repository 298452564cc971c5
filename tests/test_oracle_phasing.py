"""Pins the oracle's projections, constraints and iterations (oracle/spimage_oracle.c).

Reference semantics restated by the oracle: src/phasing.c:1026-1274 and the
scheduler src/phasing.c:895-1024.  The checks come from the mathematics and
from what the reference's own tests assert, not from the oracle's code:
  * modulus projection: |out| = amp on measured pixels, phase kept,
    unmeasured pixels untouched, zero modulus -> (amp, 0), margins clamp
    (reference tests/ProjTests.c:4-31 assert magnitude and phase);
  * every algorithm leaves a consistent solution (support- and
    Fourier-consistent) unchanged (fixed point);
  * ER zeroes the outside of the support; HIO and ER agree inside it; HIO is
    linear in beta outside it; RAAR with beta = 0 is the plain projection;
  * a float64 textbook implementation through numpy's FFT (Fienup's HIO,
    Luke's RAAR, Elser's difference map) agrees to 1e-12;
  * the reference's convergence test (tests/PhasingTests.c:10-59): a 4x4
    random object in an 8x8 frame converges under HIO;
  * Elser's gamma values at beta = 0.8 (src/phasing.c:21-43);
  * the support scheduler enumerated by hand from src/phasing.c:1002-1016.
"""
import numpy as np
import pytest

from oracle.oracle import DIFFMAP, ER, HIO, RAAR, SUP_STATIC


def make_case(shape, seed=0, beamstop=True):
    rng = np.random.default_rng(seed)
    obj = np.zeros(shape)
    sl = tuple(slice(0, max(2, s // 4)) for s in shape)
    obj[sl] = rng.random(obj[sl].shape)
    amp = np.abs(np.fft.fftn(obj))
    meas = np.ones(shape, bool)
    if beamstop:
        meas.flat[:3] = False
    sup = np.zeros(shape, bool)
    sup[tuple(slice(0, max(2, s // 4) + 2) for s in shape)] = True
    g0 = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * 0.1
    return obj, amp, meas, sup, g0


def P_M(x, amp, meas):
    G = np.fft.fftn(x)
    with np.errstate(all="ignore"):
        m = amp / np.abs(G)
    out = np.where(np.isfinite(m), G * m, amp + 0j)
    return np.where(meas, out, G)


def textbook(alg, g0, amp, meas, sup, beta, gamma1=0.0, gamma2=0.0):
    N = g0.size
    if alg == DIFFMAP:
        f1 = np.where(sup, g0, -gamma1 * g0)
        Pf1 = np.fft.ifftn(P_M(f1, amp, meas))
        Prho = np.fft.ifftn(P_M(g0, amp, meas))
        g1 = np.where(sup, g0 + beta * ((1 + gamma2) * Prho - gamma2 * g0), g0)
        return g1 - beta * Pf1
    gp = np.fft.ifftn(P_M(g0, amp, meas))
    if alg == ER:
        return np.where(sup, gp, 0)
    if alg == HIO:
        return np.where(sup, gp, g0 - beta * gp)
    return np.where(sup, gp, (1 - 2 * beta) * gp + beta * g0)


def run_one(o, alg_id, g0, amp, meas, sup, beta=0.9, constraints=0, gammas=(0.0, 0.0), n=1):
    ph = o.phaser(g0.shape)
    ph.set_inputs(amp, meas, sup, g0)
    a = o.alg(alg_id, constraints, beta=[(0, beta)], gamma1=gammas[0], gamma2=gammas[1])
    assert ph.iterate(a, n) == 0
    return ph


@pytest.mark.parametrize("alg", [ER, HIO, RAAR, DIFFMAP])
@pytest.mark.parametrize("shape", [(16, 16), (12, 10), (6, 8, 10)])
def test_one_iteration_matches_textbook(o64, alg, shape):
    obj, amp, meas, sup, g0 = make_case(shape, seed=len(shape) + alg)
    gam = o64.dm_gammas(0.8) if alg == DIFFMAP else (0.0, 0.0)
    ph = run_one(o64, alg, g0, amp, meas, sup, beta=0.8, gammas=gam)
    ref = textbook(alg, g0, amp, meas, sup, 0.8, *gam)
    assert np.linalg.norm(ph.g1 - ref) / np.linalg.norm(ref) < 1e-12


@pytest.mark.parametrize("alg", [ER, HIO, RAAR, DIFFMAP])
def test_solution_is_fixed_point(o64, alg):
    obj, amp, meas, sup, _ = make_case((16, 16), seed=5, beamstop=False)
    gam = o64.dm_gammas(0.9)
    ph = run_one(o64, alg, obj.astype(complex), amp, meas, sup, beta=0.9, gammas=gam)
    assert np.max(np.abs(ph.g1 - obj)) < 1e-12


def test_er_hio_raar_relations(o64):
    obj, amp, meas, sup, g0 = make_case((16, 16), seed=9)
    er = run_one(o64, ER, g0, amp, meas, sup).g1.copy()
    assert np.all(er[~sup] == 0)
    h1 = run_one(o64, HIO, g0, amp, meas, sup, beta=0.3).g1.copy()
    h2 = run_one(o64, HIO, g0, amp, meas, sup, beta=0.7).g1.copy()
    assert np.allclose(h1[sup], er[sup], atol=1e-14)
    gp = run_one(o64, RAAR, g0, amp, meas, sup, beta=0.0).g1.copy()  # RAAR(0) = P_M everywhere
    assert np.allclose(gp[sup], er[sup], atol=1e-14)
    # outside: HIO(b) = g0 - b gp, so HIO(0.3) - HIO(0.7) = 0.4 gp
    assert np.allclose((h1 - h2)[~sup], 0.4 * gp[~sup], atol=1e-13)


def test_modulus_projection_closed_form(o64):
    rng = np.random.default_rng(2)
    shape = (8, 8)
    G = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    G[0, 1] = 0  # zero modulus on a measured pixel -> (amp, 0)
    amp = rng.random(shape) + 0.5
    meas = np.ones(shape, bool)
    meas[3, 3] = False
    flags = (meas * 2).astype(np.int32)
    a = G.astype(np.complex128).copy()
    z = np.zeros(shape)
    rc = o64.lib.or_module_projection(a.ctypes.data, amp.ctypes.data, z.ctypes.data, z.ctypes.data,
                                      flags.ctypes.data, a.size, 0, 0)
    assert rc == 0
    m = meas.copy()
    m[0, 1] = False
    assert np.allclose(np.abs(a[m]), amp[m], rtol=1e-14)
    assert np.allclose(np.angle(a[m]), np.angle(G[m]), atol=1e-12)
    assert a[3, 3] == G[3, 3]
    assert a[0, 1] == amp[0, 1] + 0j
    # margins clamp |G| into [lo, hi]
    lo, hi = amp * 0.9, amp * 1.1
    b = G.copy()
    rc = o64.lib.or_module_projection(b.ctypes.data, amp.ctypes.data, lo.ctypes.data, hi.ctypes.data,
                                      flags.ctypes.data, b.size, 32, 1)
    expect = np.clip(np.abs(G), lo, hi)
    ok = meas & (np.abs(G) > 0)
    assert np.allclose(np.abs(b[ok]), expect[ok], rtol=1e-13)
    # margins requested without margin maps: the reference is fatal, the oracle reports it
    assert o64.lib.or_module_projection(b.ctypes.data, amp.ctypes.data, lo.ctypes.data, hi.ctypes.data,
                                        flags.ctypes.data, b.size, 32, 0) == -1


@pytest.mark.parametrize("c,expect", [
    (1, lambda g: g.real + 0j),
    (2, lambda g: np.maximum(g.real, 0) + 0j),
    (2 | 8, lambda g: np.abs(g.real) + 0j),
    (4, lambda g: np.maximum(g.real, 0) + 1j * np.maximum(g.imag, 0)),
    (4 | 8, lambda g: np.abs(g.real) + 1j * np.abs(g.imag)),
])
def test_support_constraints_closed_form(o64, c, expect):
    rng = np.random.default_rng(4)
    g = rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6))
    sup = rng.random((6, 6)) > 0.5
    flags = sup.astype(np.int32)
    a = g.copy()
    o64.lib.or_apply_constraints(a.ctypes.data, flags.ctypes.data, a.size, c)
    assert np.array_equal(a[sup], expect(g)[sup])
    assert np.array_equal(a[~sup], g[~sup])


def test_dm_gammas_elser(o64):
    # gamma1 = -(4 + (2+b) s + b s^2) / (b (4 - s + s^2)), gamma2 = (3-b)/(2b), s = 0.1
    g1, g2 = o64.dm_gammas(0.8)
    assert abs(g1 - (-(4 + 2.8 * 0.1 + 0.8 * 0.01) / (0.8 * (4 - 0.1 + 0.01)))) < 1e-6
    assert abs(g2 - 2.2 / 1.6) < 1e-12


def test_reference_convergence_case(o32):
    """tests/PhasingTests.c:10-59: 4x4 random object in an 8x8 frame converges under HIO."""
    solved = False
    for attempt in range(10):
        rng = np.random.default_rng(100 + attempt)
        obj = np.zeros((8, 8))
        obj[:4, :4] = rng.random((4, 4))
        amp = np.abs(np.fft.fftn(obj)).astype(np.float32)
        meas = np.ones((8, 8), bool)
        sup = obj != 0
        g = np.fft.ifftn(amp * np.exp(2j * np.pi * rng.random((8, 8)))).astype(np.complex64)
        ph = o32.phaser((8, 8))
        ph.set_inputs(amp, meas, sup, g)
        alg = o32.alg(HIO, 0, beta=[(0, 0.9)])
        for _ in range(300):
            prev = ph.g1.copy()
            ph.iterate(alg, 1)
            if np.sum(np.abs(ph.g1 - prev) ** 2) <= 1e-10:
                solved = True
                break
        if solved:
            break
    assert solved


def test_scheduler_matches_hand_enumeration(o64):
    """period 5: 12 iterations from 0 -> updates at 4 and 9; a later call of 4
    from 12 -> update at 14; calls of 4 then 1 from 20 -> update at 24."""
    obj, amp, meas, sup, g0 = make_case((8, 8), seed=1)
    ph = o64.phaser((8, 8))
    ph.set_inputs(amp, meas, sup, g0)
    a = o64.alg(ER)
    s = o64.sup(SUP_STATIC, period=5)
    ph.iterate(a, 12, s)
    assert ph.iteration == 12 and ph.support_log() == [4, 9]
    ph.iterate(a, 4, s)
    assert ph.iteration == 16 and ph.support_log() == [4, 9, 14]
    ph.iteration = 20
    ph.iterate(a, 4, s)
    assert ph.support_log() == [4, 9, 14]
    ph.iterate(a, 1, s)
    assert ph.iteration == 25 and ph.support_log() == [4, 9, 14, 24]
