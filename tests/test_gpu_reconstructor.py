"""The Reconstructor front-end (libspimage_amd/reconstructor.py): the
reference's interface (src/_spimage_reconstructor.py), with reconstruct_loop
advancing all random starts together in one batch.  Reconstruction r of the
loop must equal reconstruct(seed=seed + r) bit for bit, images and scores at
every output point."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _configure(shape, kind):
    from libspimage_amd import workloads
    from libspimage_amd.reconstructor import Reconstructor
    p = workloads.make_pattern(shape, n_gauss=6, radius=tuple(s / 6 for s in shape), photons=1e6, beamstop=2, seed=77)
    R = Reconstructor()
    # the reference takes the centred (non-shifted) pattern
    R.set_intensities(np.fft.fftshift(p.amplitudes.astype(np.float64) ** 2))
    R.set_mask(np.fft.fftshift(p.mask.astype(bool)))
    R.set_number_of_iterations(60)
    R.set_number_of_outputs_images(4)
    R.set_number_of_outputs_scores(7)
    R.set_initial_support(radius=min(shape) / 5)
    if kind == "area":
        R.set_support_algorithm("area", update_period=10, blur_init=3.0, blur_final=1.0, area_init=0.12,
                                area_final=0.08)
        R.append_phasing_algorithm("raar", number_of_iterations=45, beta_init=0.9, beta_final=0.8)
        R.append_phasing_algorithm("er", number_of_iterations=15)
    else:
        R.append_support_algorithm("static", number_of_iterations=30)
        R.append_support_algorithm("threshold", number_of_iterations=30, update_period=10, blur_init=2.0,
                                   blur_final=1.0, threshold_init=0.2, threshold_final=0.15)
        R.set_phasing_algorithm("diffmap", beta_init=0.8, beta_final=0.8, gamma1=float("inf"), gamma2=float("inf"),
                                constraints=["enforce_real"])
    return R


@pytest.mark.parametrize("shape,kind", [((64, 64), "area"), ((128, 128), "area"), ((64, 64), "threshold"),
                                        ((16, 16, 16), "area")])
def test_loop_equals_single_reconstructions(sp, shape, kind):
    R = _configure(shape, kind)
    loop = R.reconstruct_loop(3, full_output=True, seed=500)
    assert loop["real_space_final"].shape == (3,) + shape
    for r in range(3):
        one = R.reconstruct(seed=500 + r)
        multi = loop["single_outputs"][r]
        for key in ("real_space", "support", "fourier_space", "mask"):
            assert np.array_equal(one[key], multi[key]), (r, key)
        for key in ("real_error", "fourier_error", "support_size"):
            # the real error at iteration 0 is 0/0 (no model before projection yet) in both
            assert np.allclose(one[key], multi[key], rtol=1e-6, atol=0, equal_nan=True), (r, key, one[key], multi[key])
        assert np.array_equal(loop["real_space_final"][r], one["real_space"][-1])
        assert np.array_equal(loop["support_final"][r], one["support"][-1])
        assert loop["scores_final"]["fourier_error"][r] == pytest.approx(one["fourier_error"][-1], rel=1e-6)
        assert list(one["iteration_index_scores"]) == [0, 10, 20, 30, 40, 50, 60]
    # distinct starts give distinct reconstructions
    assert not np.array_equal(loop["real_space_final"][0], loop["real_space_final"][1])


@pytest.mark.parametrize("shape", [(47, 64), (49, 64)])
def test_loop_on_odd_lengths(sp, shape):
    """An axis without a Stockham plan (47): the front-end runs through the
    Bluestein path and reconstruct_loop falls back to single reconstructions
    (batches do not take such shapes).  49 = 7^2 has a plan with radix-7
    stages and runs batched.  The schedule test below checks both paths
    against the oracle."""
    R = _configure(shape, "area")
    loop = R.reconstruct_loop(2, full_output=True, seed=900)
    assert loop["real_space_final"].shape == (2,) + shape
    for r in range(2):
        one = R.reconstruct(seed=900 + r)
        multi = loop["single_outputs"][r]
        for key in ("real_space", "support", "fourier_space", "mask"):
            assert np.array_equal(one[key], multi[key]), (r, key)
    assert not np.array_equal(loop["real_space_final"][0], loop["real_space_final"][1])


def test_reconstruction_converges(sp):
    """A phasing run through the front-end concentrates the model in the
    support: the real-space error falls (at iteration 0 the random-phase
    start has exact moduli, so the Fourier error starts at 0 instead)."""
    R = _configure((64, 64), "area")
    out = R.reconstruct(seed=3)
    assert out["fourier_error"][0] < 1e-5
    assert out["real_error"][-1] < 0.5 * out["real_error"][1]
    assert 0.05 < out["support_size"][-1] < 0.15


def _oracle_reconstruct(o, A, M, S, start, total, phasings, supports, out_img, out_sco):
    """The reference's reconstruct() loop (src/_spimage_reconstructor.py:
    553-675) driving the oracle's phaser: iterate to each stop point, then
    switch the phasing algorithm and the support array where their segments
    end, then record images and scores."""
    oph = o.phaser(A.shape)
    oph.set_inputs(A, M, S, start)
    i_alg = i_sup = a0 = s0 = done = 0
    imgs, sups, scores = [], [], []
    for it in range(total + 1):
        ca = it - a0 == phasings[i_alg][0] and it != total
        cs = it - s0 == supports[i_sup][0] and it != total
        if not (ca or cs or it in out_img or it in out_sco):
            continue
        if it:
            assert oph.iterate(phasings[i_alg][1], it - done, supports[i_sup][1]) == 0
            done = it
        if ca:
            i_alg, a0 = i_alg + 1, it
        if cs:
            i_sup, s0 = i_sup + 1, it
        if it in out_img:
            imgs.append(np.array(oph.gp if it else oph.g1))
            sups.append((oph.flags & 1).astype(bool))
        if it in out_sco:
            ef, _ = oph.efourier()
            scores.append((ef, oph.ereal(), oph.support_fraction()))
    return imgs, sups, np.array(scores)


@pytest.mark.parametrize("shape,center", [((64, 64), False), ((64, 64), True), ((47, 64), False), ((49, 64), False)])
def test_reconstruct_schedule_matches_oracle(sp, o32, shape, center):
    """reconstruct() against the oracle driven by the reference's loop, from
    the same start model: ER 20, RAAR (beta 0.9 -> 0.8) 10, ER 10; area
    shrinkwrap every 7 iterations (blur 3 -> 1 px, area 0.15 -> 0.10, with
    the centre-image update when `center`) for 28 iterations, then static.
    Algorithm and support switches, Bezier schedules and output points must
    line up: scores within 1e-3 relative, supports equal, images within
    relative L2 1e-3 at every output point.  47 x 64 runs the Bluestein
    path (unfused iterations), 49 x 64 the fused passes with radix-7
    stages."""
    from libspimage_amd import workloads
    from libspimage_amd.reconstructor import Reconstructor
    from oracle.oracle import SUP_AREA, SUP_CENTRE, SUP_STATIC
    p = workloads.make_pattern(shape, n_gauss=6, radius=10, photons=1e7, beamstop=2, seed=78)
    R = Reconstructor()
    R.set_intensities(np.fft.fftshift(p.amplitudes.astype(np.float64) ** 2))
    R.set_mask(np.fft.fftshift(p.mask.astype(bool)))
    total = 40
    R.set_number_of_iterations(total)
    R.set_number_of_outputs_images(3)
    R.set_number_of_outputs_scores(5)
    R.set_initial_support(radius=14)
    R.append_support_algorithm("area", number_of_iterations=28, update_period=7, blur_init=3.0, blur_final=1.0,
                               area_init=0.15, area_final=0.10, center_image=center)
    R.append_support_algorithm("static", number_of_iterations=12)
    R.append_phasing_algorithm("er", number_of_iterations=20)
    R.append_phasing_algorithm("raar", number_of_iterations=10, beta_init=0.9, beta_final=0.8)
    R.append_phasing_algorithm("er", number_of_iterations=10)
    out = R.reconstruct(seed=11)
    A, M, S, _, _ = R._prepare()
    start = np.fft.ifftshift(out["real_space"][0]).astype(np.complex64)
    area = o32.sup(SUP_AREA, period=7, blur=[(0, 3.0), (28, 1.0)], values=[(0, 0.15), (28, 0.10)])
    sup0 = [area, o32.sup(SUP_CENTRE, period=7)] if center else area
    supports = [(28, sup0), (12, o32.sup(SUP_STATIC, period=12))]
    phasings = [(20, o32.alg(4, 0)), (10, o32.alg(2, 0, beta=[(20, 0.9), (30, 0.8)])), (10, o32.alg(4, 0))]
    ii, isc = out["iteration_index_images"], out["iteration_index_scores"]
    imgs, sups, scores = _oracle_reconstruct(o32, A, M.astype(bool), S.astype(bool), start, total, phasings,
                                             supports, set(ii.tolist()), set(isc.tolist()))
    for k in range(len(ii)):
        gimg = np.fft.ifftshift(out["real_space"][k])
        gsup = np.fft.ifftshift(out["support"][k])
        assert np.array_equal(gsup, sups[k]), (k, int(ii[k]))
        e = float(np.linalg.norm(gimg - imgs[k]) / np.linalg.norm(imgs[k]))
        print(f"image at iteration {ii[k]}: relative L2 {e:.2e}")
        assert e <= 1e-3, (k, e)
    got = np.stack([out["fourier_error"], out["real_error"], out["support_size"]], 1)
    # E_F at iteration 0 is rounding noise (about 1e-7): the random-phase
    # start has the measured moduli exactly; hence the absolute floor
    ok = np.isclose(got, scores, rtol=1e-3, atol=1e-6, equal_nan=True)
    assert ok.all(), (got, scores)
