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


def test_reconstruction_converges(sp):
    """A phasing run through the front-end concentrates the model in the
    support: the real-space error falls (at iteration 0 the random-phase
    start has exact moduli, so the Fourier error starts at 0 instead)."""
    R = _configure((64, 64), "area")
    out = R.reconstruct(seed=3)
    assert out["fourier_error"][0] < 1e-5
    assert out["real_error"][-1] < 0.5 * out["real_error"][1]
    assert 0.05 < out["support_size"][-1] < 0.15
