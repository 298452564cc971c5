"""Reconstructor configuration checks (no device work)."""
import numpy as np
import pytest


def test_incomplete_configuration_returns_none():
    from libspimage_amd.reconstructor import Reconstructor
    R = Reconstructor()
    assert R.reconstruct() is None          # no intensities, support, algorithms, outputs
    R.set_intensities(np.ones((8, 8)))
    R.set_number_of_iterations(10)
    R.set_number_of_outputs(2)
    R.set_initial_support(radius=2)
    R.set_support_algorithm("area", update_period=5, blur_init=2, blur_final=1, area_init=0.2, area_final=0.1)
    assert R.reconstruct_loop(2) is None    # still no phasing algorithm
    R.append_phasing_algorithm("hio", number_of_iterations=4, beta_init=0.9, beta_final=0.9)
    assert not R._ready()                   # 4 phasing iterations against 10 in total


def test_schedule_matches_reference_loop():
    """The stops of the reference's loop (src/_spimage_reconstructor.py:578-611):
    output points and algorithm / support changes."""
    from libspimage_amd.reconstructor import Reconstructor
    R = Reconstructor()
    R.set_number_of_iterations(100)
    R.set_number_of_outputs_images(3)
    R.set_number_of_outputs_scores(5)
    stops = list(R._schedule([(40, "s0"), (60, "s1")], [(80, "a0"), (20, "a1")]))
    assert [s[0] for s in stops] == [0, 25, 40, 50, 75, 80, 100]
    assert [s for s in stops if s[1] is not None] == [(80, 1, None, False, False)]
    assert [s for s in stops if s[2] is not None] == [(40, None, 1, False, False)]
    assert stops[-1] == (100, None, None, True, True)


def test_cpu_engine_refused():
    from libspimage_amd.reconstructor import Reconstructor
    with pytest.raises(ValueError):
        Reconstructor(use_gpu=False)
