"""Host-side workload generators (no GPU)."""
import numpy as np

from libspimage_amd import workloads


def test_smooth_sizes_are_exactly_the_5_smooth_lengths():
    got = workloads.smooth_sizes(128, 1024)

    def smooth(n):
        for p in (2, 3, 5):
            while n % p == 0:
                n //= p
        return n == 1

    assert got == [n for n in range(128, 1025) if smooth(n)]


def test_serving_trace_is_seeded_and_shaped():
    a = workloads.serving_trace(200, 2.0, seed=7)
    b = workloads.serving_trace(200, 2.0, seed=7)
    assert a == b
    assert 300 < len(a) < 500  # Poisson(400)
    t = np.array([x[0] for x in a])
    assert np.all(np.diff(t) > 0) and t[-1] <= 2.0
    frac3 = np.mean([len(x[1]) == 3 for x in a])
    assert 0.08 < frac3 < 0.22
    assert {x[2] for x in a} == {"er", "hio", "raar"}
    its = np.array([x[3] for x in a])
    assert its.min() >= 100 and its.max() <= 1000 and np.all(its % 50 == 0)
    assert len({x[4] for x in a}) == len(a)
