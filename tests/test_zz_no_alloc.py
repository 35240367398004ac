"""csrst_sf_current_density_noalloc (static per-GPU scratch, no allocation)
against csrst_sf_current_density through the C ABI: two calls in a row, the
second with fewer particles (the scratch keeps the first call's contents),
bit for bit, and the size limit reported as an error."""
import numpy as np
import pytest

from csrst_amd import shapefn as SF
from csrst_amd._lib import CsrstError

pytestmark = pytest.mark.gpu


def trajectories(n, seed, scale=(0.7, 1.3)):
    rng = np.random.default_rng(seed)
    oldR = rng.uniform(8, 20, (n, 2)) * np.array(scale)
    R = oldR + rng.uniform(-0.99, 0.99, (n, 2)) * np.array(scale)
    return oldR, R, rng.uniform(-1, 1, n)


@pytest.mark.parametrize("order", [0, 1, 2])
def test_noalloc_matches_allocating_entry(gpu, order):
    dx, dt = (0.7, 1.3), 0.37
    box, org = (-6, -6, 40, 40), (0.0, 0.0)
    for n, seed in ((12000, 3), (2500, 4)):
        oldR, R, q = trajectories(n, seed + order)
        for comp in (0, 1):
            a = SF.getCurrentDensity(order, comp, box, org, dx, dt, oldR, R, q)
            b = SF.getCurrentDensity(order, comp, box, org, dx, dt, oldR, R, q, noalloc=True)
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (order, n, comp)
            assert np.abs(a).max() > 0


def test_noalloc_size_limit(gpu):
    oldR, R, q = trajectories(200000, 9)
    with pytest.raises(CsrstError) as e:
        SF.getCurrentDensity(2, 0, (-6, -6, 40, 40), (0.0, 0.0), (0.7, 1.3), 0.37, oldR, R, q, noalloc=True)
    assert e.value.code == -1 and "scratch" in str(e.value)
