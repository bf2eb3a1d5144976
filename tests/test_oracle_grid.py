"""The checker's photon bucket grid gives exactly the brute-force gather:
same photons found, same sums in the same order (bit-identical estimates),
for range and k-nearest gathers, both estimates, radii from tiny to larger
than the map, and queries outside the map."""
import numpy as np
import pytest

import oracle
from scenes import build


def _photons(rng, n):
    ph = np.zeros((n, 12), np.float32)
    ph[:, 0:3] = rng.uniform(-1, 1, (n, 3)) * np.array([3.0, 0.5, 2.0])
    d = rng.normal(size=(n, 3))
    ph[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    nn = rng.normal(size=(n, 3))
    ph[:, 6:9] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    ph[:, 9:12] = rng.uniform(0, 1, (n, 3))
    return ph


@pytest.mark.parametrize("prec", [32, 64])
def test_grid_gather_equals_brute_force(prec):
    rng = np.random.default_rng(prec)
    b = build()
    grid = oracle.OracleScene(b.ptr(), prec)
    brute = oracle.OracleScene(b.ptr(), prec, oracle.BRUTE_GATHER)
    ph = _photons(rng, 8_000)
    # photons on cell borders too: coordinates that are multiples of the radius
    ph[:800, 0:3] = np.round(ph[:800, 0:3] / 0.05) * 0.05
    for kind in (0, 1):
        grid.set_photons(kind, ph)
        brute.set_photons(kind, ph)
    q = np.zeros((1000, 10), np.float32)
    q[:, 0:3] = rng.uniform(-1.2, 1.2, (1000, 3)) * np.array([3.0, 0.5, 2.0])
    q[:50, 0:3] = ph[:50, 0:3]                        # on photons
    q[50:100, 0:3] = np.array([40.0, -30.0, 7.0])     # far outside the map
    qn = rng.normal(size=(1000, 3))
    q[:, 3:6] = qn / np.linalg.norm(qn, axis=1, keepdims=True)
    for radius in (1e-3, 0.05, 0.2, 10.0):
        for k in ((0, 8, 200) if radius < 1 else (0, 200)):
            for mode in (0, 1):
                for kind in (0, 1):
                    a = grid.gather(kind, q, radius, k, mode)
                    b_ = brute.gather(kind, q, radius, k, mode)
                    assert np.array_equal(a.view(np.uint32), b_.view(np.uint32)), (radius, k, mode, kind)
    # the test has teeth: the estimates are not all zero
    assert (grid.gather(1, q, 0.2, 0, 1) != 0).any()
