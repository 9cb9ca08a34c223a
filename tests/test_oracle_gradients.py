"""The oracle's XC nuclear gradients (xqo_gradients) pinned three ways, on the CPU:

- against the compiled reference (DFTIntegrator_Integrate with gradients3, built from the reference's own sources by
  `make -C oracle ref`): water with BLYP and PBE, spherical and Cartesian, and a water dimer;
- against central finite differences of the oracle's E_xc at fixed P (the grid moves with its atoms, as the weight terms
  assume), which a dropped term, a wrong sign or a swapped index fails;
- translational invariance: the gradients of all atoms sum to zero.
"""

import numpy as np
import pytest

from helpers import cluster_system, ids, random_density, water_system


def _oracle():
    import oracle
    return oracle.oracle()


@pytest.mark.parametrize("functional", ["blyp", "pbe"])
@pytest.mark.parametrize("spherical", [True, False])
def test_water_gradients_match_reference(functional, spherical):
    import oracle
    if not oracle.reference_available():
        pytest.skip("compiled reference not built (make -C oracle ref)")
    Z, xyz, spec, n = water_system(spherical)
    P = random_density(n, 5, seed=11)
    g0 = oracle.reference().setup(2, Z, xyz, spec, spherical, ids(functional)).gradients(P, 1)
    g = _oracle().setup(2, Z, xyz, spec, spherical, ids(functional)).gradients(P, 4)
    assert np.max(np.abs(g - g0)) <= 1e-11
    assert np.max(np.abs(g0)) > 1e-2          # a real, non-vanishing gradient


def test_dimer_gradients_match_reference():
    import oracle
    if not oracle.reference_available():
        pytest.skip("compiled reference not built (make -C oracle ref)")
    Z, xyz, spec, n = cluster_system(2, 4.5, 3, False)
    P = random_density(n, 10, seed=12)
    g0 = oracle.reference().setup(2, Z, xyz, spec, False, ids("pbe")).gradients(P, 1)
    g = _oracle().setup(2, Z, xyz, spec, False, ids("pbe")).gradients(P, 8)
    assert np.max(np.abs(g - g0)) <= 1e-11


def test_gradients_are_finite_differences_of_exc():
    Z, xyz, spec, n = water_system(False)
    P = random_density(n, 5, seed=13)
    lib = _oracle()
    g = lib.setup(2, Z, xyz, spec, False, ids("blyp")).gradients(P, 8)
    h = 1.0e-4
    for atom in range(len(Z)):
        for c in range(3):
            e = []
            for sgn in (1.0, -1.0):
                x = xyz.copy()
                x[atom, c] += sgn * h
                e.append(lib.setup(2, Z, x, spec, False, ids("blyp")).integrate(P, 8)[0])
            fd = (e[0] - e[1]) / (2.0 * h)
            # Central differences with h = 1e-4 bohr agree to about 2e-9 here; a dropped weight term is 1e-3 or more.
            assert abs(fd - g[atom, c]) <= 2e-8 * max(1.0, abs(g[atom, c])), (atom, c, fd, g[atom, c])


def test_gradients_sum_to_zero():
    Z, xyz, spec, n = cluster_system(3, 5.5, 9, True)
    P = random_density(n, 12, seed=14)
    g = _oracle().setup(2, Z, xyz, spec, True, ids("pbe")).gradients(P, 8)
    assert np.max(np.abs(g.sum(axis=0))) <= 1e-11
    assert np.max(np.abs(g)) > 1e-3
