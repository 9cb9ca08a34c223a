"""Unrestricted (UKS) quadrature in the oracle (xqo_integrate_uks), on the CPU.

- against the compiled reference's DFTIntegrator_Integrate with alpha and beta densities and Fock matrices
  (DFTIntegrator.c:281-402): water, spherical and Cartesian, BLYP, PBE, LDA and B3LYP, and a water dimer;
- the closed-shell limit: P_a = P_b = P/2 gives the restricted E_xc, N_e and V_a = V_b = V;
- dE_xc/dP_a,mn = (2 - delta_mn) V_a,mn by central differences (a dropped cross term v_sigma_ab fails it).
"""

import numpy as np
import pytest

from helpers import cluster_system, random_density, water_system
from pdynamo3_amd.definitions import functional_ids


def _oracle():
    import oracle
    return oracle.oracle()


def _pair(n, seed):
    return random_density(n, 5, seed=seed, scale=0.2), random_density(n, 4, seed=seed + 1, scale=0.2)


@pytest.mark.parametrize("name", ["blyp", "pbe", "lda", "b3lyp"])
@pytest.mark.parametrize("spherical", [True, False])
def test_water_uks_matches_reference(name, spherical):
    import oracle
    if not oracle.reference_available():
        pytest.skip("compiled reference not built (make -C oracle ref)")
    ids = list(functional_ids(name))
    Z, xyz, spec, n = water_system(spherical)
    Pa, Pb = _pair(n, 11)
    e0, n0, Va0, Vb0, _ = oracle.reference().setup(2, Z, xyz, spec, spherical, ids).integrate_uks(Pa, Pb, 1)
    e, ne, Va, Vb, _ = _oracle().setup(2, Z, xyz, spec, spherical, ids).integrate_uks(Pa, Pb, 4)
    assert abs(e - e0) <= 1e-11 and abs(ne - n0) <= 1e-11
    assert np.max(np.abs(Va - Va0)) <= 1e-12 and np.max(np.abs(Vb - Vb0)) <= 1e-12


def test_dimer_uks_matches_reference():
    import oracle
    if not oracle.reference_available():
        pytest.skip("compiled reference not built (make -C oracle ref)")
    Z, xyz, spec, n = cluster_system(2, 4.5, 3, False)
    Pa, Pb = _pair(n, 31)
    ids = list(functional_ids("pbe"))
    e0, n0, Va0, Vb0, _ = oracle.reference().setup(2, Z, xyz, spec, False, ids).integrate_uks(Pa, Pb, 1)
    e, ne, Va, Vb, _ = _oracle().setup(2, Z, xyz, spec, False, ids).integrate_uks(Pa, Pb, 8)
    assert abs(e - e0) <= 1e-11 and abs(ne - n0) <= 1e-11
    assert np.max(np.abs(Va - Va0)) <= 1e-12 and np.max(np.abs(Vb - Vb0)) <= 1e-12


@pytest.mark.parametrize("name", ["blyp", "pbe", "lda", "b3lyp"])
def test_closed_shell_limit(name):
    ids = list(functional_ids(name))
    Z, xyz, spec, n = water_system(True)
    P = random_density(n, 5, seed=13)
    h = _oracle().setup(2, Z, xyz, spec, True, ids)
    e, ne, V, _ = h.integrate(P, 4)
    eu, neu, Va, Vb, _ = h.integrate_uks(P / 2, P / 2, 4)
    assert abs(eu - e) <= 1e-11 * max(1.0, abs(e)) and abs(neu - ne) <= 1e-11
    assert np.max(np.abs(Va - V)) <= 1e-11 and np.max(np.abs(Vb - V)) <= 1e-11


def test_alpha_potential_is_derivative():
    Z, xyz, spec, n = water_system(False)
    Pa, Pb = _pair(n, 41)
    h = _oracle().setup(2, Z, xyz, spec, False, list(functional_ids("pbe")))
    _, _, Va, _, _ = h.integrate_uks(Pa, Pb, 8)
    rng = np.random.default_rng(5)
    iu = np.tril_indices(n)
    for k in rng.choice(len(Pa), 6, replace=False):
        d = 1e-5
        Pp, Pm = Pa.copy(), Pa.copy()
        Pp[k] += d
        Pm[k] -= d
        fd = (h.integrate_uks(Pp, Pb, 8)[0] - h.integrate_uks(Pm, Pb, 8)[0]) / (2 * d)
        factor = 1.0 if iu[0][k] == iu[1][k] else 2.0
        assert abs(fd - factor * Va[k]) <= 1e-7 * max(1.0, abs(fd)), (k, fd, Va[k])
