"""Unrestricted (UKS) quadrature in the oracle (xqo_integrate_uks), on the CPU.

- against the compiled reference's DFTIntegrator_Integrate with alpha and beta densities and Fock matrices
  (DFTIntegrator.c:281-402): water, spherical and Cartesian, BLYP, PBE, LDA and B3LYP, and a water dimer;
- the closed-shell limit: P_a = P_b = P/2 gives the restricted E_xc, N_e and V_a = V_b = V;
- dE_xc/dP_a,mn = (2 - delta_mn) V_a,mn by central differences (a dropped cross term v_sigma_ab fails it);
- the unrestricted XC nuclear gradients: against the compiled reference, the closed-shell limit (the restricted
  gradients), central differences of the unrestricted E_xc in the atom positions, translational invariance.
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


# Unrestricted XC nuclear gradients (xqo_gradients_uks, DFTIntegrator.c:405-487).

@pytest.mark.parametrize("name", ["blyp", "pbe", "lda", "b3lyp"])
@pytest.mark.parametrize("spherical", [True, False])
def test_water_uks_gradients_match_reference(name, spherical):
    import oracle
    if not oracle.reference_available():
        pytest.skip("compiled reference not built (make -C oracle ref)")
    ids = list(functional_ids(name))
    Z, xyz, spec, n = water_system(spherical)
    Pa, Pb = _pair(n, 11)
    g0 = oracle.reference().setup(2, Z, xyz, spec, spherical, ids).gradients_uks(Pa, Pb, 1)
    g = _oracle().setup(2, Z, xyz, spec, spherical, ids).gradients_uks(Pa, Pb, 4)
    assert np.max(np.abs(g - g0)) <= 1e-11
    assert np.max(np.abs(g0)) > 1e-2


def test_dimer_uks_gradients_match_reference():
    import oracle
    if not oracle.reference_available():
        pytest.skip("compiled reference not built (make -C oracle ref)")
    Z, xyz, spec, n = cluster_system(2, 4.5, 3, False)
    Pa, Pb = _pair(n, 31)
    ids = list(functional_ids("pbe"))
    g0 = oracle.reference().setup(2, Z, xyz, spec, False, ids).gradients_uks(Pa, Pb, 1)
    g = _oracle().setup(2, Z, xyz, spec, False, ids).gradients_uks(Pa, Pb, 8)
    assert np.max(np.abs(g - g0)) <= 1e-11


@pytest.mark.parametrize("name", ["blyp", "lda"])
def test_uks_gradients_closed_shell_limit(name):
    Z, xyz, spec, n = water_system(True)
    P = random_density(n, 5, seed=13)
    h = _oracle().setup(2, Z, xyz, spec, True, list(functional_ids(name)))
    g = h.gradients(P, 4)
    gu = h.gradients_uks(P / 2, P / 2, 4)
    assert np.max(np.abs(gu - g)) <= 1e-10 * max(1.0, np.max(np.abs(g)))


def test_uks_gradients_are_finite_differences_of_exc():
    # A dropped cross term (v_sigma_ab with the other spin's gradient) or a weight term built from one spin's density
    # moves the result by 1e-3 or more; central differences with h = 1e-4 bohr agree to a few 1e-9 here.
    Z, xyz, spec, n = water_system(False)
    Pa, Pb = _pair(n, 51)
    ids = list(functional_ids("pbe"))
    lib = _oracle()
    g = lib.setup(2, Z, xyz, spec, False, ids).gradients_uks(Pa, Pb, 8)
    h = 1.0e-4
    for atom in range(len(Z)):
        for c in range(3):
            e = []
            for sgn in (1.0, -1.0):
                x = xyz.copy()
                x[atom, c] += sgn * h
                e.append(lib.setup(2, Z, x, spec, False, ids).integrate_uks(Pa, Pb, 8)[0])
            fd = (e[0] - e[1]) / (2.0 * h)
            assert abs(fd - g[atom, c]) <= 2e-8 * max(1.0, abs(g[atom, c])), (atom, c, fd, g[atom, c])


def test_uks_gradients_sum_to_zero():
    Z, xyz, spec, n = cluster_system(3, 5.5, 9, True)
    Pa, Pb = random_density(n, 12, seed=14, scale=0.2), random_density(n, 10, seed=15, scale=0.2)
    g = _oracle().setup(2, Z, xyz, spec, True, list(functional_ids("b3lyp"))).gradients_uks(Pa, Pb, 8)
    assert np.max(np.abs(g.sum(axis=0))) <= 1e-11
    assert np.max(np.abs(g)) > 1e-3


def test_fully_polarised_hydrogen_matches_reference():
    """One hydrogen atom with all density in the alpha spin (zeta = 1, the functionals' threshold branch)."""
    import oracle
    if not oracle.reference_available():
        pytest.skip("compiled reference not built (make -C oracle ref)")
    from pdynamo3_amd import basis
    Z = np.array([1], dtype=np.int32)
    xyz = np.zeros((1, 3))
    spec = basis.element_spec(basis.load_basis(), Z)
    Pa = np.array([0.30, 0.25, 0.45])    # packed 2 x 2 (two functions on H), positive definite
    Pb = np.zeros(3)
    for name in ("blyp", "pbe", "lda", "b3lyp"):
        ids = list(functional_ids(name))
        e0, n0, Va0, Vb0, _ = oracle.reference().setup(2, Z, xyz, spec, True, ids).integrate_uks(Pa, Pb, 1)
        h = _oracle().setup(2, Z, xyz, spec, True, ids)
        e, ne, Va, Vb, _ = h.integrate_uks(Pa, Pb, 2)
        assert abs(e - e0) <= 1e-12 and abs(ne - n0) <= 1e-12, name
        assert np.max(np.abs(Va - Va0)) <= 1e-12 and np.max(np.abs(Vb - Vb0)) <= 1e-12, name
        assert np.max(np.abs(h.gradients_uks(Pa, Pb, 2))) == 0.0
