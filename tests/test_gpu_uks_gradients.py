"""Unrestricted XC nuclear gradients on the GPU (xq_gradients_uks, FP64) against the oracle's (itself equal to the
compiled reference's unrestricted DFTIntegrator_Integrate with gradients3 and pinned by finite differences of the
unrestricted E_xc, tests/test_oracle_uks.py) on the same inputs.

Tolerance 1e-9 hartree/bohr, as for the restricted gradients (different summation orders).  The gradient is FP64 on either
precision's handle.
"""

import numpy as np
import pytest

from helpers import cluster_system, random_density, water_system
from pdynamo3_amd.definitions import functional_ids

pytestmark = pytest.mark.gpu


def _grid(Z, xyz, spec, spherical, precision):
    from pdynamo3_amd import integrator
    g = integrator.QuadratureGrid(Z, xyz, 2)
    g.set_basis(spec, spherical, precision)
    return g


def _pair(n, seed):
    return random_density(n, 5, seed=seed, scale=0.2), random_density(n, 4, seed=seed + 1, scale=0.2)


@pytest.mark.parametrize("name", ["blyp", "pbe", "lda", "b3lyp"])
@pytest.mark.parametrize("spherical", [True, False])
def test_water_uks_gradients_match_oracle(oracle_lib, name, spherical):
    ids = list(functional_ids(name))
    Z, xyz, spec, n = water_system(spherical)
    Pa, Pb = _pair(n, 11)
    g0 = oracle_lib.setup(2, Z, xyz, spec, spherical, ids).gradients_uks(Pa, Pb, 4)
    g = _grid(Z, xyz, spec, spherical, "fp64")
    G = g.gradients_uks(ids, Pa, Pb)
    g.close()
    print(name, spherical, "max dG", np.max(np.abs(G - g0)), "max G", np.max(np.abs(g0)))
    assert np.max(np.abs(G - g0)) <= 1e-9
    assert np.max(np.abs(g0)) > 1e-2


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cluster4_uks_gradients_match_oracle(oracle_lib, precision):
    Z, xyz, spec, n = cluster_system(4, 6.0, 5, False)
    Pa, Pb = random_density(n, 20, seed=17, scale=0.2), random_density(n, 18, seed=19, scale=0.2)
    ids = list(functional_ids("pbe"))
    g0 = oracle_lib.setup(2, Z, xyz, spec, False, ids).gradients_uks(Pa, Pb, 8)
    g = _grid(Z, xyz, spec, False, precision)
    G = g.gradients_uks(ids, Pa, Pb)
    g.close()
    print("max dG", np.max(np.abs(G - g0)), "max G", np.max(np.abs(g0)))
    assert np.max(np.abs(G - g0)) <= 1e-9


@pytest.mark.parametrize("name", ["pbe", "b3lyp"])
def test_closed_shell_limit_matches_restricted_gradients(name):
    from pdynamo3_amd import densities, molecules
    Z, xyz, spec, n = cluster_system(4, 6.0, 5, False)
    _, _, rots, _ = molecules.water_cluster(4, 6.0, 5)
    P = densities.cluster_density(rots, spherical=False)
    ids = list(functional_ids(name))
    g = _grid(Z, xyz, spec, False, "mixed")
    G = g.gradients(ids, P)
    Gu = g.gradients_uks(ids, P / 2, P / 2)
    g.close()
    print(name, "max dG", np.max(np.abs(Gu - G)), "max G", np.max(np.abs(G)))
    assert np.max(np.abs(Gu - G)) <= 1e-9


def test_h2o16_uks_gradients_sum_to_zero():
    from pdynamo3_amd import densities, molecules
    Z, xyz, spec, n = cluster_system(16, 9.0, 1, False)
    _, _, rots, _ = molecules.water_cluster(16, 9.0, 1)
    P = densities.cluster_density(rots, spherical=False)
    g = _grid(Z, xyz, spec, False, "mixed")
    G = g.gradients_uks(list(functional_ids("pbe")), 0.55 * P, 0.45 * P)
    g.close()
    print("sum", G.sum(axis=0), "max", np.max(np.abs(G)))
    assert np.max(np.abs(G.sum(axis=0))) <= 1e-9
    assert np.max(np.abs(G)) > 1e-3


def test_uks_gradients_increment_and_zero_density():
    Z, xyz, spec, n = water_system(True)
    ids = list(functional_ids("blyp"))
    g = _grid(Z, xyz, spec, True, "fp64")
    zero = np.zeros(n * (n + 1) // 2)
    G0 = np.full((3, 3), 0.5)
    g.gradients_uks(ids, zero, zero, G0)
    assert np.array_equal(G0, np.full((3, 3), 0.5))       # every block skipped: nothing added
    Pa, Pb = _pair(n, 3)
    G1 = g.gradients_uks(ids, Pa, Pb)
    G2 = np.full((3, 3), 0.25)
    g.gradients_uks(ids, Pa, Pb, G2)
    g.close()
    assert np.max(np.abs(G2 - 0.25 - G1)) <= 1e-14


def test_alternating_gradient_and_quadrature_calls(oracle_lib):
    """Restricted and unrestricted gradients and quadratures in turn on one handle give each call's own result."""
    Z, xyz, spec, n = water_system(False)
    Pa, Pb = _pair(n, 71)
    ids = list(functional_ids("pbe"))
    h = oracle_lib.setup(2, Z, xyz, spec, False, ids)
    gu0, gr0 = h.gradients_uks(Pa, Pb, 4), h.gradients(Pa + Pb, 4)
    e0 = h.integrate_uks(Pa, Pb, 4)[0]
    g = _grid(Z, xyz, spec, False, "mixed")
    for _ in range(2):
        assert np.max(np.abs(g.gradients_uks(ids, Pa, Pb) - gu0)) <= 1e-9
        assert np.max(np.abs(g.gradients(ids, Pa + Pb) - gr0)) <= 1e-9
        assert abs(g.integrate_uks(ids, Pa, Pb)[0] - e0) <= 1e-6
    g.close()


def test_drop_in_gradients_unrestricted(oracle_lib):
    """DFTGridIntegrator.Gradients with total and spin densities: alpha/beta conversion in, qcGradients3AU incremented."""
    from pdynamo3_amd import integrator
    Z, xyz, spec, n = water_system(True)
    Pa, Pb = _pair(n, 81)
    g0 = oracle_lib.setup(2, Z, xyz, spec, True, list(functional_ids("blyp"))).gradients_uks(Pa, Pb, 4)

    class Holder:
        pass
    target = Holder()
    target.scratch, target.qcState, target.qcModel, target.electronicState = Holder(), Holder(), Holder(), Holder()
    target.qcState.atomicNumbers = Z
    target.qcState.orbitalBases = integrator.OrbitalBasis("6-31g_st", Z, True)
    target.qcModel.functionalModel = integrator.FunctionalModel("blyp")
    target.electronicState.isSpinRestricted = False
    target.scratch.qcCoordinates3AU = xyz
    target.scratch.qcEnergyReport = {}
    target.scratch.onePDMP, target.scratch.onePDMQ = Holder(), Holder()
    target.scratch.onePDMP.density, target.scratch.onePDMQ.density = Pa + Pb, Pa - Pb
    target.scratch.doGradients = True
    target.scratch.qcGradients3AU = np.full((3, 3), 0.125)
    dft = integrator.DFTGridIntegrator(precision="fp64")
    dft.BuildGrid(target)
    dft.Gradients(target)
    target.scratch.dftGrid.close()
    assert np.max(np.abs(target.scratch.qcGradients3AU - 0.125 - g0)) <= 1e-9
