"""Unrestricted quadrature on the GPU (xq_integrate_uks) against the oracle's (itself equal to the compiled reference's
unrestricted DFTIntegrator_Integrate, tests/test_oracle_uks.py) on the same inputs.

Tolerances as for the restricted path: FP64 |dE| <= 1e-9 Eh and max|dV| <= 1e-9; mixed |dE| <= 1e-6 Eh and max|dV| <= 1e-7.
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


def _tol(precision):
    return (1e-9, 1e-9) if precision == "fp64" else (1e-6, 1e-7)


@pytest.mark.parametrize("name", ["blyp", "pbe", "lda", "b3lyp"])
@pytest.mark.parametrize("spherical", [True, False])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_water_uks_matches_oracle(oracle_lib, name, spherical, precision):
    ids = list(functional_ids(name))
    Z, xyz, spec, n = water_system(spherical)
    Pa = random_density(n, 5, seed=51, scale=0.2)
    Pb = random_density(n, 4, seed=52, scale=0.2)
    e0, n0, Va0, Vb0, _ = oracle_lib.setup(2, Z, xyz, spec, spherical, ids).integrate_uks(Pa, Pb, 4)
    g = _grid(Z, xyz, spec, spherical, precision)
    Va, Vb = np.zeros_like(Pa), np.zeros_like(Pb)
    e, ne, _ = g.integrate_uks(ids, Pa, Pb, Va, Vb)
    g.close()
    tolE, tolV = _tol(precision)
    print(name, spherical, precision, "dE", e - e0, "dVa", np.max(np.abs(Va - Va0)), "dVb", np.max(np.abs(Vb - Vb0)))
    assert abs(e - e0) <= tolE and abs(ne - n0) <= tolE
    assert np.max(np.abs(Va - Va0)) <= tolV and np.max(np.abs(Vb - Vb0)) <= tolV


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cluster4_uks_matches_oracle(oracle_lib, precision):
    Z, xyz, spec, n = cluster_system(4, 6.0, 5, False)
    Pa = random_density(n, 20, seed=53, scale=0.2)
    Pb = random_density(n, 18, seed=54, scale=0.2)
    ids = list(functional_ids("pbe"))
    e0, n0, Va0, Vb0, _ = oracle_lib.setup(2, Z, xyz, spec, False, ids).integrate_uks(Pa, Pb, 8)
    g = _grid(Z, xyz, spec, False, precision)
    Va, Vb = np.zeros_like(Pa), np.zeros_like(Pb)
    e, ne, _ = g.integrate_uks(ids, Pa, Pb, Va, Vb)
    g.close()
    tolE, tolV = _tol(precision)
    print(precision, "dE", e - e0, "dVa", np.max(np.abs(Va - Va0)), "dVb", np.max(np.abs(Vb - Vb0)))
    assert abs(e - e0) <= tolE and abs(ne - n0) <= tolE
    assert np.max(np.abs(Va - Va0)) <= tolV and np.max(np.abs(Vb - Vb0)) <= tolV


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_closed_shell_limit_matches_restricted(precision):
    """P_a = P_b = P/2 through the unrestricted path equals the restricted quadrature of P on the same handle."""
    Z, xyz, spec, n = water_system(True)
    P = random_density(n, 5, seed=55)
    ids = list(functional_ids("blyp"))
    g = _grid(Z, xyz, spec, True, precision)
    V = np.zeros_like(P)
    e, ne, _ = g.integrate(ids, P, V)
    Va, Vb = np.zeros_like(P), np.zeros_like(P)
    eu, neu, _ = g.integrate_uks(ids, P / 2, P / 2, Va, Vb)
    g.close()
    tol = 1e-10 if precision == "fp64" else 1e-7
    assert abs(eu - e) <= tol and abs(neu - ne) <= tol
    assert np.max(np.abs(Va - V)) <= tol and np.max(np.abs(Vb - V)) <= tol


def test_h2o16_uks_mixed_against_fp64():
    from pdynamo3_amd import densities, molecules
    Z, xyz, spec, n = cluster_system(16, 9.0, 1, False)
    _, _, rots, _ = molecules.water_cluster(16, 9.0, 1)
    P = densities.cluster_density(rots, spherical=False)
    Pa, Pb = 0.55 * P, 0.45 * P
    ids = list(functional_ids("b3lyp"))
    out = {}
    for precision in ("fp64", "mixed"):
        g = _grid(Z, xyz, spec, False, precision)
        Va, Vb = np.zeros_like(P), np.zeros_like(P)
        e, ne, _ = g.integrate_uks(ids, Pa, Pb, Va, Vb)
        g.close()
        out[precision] = (e, ne, Va, Vb)
    (e0, n0, Va0, Vb0), (e1, n1, Va1, Vb1) = out["fp64"], out["mixed"]
    print("dE", e1 - e0, "dVa", np.max(np.abs(Va1 - Va0)), "dVb", np.max(np.abs(Vb1 - Vb0)), "E", e0, "N", n0)
    assert abs(e1 - e0) <= 1e-6 and abs(n1 - n0) <= 1e-6
    assert np.max(np.abs(Va1 - Va0)) <= 1e-7 and np.max(np.abs(Vb1 - Vb0)) <= 1e-7
    assert abs(n0 - 160.0) <= 1e-3


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_alternating_restricted_and_unrestricted_calls(precision):
    """Regression: restricted and unrestricted calls in turn on one handle (each launch refills the handle's job table;
    the unrestricted call launches several times and must not overwrite the table before its copy has run)."""
    from pdynamo3_amd import densities, molecules
    Z, xyz, spec, n = cluster_system(4, 6.0, 5, False)
    _, _, rots, _ = molecules.water_cluster(4, 6.0, 5)
    P = densities.cluster_density(rots, spherical=False)
    ids = list(functional_ids("pbe"))
    g = _grid(Z, xyz, spec, False, precision)
    for _ in range(2):
        V = np.zeros_like(P)
        e, ne, _ = g.integrate(ids, P, V)
        Va, Vb = np.zeros_like(P), np.zeros_like(P)
        eu, neu, _ = g.integrate_uks(ids, P / 2, P / 2, Va, Vb)
        Va2, Vb2 = np.zeros_like(P), np.zeros_like(P)
        e2, ne2, _ = g.integrate_uks(ids, 0.55 * P, 0.45 * P, Va2, Vb2)
        assert abs(eu - e) <= 1e-7 and abs(neu - ne) <= 1e-7 and np.max(np.abs(Va - V)) <= 1e-7
        assert abs(ne2 - ne) <= 1e-7 and abs(ne - 40.0) <= 1e-3
    g.close()


def test_drop_in_fock_unrestricted(oracle_lib):
    """DFTGridIntegrator.Fock with total and spin densities: alpha/beta conversion in, total/spin Fock out."""
    from pdynamo3_amd import integrator
    Z, xyz, spec, n = water_system(True)
    Pa = random_density(n, 5, seed=61, scale=0.2)
    Pb = random_density(n, 4, seed=62, scale=0.2)
    e0, n0, Va0, Vb0, _ = oracle_lib.setup(2, Z, xyz, spec, True, list(functional_ids("blyp"))).integrate_uks(Pa, Pb, 4)

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
    target.scratch.onePDMP.fock, target.scratch.onePDMQ.fock = np.zeros_like(Pa), np.zeros_like(Pa)
    dft = integrator.DFTGridIntegrator(precision="fp64")
    dft.BuildGrid(target)
    e = dft.Fock(target)
    target.scratch.dftGrid.close()
    assert abs(e - e0) <= 1e-9
    assert np.max(np.abs(target.scratch.onePDMP.fock - 0.5 * (Va0 + Vb0))) <= 1e-9
    assert np.max(np.abs(target.scratch.onePDMQ.fock - 0.5 * (Va0 - Vb0))) <= 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fully_polarised_hydrogen_matches_oracle(oracle_lib, precision):
    """One hydrogen atom with all density in the alpha spin (zeta = 1); its XC gradient vanishes."""
    from pdynamo3_amd import basis
    Z = np.array([1], dtype=np.int32)
    xyz = np.zeros((1, 3))
    spec = basis.element_spec(basis.load_basis(), Z)
    Pa, Pb = np.array([0.30, 0.25, 0.45]), np.zeros(3)
    tolE, tolV = _tol(precision)
    g = _grid(Z, xyz, spec, True, precision)
    for name in ("blyp", "pbe", "lda", "b3lyp"):
        ids = list(functional_ids(name))
        e0, n0, Va0, Vb0, _ = oracle_lib.setup(2, Z, xyz, spec, True, ids).integrate_uks(Pa, Pb, 2)
        Va, Vb = np.zeros(3), np.zeros(3)
        e, ne, _ = g.integrate_uks(ids, Pa, Pb, Va, Vb)
        print(name, precision, "dE", e - e0, "dVa", np.max(np.abs(Va - Va0)), "dVb", np.max(np.abs(Vb - Vb0)))
        assert abs(e - e0) <= tolE and abs(ne - n0) <= tolE
        assert np.max(np.abs(Va - Va0)) <= tolV and np.max(np.abs(Vb - Vb0)) <= tolV
        assert np.max(np.abs(g.gradients_uks(ids, Pa, Pb))) <= 1e-12
    g.close()
