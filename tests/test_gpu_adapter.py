"""The pDynamo adapter on the GPU, driven with stand-ins shaped like pDynamo's objects (tests/pdynamo_standins.py) as
QCModelDFT drives its grid integrator: BuildGrid, then Fock (the Fock matrices incremented in place) and Gradients (the
gradient Coordinates3 incremented in place), restricted and unrestricted, against the FP64 oracle on the same inputs."""

import numpy as np
import pytest

import pdynamo_standins as pd
from helpers import ids, random_density

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("spherical", [True, False])
def test_restricted_fock_and_gradients_through_the_adapter(oracle_lib, spherical):
    from pdynamo3_amd import basis, molecules, pdynamo_adapter
    Z, xyz = molecules.water()
    b = basis.load_basis()
    n = basis.number_of_functions(b, Z, spherical)
    P = random_density(n, 5, seed=7)
    t = pd.target(Z, xyz, b, spherical, ids("blyp"), P)
    F0 = np.linspace(-1.0, 1.0, P.size)
    pdynamo_adapter.increment_packed(t.scratch.onePDMP.fock, F0)       # the J-filled Fock matrix it adds to
    G0 = np.full((3, 3), 0.25)
    pdynamo_adapter.increment_coordinates(t.scratch.qcGradients3AU, G0)
    gi = pdynamo_adapter.DFTGridIntegratorMI355X()
    gi.BuildGrid(t)
    E = gi.Fock(t)
    gi.Gradients(t)
    h = oracle_lib.setup(2, Z, xyz, basis.element_spec(b, Z), spherical, ids("blyp"))
    e0, n0, V0, _ = h.integrate(P)
    g0 = h.gradients(P)
    report = t.scratch.qcEnergyReport
    print("E", E, "oracle", e0, "max dV", np.max(np.abs(t.scratch.onePDMP.fock.values() - F0 - V0)))
    assert abs(E - e0) < 1e-9 and report["Quadrature Energy"] == E
    assert abs(report["Quadrature Electron Count"][0] - n0) < 1e-9
    assert report["Quadrature Points"] == h.npoints
    assert np.max(np.abs(t.scratch.onePDMP.fock.values() - F0 - V0)) < 1e-9
    assert np.max(np.abs(t.scratch.qcGradients3AU.values() - G0 - g0)) < 1e-9
    t.scratch.dftGrid.close()


def test_unrestricted_fock_through_the_adapter(oracle_lib):
    from pdynamo3_amd import basis, molecules, pdynamo_adapter
    Z, xyz = molecules.water()
    b = basis.load_basis()
    n = basis.number_of_functions(b, Z, True)
    Pa, Pb = random_density(n, 5, seed=8), random_density(n, 4, seed=9)
    t = pd.target(Z, xyz, b, True, ids("pbe"), Pa + Pb, spin_density=Pa - Pb, restricted=False)
    gi = pdynamo_adapter.DFTGridIntegratorMI355X(precision="mixed")
    gi.BuildGrid(t)
    E = gi.Fock(t)
    h = oracle_lib.setup(2, Z, xyz, basis.element_spec(b, Z), True, ids("pbe"))
    e0, n0, Va, Vb, _ = h.integrate_uks(Pa, Pb)
    FT = t.scratch.onePDMP.fock.values()
    FS = t.scratch.onePDMQ.fock.values()
    print("E", E, "oracle", e0)
    assert abs(E - e0) < 1e-6
    assert np.max(np.abs(FT - 0.5 * (Va + Vb))) < 1e-7 and np.max(np.abs(FS - 0.5 * (Va - Vb))) < 1e-7
    gi.BuildGrid(t)          # a new grid replaces the old one
    assert t.scratch.qcEnergyReport["Quadrature Points"] == h.npoints
    t.scratch.dftGrid.close()


def test_library_errors_become_the_callers_error_class():
    from pdynamo3_amd import basis, molecules, pdynamo_adapter

    class CallersError(Exception):
        pass
    Z, xyz = molecules.water()
    b = basis.load_basis()
    n = basis.number_of_functions(b, Z, True)
    t = pd.target(Z, xyz, b, True, [999], random_density(n, 5, seed=1))
    gi = pdynamo_adapter.DFTGridIntegratorMI355X(error_class=CallersError)
    gi.BuildGrid(t)
    with pytest.raises(CallersError, match="Error in DFT integrator"):
        gi.Fock(t)
    t.scratch.dftGrid.close()
