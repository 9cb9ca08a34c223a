"""The pDynamo adapter's conversions on the CPU, with stand-ins shaped like pDynamo's objects (tests/pdynamo_standins.py):
basis containers become the element specification of the parameter files, packed matrices and coordinates are read and
incremented in place, and EstimateMemory follows the reference's host-byte semantics."""

import numpy as np
import pytest

import pdynamo_standins as pd
from helpers import random_density


def test_container_gives_the_parameter_files_specification():
    from pdynamo3_amd import basis, molecules, pdynamo_adapter
    Z, _ = molecules.water()
    for label in ("6-31g_st", "def2-svp"):
        b = basis.load_basis(label)
        for spherical in (True, False):
            c = pd.GaussianBasisContainerStandIn(b, Z, spherical)
            spec, sph = pdynamo_adapter.basis_from_container(c, Z)
            assert sph == spherical
            assert np.array_equal(spec, basis.element_spec(b, Z))
            assert len(c) == basis.number_of_functions(b, Z, spherical)


def test_unnormalised_primitives_are_refused():
    from pdynamo3_amd import basis, molecules, pdynamo_adapter
    Z, _ = molecules.water()
    c = pd.GaussianBasisContainerStandIn(basis.load_basis(), Z, True)
    entry = c.uniqueEntries[8]
    state = entry.__getstate__()
    state["Normalized Primitives"] = False
    entry.__getstate__ = lambda: state
    with pytest.raises(basis.BasisError):
        pdynamo_adapter.basis_from_container(c, Z)


def test_packed_and_coordinates_round_trip_in_place():
    from pdynamo3_amd import pdynamo_adapter as ad
    P = random_density(7, 3, seed=1)
    m = pd.SymmetricMatrixStandIn(7, P)
    assert np.array_equal(ad.packed(m), P)
    d = np.arange(P.size, dtype=np.float64)
    ad.increment_packed(m, d)
    assert np.array_equal(m.values(), P + d)
    x = np.arange(12.0).reshape(4, 3)
    c = pd.Coordinates3StandIn(x)
    assert np.array_equal(ad.coordinates(c), x)
    ad.increment_coordinates(c, np.ones((4, 3)))
    assert np.array_equal(c.values(), x + 1.0)
    # numpy arrays are viewed and incremented in one step
    a = P.copy()
    ad.increment_packed(a, d)
    assert np.array_equal(a, P + d)


def test_estimate_memory_is_host_bytes():
    from pdynamo3_amd import basis, molecules, pdynamo_adapter
    Z, xa, _, _ = molecules.water_cluster(4, 9.0 * 0.25 ** (1.0 / 3.0), 1)
    b = basis.load_basis()
    state = pd.target(Z, molecules.to_bohr(xa), b, False, [106, 131], None).qcState
    gi = pdynamo_adapter.DFTGridIntegratorMI355X()
    n = len(state.orbitalBases)
    host = gi.EstimateMemory(None, state)
    # The packed P/V staging dominates; it is far below the reference's in-core figure and the device peak.
    assert host >= 8.0 * n * (n + 1) / 2
    p = 4 * 42086
    assert host < 32.0 * p + p * n * (4.0 + 8.0 * 4)
    assert gi.SummaryItems()[0] == ("Grid Accuracy", "Medium")


def test_accuracy_takes_pdynamos_enum():
    import enum
    from pdynamo3_amd import pdynamo_adapter

    class DFTGridAccuracy(enum.Enum):     # pDynamo's is a plain Enum (DFTDefinitions.py:11-17)
        High = 3
    assert pdynamo_adapter.DFTGridIntegratorMI355X(accuracy=DFTGridAccuracy.High).accuracy.name == "High"
