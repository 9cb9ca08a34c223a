"""GPU parity: the HIP library against the FP64 oracle on the same inputs (water, small water clusters).

Tolerances: FP64 path |dE| <= 1e-9 Eh and max|dV| <= 1e-9 (rounding-level differences: different summation order, device
libm); mixed path |dE| <= 1e-6 Eh and max|dV| <= 1e-7 (the target of the mixed-precision quadrature).
"""

import numpy as np
import pytest

from helpers import cluster_system, ids, random_density, water_system

pytestmark = pytest.mark.gpu


def _integrator():
    from pdynamo3_amd import integrator
    return integrator


@pytest.mark.parametrize("spherical", [True, False])
def test_water_grid_matches_oracle(oracle_lib, spherical):
    I = _integrator()
    Z, xyz, spec, n = water_system(spherical)
    g = I.QuadratureGrid(Z, xyz, 2)
    h = oracle_lib.setup(2, Z, xyz, spec, spherical, ids("blyp"))
    assert g.number_of_points == h.npoints
    assert g.number_of_blocks == h.nblocks
    gp, gba, gbc = g.points()
    op, oba, obc = h.grid()
    assert np.array_equal(gba, oba) and np.array_equal(gbc, obc)
    assert np.max(np.abs(gp[:, :3] - op[:, :3])) <= 1e-12
    assert np.max(np.abs(gp[:, 3] - op[:, 3]) / np.maximum(np.abs(op[:, 3]), 1e-300)) <= 1e-11


@pytest.mark.parametrize("functional", ["blyp", "pbe"])
@pytest.mark.parametrize("spherical", [True, False])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_water_quadrature_matches_oracle(oracle_lib, functional, spherical, precision):
    I = _integrator()
    Z, xyz, spec, n = water_system(spherical)
    P = random_density(n, 5, seed=11)
    h = oracle_lib.setup(2, Z, xyz, spec, spherical, ids(functional))
    e0, n0, V0, _ = h.integrate(P)
    g = I.QuadratureGrid(Z, xyz, 2)
    g.set_basis(spec, spherical, precision)
    V = np.zeros_like(P)
    e, ne, stats = g.integrate(ids(functional), P, V)
    tolE, tolV = (1e-9, 1e-9) if precision == "fp64" else (1e-6, 1e-7)
    print("dE", e - e0, "dN", ne - n0, "max dV", np.max(np.abs(V - V0)), stats)
    assert abs(e - e0) <= tolE
    assert abs(ne - n0) <= tolE
    assert np.max(np.abs(V - V0)) <= tolV
    assert stats["bf_values"] == pytest.approx(np.sum([len(h.block_values(b)[0]) * c for b, c in
                                                       enumerate(h.grid()[2])]))


def test_fock_is_incremented(oracle_lib):
    I = _integrator()
    Z, xyz, spec, n = water_system(True)
    P = random_density(n, 5, seed=3)
    g = I.QuadratureGrid(Z, xyz, 2)
    g.set_basis(spec, True, "fp64")
    V1 = np.zeros_like(P)
    g.integrate(ids("blyp"), P, V1)
    V2 = np.full_like(P, 0.25)
    g.integrate(ids("blyp"), P, V2)
    assert np.max(np.abs(V2 - 0.25 - V1)) <= 1e-14


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cluster4_matches_oracle(oracle_lib, precision):
    I = _integrator()
    Z, xyz, spec, n = cluster_system(4, 6.0, seed=5)
    P = random_density(n, 20, seed=7, scale=0.15)
    h = oracle_lib.setup(2, Z, xyz, spec, True, ids("pbe"))
    e0, n0, V0, _ = h.integrate(P, threads=8)
    g = I.QuadratureGrid(Z, xyz, 2)
    assert g.number_of_points == h.npoints
    g.set_basis(spec, True, precision)
    V = np.zeros_like(P)
    e, ne, stats = g.integrate(ids("pbe"), P, V)
    tolE, tolV = (1e-9, 1e-9) if precision == "fp64" else (1e-6, 1e-7)
    print("dE", e - e0, "dN", ne - n0, "max dV", np.max(np.abs(V - V0)), stats)
    assert abs(e - e0) <= tolE
    assert np.max(np.abs(V - V0)) <= tolV
