"""XC nuclear gradients on the GPU (xq_gradients, FP64) against the oracle (itself pinned against the compiled reference and
finite differences of E_xc in test_oracle_gradients.py), on the same inputs.

Tolerance 1e-9 hartree/bohr (different summation orders; the weight terms are formed in O(nAtoms^2) per point instead of
the reference's O(nAtoms^3)).  The gradient is FP64 on either precision's handle.
"""

import numpy as np
import pytest

from helpers import cluster_system, ids, random_density, water_system

pytestmark = pytest.mark.gpu


def _grid(Z, xyz, spec, spherical, precision):
    from pdynamo3_amd import integrator
    g = integrator.QuadratureGrid(Z, xyz, 2)
    g.set_basis(spec, spherical, precision)
    return g


@pytest.mark.parametrize("functional", ["blyp", "pbe"])
@pytest.mark.parametrize("spherical", [True, False])
def test_water_gradients_match_oracle(oracle_lib, functional, spherical):
    Z, xyz, spec, n = water_system(spherical)
    P = random_density(n, 5, seed=11)
    g0 = oracle_lib.setup(2, Z, xyz, spec, spherical, ids(functional)).gradients(P, 4)
    g = _grid(Z, xyz, spec, spherical, "fp64")
    G = g.gradients(ids(functional), P)
    g.close()
    print("max dG", np.max(np.abs(G - g0)))
    assert np.max(np.abs(G - g0)) <= 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cluster4_gradients_match_oracle(oracle_lib, precision):
    Z, xyz, spec, n = cluster_system(4, 6.0, 5, False)
    P = random_density(n, 20, seed=17)
    g0 = oracle_lib.setup(2, Z, xyz, spec, False, ids("pbe")).gradients(P, 8)
    g = _grid(Z, xyz, spec, False, precision)
    G = g.gradients(ids("pbe"), P)
    g.close()
    print("max dG", np.max(np.abs(G - g0)), "max G", np.max(np.abs(g0)))
    assert np.max(np.abs(G - g0)) <= 1e-9


def test_h2o16_gradients_sum_to_zero():
    from pdynamo3_amd import densities
    Z, xyz, spec, n = cluster_system(16, 9.0, 1, False)
    from pdynamo3_amd import molecules
    _, _, rots, _ = molecules.water_cluster(16, 9.0, 1)
    P = densities.cluster_density(rots, spherical=False)
    g = _grid(Z, xyz, spec, False, "mixed")
    G = g.gradients(ids("pbe"), P)
    g.close()
    print("sum", G.sum(axis=0), "max", np.max(np.abs(G)))
    assert np.max(np.abs(G.sum(axis=0))) <= 1e-9
    assert np.max(np.abs(G)) > 1e-3


def test_gradients_increment_and_zero_density():
    Z, xyz, spec, n = water_system(True)
    g = _grid(Z, xyz, spec, True, "fp64")
    G0 = np.full((3, 3), 0.5)
    g.gradients(ids("blyp"), np.zeros(n * (n + 1) // 2), G0)
    assert np.array_equal(G0, np.full((3, 3), 0.5))       # every block skipped: nothing added
    P = random_density(n, 5, seed=3)
    G1 = g.gradients(ids("blyp"), P)
    G2 = np.full((3, 3), 0.25)
    g.gradients(ids("blyp"), P, G2)
    g.close()
    assert np.max(np.abs(G2 - 0.25 - G1)) <= 1e-14


def test_methane_def2svp_gradients_match_oracle(oracle_lib):
    from pdynamo3_amd import basis, molecules
    from helpers import methane_angstrom
    Zm, xa = methane_angstrom()
    xyz = molecules.to_bohr(xa)
    b = basis.load_basis("def2-svp")
    spec, n = basis.element_spec(b, Zm), basis.number_of_functions(b, Zm, True)
    P = random_density(n, 5, seed=24)
    g0 = oracle_lib.setup(2, Zm, xyz, spec, True, ids("blyp")).gradients(P, 8)
    g = _grid(Zm, xyz, spec, True, "fp64")
    G = g.gradients(ids("blyp"), P)
    g.close()
    print("max dG", np.max(np.abs(G - g0)))
    assert np.max(np.abs(G - g0)) <= 1e-9


@pytest.mark.parametrize("name", ["lda", "b3lyp"])
def test_lda_and_b3lyp_gradients_match_oracle(oracle_lib, name):
    from pdynamo3_amd.definitions import functional_ids
    ids = list(functional_ids(name))
    Z, xyz, spec, n = water_system(True)
    P = random_density(n, 5, seed=26)
    g0 = oracle_lib.setup(2, Z, xyz, spec, True, ids).gradients(P, 4)
    g = _grid(Z, xyz, spec, True, "mixed")
    G = g.gradients(ids, P)
    g.close()
    print(name, "max dG", np.max(np.abs(G - g0)))
    assert np.max(np.abs(G - g0)) <= 1e-9
