"""Synthetic densities: monomer fixture and rotated cluster densities."""

import numpy as np
import pytest

from helpers import unpacked
from pdynamo3_amd import basis, densities, molecules


@pytest.mark.parametrize("spherical", [True, False])
def test_monomer_density_fixture(spherical):
    P, meta = densities.load_monomer_density(spherical=spherical)
    n = basis.number_of_functions(basis.load_basis(), [8, 1, 1], spherical)
    assert P.shape == (n, n)
    assert np.allclose(P, P.T)
    assert meta["electrons_tr_PS"] == pytest.approx(10.0, abs=1e-9)
    assert np.min(np.linalg.eigvalsh(P)) > -1e-10


def test_spherical_rotations_are_orthogonal():
    R = molecules.random_rotation(np.random.default_rng(1))
    for l in (1, 2, 3):
        D = densities.shell_rotation(l, R, True)
        assert np.allclose(D @ D.T, np.eye(2 * l + 1), atol=1e-12)


def test_p_rotation_is_the_rotation_matrix():
    R = molecules.random_rotation(np.random.default_rng(2))
    D = densities.shell_rotation(1, R, False)     # Cartesian p: x, y, z; A(R^T u) = R^T u
    assert np.allclose(D, R.T, atol=1e-12)


def test_identity_rotation_reproduces_the_monomer():
    P1, _ = densities.load_monomer_density(spherical=True)
    Pc = densities.cluster_density([np.eye(3), np.eye(3)], spherical=True)
    M = unpacked(Pc, 2 * P1.shape[0])
    m = P1.shape[0]
    assert np.allclose(M[:m, :m], P1) and np.allclose(M[m:, m:], P1) and np.allclose(M[:m, m:], 0.0)
