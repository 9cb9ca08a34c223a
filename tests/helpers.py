"""Shared inputs for the tests: systems, densities and packed-matrix helpers."""

import numpy as np

from pdynamo3_amd import basis, molecules
from pdynamo3_amd.definitions import FUNCTIONAL_IDS


def packed(M):
    """Lower triangle of a square matrix, row-major (SymmetricMatrix layout)."""
    return np.ascontiguousarray(M[np.tril_indices(M.shape[0])])


def unpacked(v, n):
    M = np.zeros((n, n))
    M[np.tril_indices(n)] = v
    return M + np.tril(M, -1).T


def random_density(n, nocc, seed, scale=0.3):
    """A symmetric positive semi-definite density 2 C C^T (C random, n x nocc), packed."""
    rng = np.random.default_rng(seed)
    C = rng.normal(size=(n, nocc)) * scale
    return packed(2.0 * C @ C.T)


def water_system(spherical=True, label="6-31g_st"):
    Z, xyz = molecules.water()
    b = basis.load_basis(label)
    return Z, xyz, basis.element_spec(b, Z), basis.number_of_functions(b, Z, spherical)


def methane_angstrom():
    """CH4, C at the origin, C-H 1.09 Angstrom along the tetrahedral directions (slightly distorted)."""
    d = 1.09 / np.sqrt(3.0)
    xyz = np.array([[0.0, 0.0, 0.0], [d, d, d], [-d, -d, d], [-d, d, -d], [d, -d, -d * 1.02]])
    return np.array([6, 1, 1, 1, 1], dtype=np.int32), xyz


def cluster_system(n_waters, cube, seed, spherical=True):
    Z, xyz_a, _, _ = molecules.water_cluster(n_waters, cube, seed)
    xyz = molecules.to_bohr(xyz_a)
    b = basis.load_basis("6-31g_st")
    return Z, xyz, basis.element_spec(b, Z), basis.number_of_functions(b, Z, spherical)


def ids(name):
    return list(FUNCTIONAL_IDS[name])
