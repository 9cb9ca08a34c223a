"""Stand-ins shaped like the pDynamo objects the grid integrator reads and writes, for testing the adapter without pDynamo.

Each keeps the Python surface the adapter uses and no numpy interface, so the element-wise paths are exercised:
  SymmetricMatrixStandIn   .extent, .block (RealBlock: [k] get/set, len), packed lower triangle (SymmetricMatrix.pyx)
  Coordinates3StandIn      .rows, [i, j] get/set (RealArray2D)
  GaussianBasisStandIn     __getstate__() -> the parameter-file mapping, .isSpherical (GaussianBasis.pyx:67-91)
  GaussianBasisContainerStandIn   .uniqueEntries {Z: entry}, .label, len() = number of functions
  DFTFunctionalModelStandIn       .ids (sequence of libxc IDs), .order
"""

import types

import numpy as np


class RealBlockStandIn:
    def __init__(self, values):
        self._v = [float(x) for x in values]

    def __getitem__(self, k):
        return self._v[k]

    def __setitem__(self, k, value):
        self._v[k] = float(value)

    def __len__(self):
        return len(self._v)


class SymmetricMatrixStandIn:
    def __init__(self, n, packed_values=None):
        self.extent = n
        size = n * (n + 1) // 2
        self.block = RealBlockStandIn(np.zeros(size) if packed_values is None else packed_values)
        self.size = size

    def values(self):
        return np.asarray(self.block._v)


class Coordinates3StandIn:
    def __init__(self, xyz):
        self._v = [[float(c) for c in row] for row in np.asarray(xyz).reshape(-1, 3)]
        self.rows = len(self._v)

    def __getitem__(self, ij):
        return self._v[ij[0]][ij[1]]

    def __setitem__(self, ij, value):
        self._v[ij[0]][ij[1]] = float(value)

    def values(self):
        return np.asarray(self._v)


class GaussianBasisStandIn:
    def __init__(self, z, element, spherical):
        self.atomicNumber = z
        self.isSpherical = spherical
        self._element = element

    def __getstate__(self):
        e = self._element
        return {"Atomic Number": self.atomicNumber, "Basis Type": "Orbital", "Is Spherical": self.isSpherical,
                "Normalized Primitives": True, "Primitive Fields": ["Exponent", "Coefficients"],
                "Shells": [{"Type": s["type"], "Primitives": [[x, list(c)] for x, c in zip(s["exponents"],
                                                                                          s["coefficients"])]}
                           for s in e["shells"]]}


class GaussianBasisContainerStandIn:
    def __init__(self, basis, atomic_numbers, spherical):
        from pdynamo3_amd.basis import number_of_functions
        self.label = basis["label"]
        self.uniqueEntries = {int(z): GaussianBasisStandIn(int(z), basis["elements"][str(int(z))], spherical)
                              for z in set(int(a) for a in atomic_numbers)}
        self._n = number_of_functions(basis, atomic_numbers, spherical)

    def __len__(self):
        return self._n


class DFTFunctionalModelStandIn:
    def __init__(self, ids, order=1):
        self.ids = list(ids)
        self.order = order


def target(Z, xyz_bohr, basis, spherical, ids, density, spin_density=None, restricted=True):
    """A QC target as QCModelDFT hands it to the grid integrator: qcState, qcModel, electronicState and scratch."""
    container = GaussianBasisContainerStandIn(basis, Z, spherical)
    n = len(container)
    scratch = types.SimpleNamespace(qcCoordinates3AU=Coordinates3StandIn(xyz_bohr), qcEnergyReport={},
                                    onePDMP=types.SimpleNamespace(density=SymmetricMatrixStandIn(n, density),
                                                                  fock=SymmetricMatrixStandIn(n)),
                                    doGradients=True, qcGradients3AU=Coordinates3StandIn(np.zeros((len(Z), 3))))
    if spin_density is not None:
        scratch.onePDMQ = types.SimpleNamespace(density=SymmetricMatrixStandIn(n, spin_density),
                                                fock=SymmetricMatrixStandIn(n))
    return types.SimpleNamespace(
        qcState=types.SimpleNamespace(atomicNumbers=[int(z) for z in Z], orbitalBases=container),
        qcModel=types.SimpleNamespace(functionalModel=DFTFunctionalModelStandIn(ids)),
        electronicState=types.SimpleNamespace(isSpinRestricted=restricted),
        scratch=scratch)
