"""Drop-in grid integrator for pDynamo3's QCModelDFT: pDynamo's own objects in, the GPU quadrature underneath.

QCModelDFT takes its grid integrator as an attribute (pMolecule/QCModel/QCModelDFT.py:51, default built at :341-344) and
calls BuildGrid, EstimateMemory, Fock, Gradients and SummaryItems on it (pMolecule/QCModel/DFTGridIntegrator.py:17-93).
DFTGridIntegratorMI355X has those methods and reads and writes the objects the reference's integrator reads and writes:

  target.qcState.atomicNumbers        IntegerArray1D (any sequence of ints)
  target.scratch.qcCoordinates3AU     Coordinates3, n x 3 in bohr (RealArray2D: .rows and [i, j])
  target.qcState.orbitalBases         GaussianBasisContainer: .uniqueEntries {Z: GaussianBasis}; each entry's
                                      __getstate__() is the mapping of its parameter file (GaussianBasis.pyx:67-91) and
                                      .isSpherical its function type; len() is the number of functions
  target.qcModel.functionalModel      DFTFunctionalModel: .ids (IntegerArray1D of libxc IDs), .order
  scratch.onePDMP.density / .fock     SymmetricMatrix: packed lower triangle in .block (RealBlock, [k] get and set),
  scratch.onePDMQ.density / .fock     .extent; the Fock matrices are INCREMENTED in place, as DFTIntegrator_Integrate does
  scratch.qcGradients3AU              Coordinates3, incremented in place
  scratch.qcEnergyReport              the same keys as DFTGridIntegrator.Fock writes
  target.electronicState.isSpinRestricted

Errors: a non-zero status of the library raises QCModelError("Error in DFT integrator ..."), as DFTIntegrator.pyx:72
does; the caller may pass its own exception class (pDynamo's QCModelError) as `error_class`.

EstimateMemory returns, as the reference's does, the HOST bytes that the integrator keeps for a job (QCModelDFT.CheckMemory
adds it to its other host estimates and compares the sum with maximumMemory, QCModelDFT.py:73-94).  The grid, the basis
values and the work arrays live in device memory, which the library checks itself (status 7 when it is short);
EstimateDeviceMemory gives that peak for admission control.

Element access to RealBlock and RealArray2D goes through Python item calls (pDynamo's arrays expose no buffer protocol);
objects that numpy can view directly (numpy arrays, anything with __array_interface__) are read and written in one step.
"""

import numpy as np

from .basis import element_from_mapping, element_spec, BasisError
from .definitions import DFTGridAccuracy
from .integrator import QCModelError, QuadratureGrid, estimate_bytes, estimated_points

# Values per point of each function for a functional of the given order, as the reference counts them
# (DFTGridIntegrator.py:10-13).
_NUMBER_OF_FUNCTIONS = {0: 1, 1: 4, 2: 10}


# ---------------------------------------------------------------------------------------------------------------------
# Conversions between pDynamo's objects and arrays.
# ---------------------------------------------------------------------------------------------------------------------
def _is_numpy_viewable(obj):
    return isinstance(obj, np.ndarray) or hasattr(obj, "__array_interface__")


def atomic_numbers(obj):
    """IntegerArray1D (or any sequence) of atomic numbers as int32."""
    return np.ascontiguousarray([int(z) for z in obj], dtype=np.int32)


def coordinates(c3):
    """Coordinates3 (n x 3) as a float64 array."""
    if _is_numpy_viewable(c3):
        return np.ascontiguousarray(c3, dtype=np.float64).reshape(-1, 3)
    return np.asarray([[c3[i, j] for j in range(3)] for i in range(c3.rows)], dtype=np.float64)


def increment_coordinates(c3, g):
    """c3 += g (Coordinates3 incremented in place, as the reference increments its gradients)."""
    if _is_numpy_viewable(c3):
        view = np.asarray(c3)
        view += g.reshape(view.shape)
        return
    for i in range(g.shape[0]):
        for j in range(3):
            c3[i, j] = c3[i, j] + float(g[i, j])


def packed(matrix):
    """A SymmetricMatrix's packed lower triangle as float64 (i(i+1)/2 + j, SymmetricMatrix.h:43-46)."""
    if _is_numpy_viewable(matrix):
        return np.ascontiguousarray(matrix, dtype=np.float64).ravel()
    block = matrix.block
    n = matrix.extent
    size = n * (n + 1) // 2
    return np.fromiter((block[k] for k in range(size)), dtype=np.float64, count=size)


def increment_packed(matrix, values):
    """matrix += values, on the SymmetricMatrix's own storage."""
    if _is_numpy_viewable(matrix):
        view = np.asarray(matrix)
        view.reshape(-1)[:] += values
        return
    block = matrix.block
    for k in range(values.size):
        block[k] = block[k] + float(values[k])


def basis_from_container(container, numbers):
    """(element specification, spherical) of a GaussianBasisContainer: each unique entry's parameter mapping
    (GaussianBasis.__getstate__ returns the exponents and coefficients as the parameter file gives them, with the
    'Normalized Primitives' flag the library's unnormalisation assumes) and its function type."""
    entries = container.uniqueEntries
    basis = {"label": getattr(container, "label", None) or "container", "elements": {}}
    spherical = set()
    for z in sorted(set(int(a) for a in numbers)):
        entry = entries[z]
        mapping = entry.__getstate__()
        element = element_from_mapping(mapping)
        if not element["normalizedPrimitives"]:
            raise BasisError("Basis entry for Z = {:d} has unnormalised primitives; the GPU integrator takes "
                             "primitive-normalised coefficients.".format(z))
        basis["elements"][str(z)] = element
        spherical.add(bool(getattr(entry, "isSpherical", mapping.get("Is Spherical", True))))
    if len(spherical) != 1:
        raise BasisError("The container mixes spherical and Cartesian entries.")
    return element_spec(basis, numbers), spherical.pop()


def functional_ids(functional_model):
    return [int(i) for i in functional_model.ids]


# ---------------------------------------------------------------------------------------------------------------------
# The plug-in.
# ---------------------------------------------------------------------------------------------------------------------
class DFTGridIntegratorMI355X:
    """Drop-in for pMolecule.QCModel.DFTGridIntegrator.DFTGridIntegrator with the quadrature on one MI355X.

    accuracy  : DFTGridAccuracy (default Medium, DFTGridIntegrator.py:22)
    inCore    : accepted for interface parity; the basis values always stay in device memory across calls
    precision : 'fp64' (default; matches the reference's FP64 quadrature to rounding) or 'mixed'
    """

    def __init__(self, accuracy=DFTGridAccuracy.Medium, inCore=False, precision="fp64", error_class=QCModelError):
        # pDynamo's DFTGridAccuracy is a plain Enum with the same values (DFTDefinitions.py:11-17).
        self.accuracy = DFTGridAccuracy(int(getattr(accuracy, "value", accuracy)))
        self.inCore = inCore
        self.precision = precision
        self.error_class = error_class

    def _call(self, fn, *args):
        try:
            return fn(*args)
        except QCModelError as exc:
            if self.error_class is QCModelError:
                raise
            raise self.error_class(str(exc)) from exc

    def _grid(self, target):
        """The job's device grid with its basis values in this integrator's precision."""
        scratch = target.scratch
        grid = getattr(scratch, "dftGrid", None)
        if not isinstance(grid, QuadratureGrid):
            raise self.error_class("Error in DFT integrator (no grid: call BuildGrid first).")
        if grid.number_of_functions == 0 or grid.precision != self.precision:
            numbers = atomic_numbers(target.qcState.atomicNumbers)
            spec, spherical = basis_from_container(target.qcState.orbitalBases, numbers)
            self._call(grid.set_basis, spec, spherical, self.precision)
            if grid.number_of_functions != len(target.qcState.orbitalBases):
                raise self.error_class("Error in DFT integrator (basis has {:d} functions, the container {:d}).".format(
                    grid.number_of_functions, len(target.qcState.orbitalBases)))
        return grid

    def BuildGrid(self, target):
        """Build the grid (DFTGridIntegrator.py:25-29); any previous grid of the job is released."""
        old = getattr(target.scratch, "dftGrid", None)
        if isinstance(old, QuadratureGrid):
            old.close()
        numbers = atomic_numbers(target.qcState.atomicNumbers)
        xyz = coordinates(target.scratch.qcCoordinates3AU)
        grid = self._call(QuadratureGrid, numbers, xyz, self.accuracy)
        target.scratch.dftGrid = grid
        target.scratch.qcEnergyReport["Quadrature Points"] = grid.number_of_points

    def EstimateMemory(self, qcModel, qcState):
        """Host bytes this integrator keeps for a job (the quantity DFTGridIntegrator.EstimateMemory estimates for the
        reference, DFTGridIntegrator.py:31-39): the packed staging of P and V (8 n(n+1)/2 bytes) and, per block of 128
        points, its host table entry and partial sums (64 bytes).  The in-core basis values, the reference's
        f (4 + 8 r) bytes, are in device memory here (EstimateDeviceMemory)."""
        p = float(estimated_points(self.accuracy, atomic_numbers(qcState.atomicNumbers)))
        n = float(len(qcState.orbitalBases))
        return 8.0 * n * (n + 1.0) / 2.0 + 64.0 * (p / 128.0 + 1.0)

    def EstimateDeviceMemory(self, qcModel, qcState):
        """Peak device bytes of a job (grid, basis values in this precision, work arrays), for admission control; the
        reference's in-core figure f (4 + 8 r) is the host counterpart of its basis-value part."""
        numbers = atomic_numbers(qcState.atomicNumbers)
        p = estimated_points(self.accuracy, numbers)
        n = len(qcState.orbitalBases)
        return estimate_bytes(p, n, float(n), self.precision)

    def Fock(self, target):
        """Add V_xc into the Fock matrices and return E_xc (DFTGridIntegrator.py:41-70).  Unrestricted states
        (onePDMQ present, not spin restricted): alpha = (T + S)/2 and beta = (T - S)/2 go to the GPU and the total and
        spin Fock matrices receive (F_a + F_b)/2 and (F_a - F_b)/2 (DFTIntegrator.pyx:82-107)."""
        scratch = target.scratch
        grid = self._grid(target)
        ids = functional_ids(target.qcModel.functionalModel)
        restricted = target.electronicState.isSpinRestricted
        onePDMQ = getattr(scratch, "onePDMQ", None)
        if not restricted and onePDMQ is None:
            raise self.error_class("Error in DFT integrator (an unrestricted state needs onePDMQ).")
        T = packed(scratch.onePDMP.density)
        if restricted:
            V = np.zeros_like(T)
            eQuad, rhoQuad, stats = self._call(grid.integrate, ids, T, V)
            increment_packed(scratch.onePDMP.fock, V)
        else:
            S = packed(onePDMQ.density)
            Fa, Fb = np.zeros_like(T), np.zeros_like(T)
            eQuad, rhoQuad, stats = self._call(grid.integrate_uks, ids, 0.5 * (T + S), 0.5 * (T - S), Fa, Fb)
            increment_packed(scratch.onePDMP.fock, 0.5 * (Fa + Fb))
            increment_packed(onePDMQ.fock, 0.5 * (Fa - Fb))
        report = scratch.qcEnergyReport
        report["Quadrature Electron Count"] = (rhoQuad, "{:.3f}")
        report["Quadrature Energy"] = eQuad
        n = stats["bf_values"]
        o = float(len(target.qcState.orbitalBases))
        report["Quadrature BF Values"] = n
        report["Quadrature BF Sparsity (%)"] = ((1.0 - n / (grid.number_of_points * o)) * 100.0, "{:.1f}")
        report["Quadrature BF Storage (GB)"] = (grid.device_bytes() / 1.0e9, "{:.3f}")
        return eQuad

    def Gradients(self, target):
        """Add the XC terms to scratch.qcGradients3AU when scratch.doGradients is set (DFTGridIntegrator.py:72-88)."""
        scratch = target.scratch
        if not getattr(scratch, "doGradients", False):
            return
        grid = self._grid(target)
        ids = functional_ids(target.qcModel.functionalModel)
        T = packed(scratch.onePDMP.density)
        if target.electronicState.isSpinRestricted:
            G = self._call(grid.gradients, ids, T)
        else:
            onePDMQ = getattr(scratch, "onePDMQ", None)
            if onePDMQ is None:
                raise self.error_class("Error in DFT integrator (an unrestricted state needs onePDMQ).")
            S = packed(onePDMQ.density)
            G = self._call(grid.gradients_uks, ids, 0.5 * (T + S), 0.5 * (T - S))
        increment_coordinates(scratch.qcGradients3AU, G)

    def SummaryItems(self):
        """Summary items (DFTGridIntegrator.py:90-93), and the precision of the GPU path."""
        return [("Grid Accuracy", "{:s}".format(self.accuracy.name)),
                ("Save Grid Function Data", "{:s}".format(repr(self.inCore))),
                ("Quadrature Precision", self.precision)]
