"""Host-side mirror of pDynamo3's DFT grid integrator on top of libxcquad.

QuadratureGrid wraps one device-resident grid (points, weights, screened basis values); DFTGridIntegrator keeps the
reference's plugin interface (pMolecule/QCModel/DFTGridIntegrator.py:17-93): BuildGrid(target), EstimateMemory(qcModel,
qcState), Fock(target), Gradients(target), SummaryItems(), with the same report keys and the same error behaviour (any
non-zero status raises QCModelError("Error in DFT integrator."), DFTIntegrator.pyx:72).
"""

import ctypes

import numpy as np

from . import _lib
from .basis import element_spec, load_basis, number_of_functions
from .definitions import (DFTGridAccuracy, LDA_C_VWN, LDA_C_VWN_RPA, LDA_X, exact_exchange_scaling,
                          functional_ids)

PRECISIONS = {"fp64": 0, "mixed": 1}
NSTATS = 11         # XQ_NSTATS of include/xcquad.h


class QCModelError(Exception):
    """Error raised by the quadrature (pMolecule/QCModel/QCModelError.py)."""


def _check(status, what):
    if status != 0:
        raise QCModelError("Error in DFT integrator ({:s}: status {:d}, {:s}).".format(what, status, _lib.last_error()))


class QuadratureGrid:
    """One molecule's grid and basis values on the device."""

    def __init__(self, atomic_numbers, xyz_bohr, accuracy=DFTGridAccuracy.Medium):
        self.atomic_numbers = _lib.i32(atomic_numbers)
        self.number_of_atoms = len(self.atomic_numbers)
        self.xyz = _lib.f64(xyz_bohr).reshape(-1, 3)
        self.accuracy = DFTGridAccuracy(accuracy)
        self._h = ctypes.c_void_p()
        _check(_lib.lib.xq_grid_build(int(self.accuracy), len(self.atomic_numbers), _lib.iptr(self.atomic_numbers),
                                      _lib.dptr(self.xyz), ctypes.byref(self._h)), "grid build")
        n_p, n_b = ctypes.c_int(), ctypes.c_int()
        _check(_lib.lib.xq_grid_counts(self._h, ctypes.byref(n_p), ctypes.byref(n_b)), "grid counts")
        self.number_of_points, self.number_of_blocks = n_p.value, n_b.value
        self.number_of_functions = 0
        self.precision = None

    def points(self):
        """(x, y, z, w) per point [G, 4], block atoms [B] and block sizes [B]."""
        xyzw = np.zeros((self.number_of_points, 4))
        ba = np.zeros(self.number_of_blocks, dtype=np.int32)
        bc = np.zeros(self.number_of_blocks, dtype=np.int32)
        _check(_lib.lib.xq_grid_points(self._h, _lib.dptr(xyzw), _lib.iptr(ba), _lib.iptr(bc)), "grid points")
        return xyzw, ba, bc

    def set_basis(self, spec, spherical=True, precision="fp64"):
        """Evaluate and store the screened basis values (precision 'fp64' or 'mixed')."""
        spec = _lib.f64(spec)
        nf = ctypes.c_int()
        _check(_lib.lib.xq_basis_set(self._h, spec.size, _lib.dptr(spec), int(bool(spherical)), PRECISIONS[precision],
                                     ctypes.byref(nf)), "basis")
        self.number_of_functions = nf.value
        self.precision = precision

    def integrate(self, ids, density, fock=None):
        """One quadrature: returns (E_xc, N_e, stats); fock (packed, float64) is incremented in place when given."""
        ids = _lib.i32(ids)
        P = _lib.f64(density)
        n = self.number_of_functions
        if P.size != n * (n + 1) // 2:
            raise QCModelError("Density has {:d} entries; {:d} expected.".format(P.size, n * (n + 1) // 2))
        V = fock if fock is not None else np.zeros_like(P)
        if V.dtype != np.float64 or not V.flags.c_contiguous or V.size != P.size:
            raise QCModelError("Fock matrix must be a contiguous float64 array of the density's size.")
        e, ne = ctypes.c_double(), ctypes.c_double()
        stats = np.zeros(NSTATS)
        _check(_lib.lib.xq_integrate(self._h, ids.size, _lib.iptr(ids), _lib.dptr(P), ctypes.byref(e), ctypes.byref(ne),
                                     _lib.dptr(V), _lib.dptr(stats)), "integrate")
        return e.value, ne.value, {"bf_values": stats[0], "blocks_skipped": int(stats[1]), "device_ms": stats[2],
                                   "rho_ms": stats[3], "fock_ms": stats[4], "bf_squares": stats[5],
                                   "rho_bytes": stats[6], "fock_bytes": stats[7], "bf_squares_padded": stats[8],
                                   "product_ms": stats[9], "product_bytes": stats[10]}

    def integrate_uks(self, ids, density_a, density_b, fock_a=None, fock_b=None):
        """One unrestricted quadrature (alpha and beta densities, packed): returns (E_xc, N_e, stats); fock_a and fock_b
        (packed float64) are incremented in place when given."""
        ids = _lib.i32(ids)
        Pa, Pb = _lib.f64(density_a), _lib.f64(density_b)
        n = self.number_of_functions
        if Pa.size != n * (n + 1) // 2 or Pb.size != Pa.size:
            raise QCModelError("Densities have {:d} and {:d} entries; {:d} expected.".format(Pa.size, Pb.size,
                                                                                          n * (n + 1) // 2))
        Va = fock_a if fock_a is not None else np.zeros_like(Pa)
        Vb = fock_b if fock_b is not None else np.zeros_like(Pb)
        for V in (Va, Vb):
            if V.dtype != np.float64 or not V.flags.c_contiguous or V.size != Pa.size:
                raise QCModelError("Fock matrices must be contiguous float64 arrays of the density's size.")
        e, ne = ctypes.c_double(), ctypes.c_double()
        stats = np.zeros(NSTATS)
        _check(_lib.lib.xq_integrate_uks(self._h, ids.size, _lib.iptr(ids), _lib.dptr(Pa), _lib.dptr(Pb), ctypes.byref(e),
                                         ctypes.byref(ne), _lib.dptr(Va), _lib.dptr(Vb), _lib.dptr(stats)),
               "integrate_uks")
        return e.value, ne.value, {"bf_values": stats[0], "blocks_skipped": int(stats[1]), "device_ms": stats[2]}

    def gradients(self, ids, density, gradients=None):
        """XC nuclear gradients (n_atoms x 3, hartree/bohr) at the density; `gradients` (contiguous float64, n_atoms x 3)
        is incremented in place when given, as the reference increments its Coordinates3."""
        ids = _lib.i32(ids)
        P = _lib.f64(density)
        n = self.number_of_functions
        if P.size != n * (n + 1) // 2:
            raise QCModelError("Density has {:d} entries; {:d} expected.".format(P.size, n * (n + 1) // 2))
        G = gradients if gradients is not None else np.zeros((self.number_of_atoms, 3))
        if G.dtype != np.float64 or not G.flags.c_contiguous or G.size != 3 * self.number_of_atoms:
            raise QCModelError("Gradients must be a contiguous float64 array of n_atoms x 3.")
        _check(_lib.lib.xq_gradients(self._h, ids.size, _lib.iptr(ids), _lib.dptr(P), _lib.dptr(G)), "gradients")
        return G

    def gradients_uks(self, ids, density_a, density_b, gradients=None):
        """Unrestricted XC nuclear gradients (n_atoms x 3) at the alpha and beta densities (packed); `gradients` is
        incremented in place when given."""
        ids = _lib.i32(ids)
        Pa, Pb = _lib.f64(density_a), _lib.f64(density_b)
        n = self.number_of_functions
        if Pa.size != n * (n + 1) // 2 or Pb.size != Pa.size:
            raise QCModelError("Densities have {:d} and {:d} entries; {:d} expected.".format(Pa.size, Pb.size,
                                                                                          n * (n + 1) // 2))
        G = gradients if gradients is not None else np.zeros((self.number_of_atoms, 3))
        if G.dtype != np.float64 or not G.flags.c_contiguous or G.size != 3 * self.number_of_atoms:
            raise QCModelError("Gradients must be a contiguous float64 array of n_atoms x 3.")
        _check(_lib.lib.xq_gradients_uks(self._h, ids.size, _lib.iptr(ids), _lib.dptr(Pa), _lib.dptr(Pb), _lib.dptr(G)),
               "gradients_uks")
        return G

    def device_bytes(self):
        b = ctypes.c_longlong()
        _check(_lib.lib.xq_device_bytes(self._h, ctypes.byref(b)), "device bytes")
        return b.value

    def close(self):
        if self._h:
            _lib.lib.xq_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class QuadratureBatch:
    """Quadratures of several grids (one per job) in one set of launches: xq_integrate_batch.

    Each job's E_xc and N_e equal those of QuadratureGrid.integrate on the same grid alone; its V_xc equals it up to the
    order of the FP64 atomic additions.
    """

    def __init__(self, max_jobs):
        self.max_jobs = int(max_jobs)
        self._h = ctypes.c_void_p()
        _check(_lib.lib.xq_batch_create(self.max_jobs, ctypes.byref(self._h)), "batch create")

    def integrate(self, jobs):
        """jobs: sequence of (grid, ids, density, fock); each fock (packed float64, or None) is incremented in place.
        Returns ([(E_xc, N_e) per job], stats of the whole batch)."""
        n = len(jobs)
        if n < 1 or n > self.max_jobs:
            raise QCModelError("Batch of {:d} jobs; 1 to {:d} allowed.".format(n, self.max_jobs))
        grids = (ctypes.c_void_p * n)()
        n_ids = np.zeros(n, dtype=np.int32)
        ids = np.zeros((n, 4), dtype=np.int32)
        keep = []
        P_ptr, V_ptr = (_lib._dp * n)(), (_lib._dp * n)()
        for j, (grid, job_ids, density, fock) in enumerate(jobs):
            grids[j] = grid._h.value
            job_ids = _lib.i32(job_ids)
            n_ids[j] = job_ids.size
            ids[j, :job_ids.size] = job_ids[:4]
            P = _lib.f64(density)
            m = grid.number_of_functions
            if P.size != m * (m + 1) // 2:
                raise QCModelError("Density of job {:d} has {:d} entries; {:d} expected.".format(j, P.size, m * (m + 1) // 2))
            V = fock if fock is not None else np.zeros_like(P)
            if V.dtype != np.float64 or not V.flags.c_contiguous or V.size != P.size:
                raise QCModelError("Fock matrix of job {:d} must be a contiguous float64 array of the density's size.".format(j))
            keep += [P, V]
            P_ptr[j] = _lib.dptr(P)
            V_ptr[j] = _lib.dptr(V)
        e, ne, stats = np.zeros(n), np.zeros(n), np.zeros(NSTATS)
        _check(_lib.lib.xq_integrate_batch(self._h, n, grids, _lib.iptr(n_ids), _lib.iptr(ids), P_ptr, _lib.dptr(e),
                                           _lib.dptr(ne), V_ptr, _lib.dptr(stats)), "batch integrate")
        return [(float(e[j]), float(ne[j])) for j in range(n)], {
            "bf_values": stats[0], "blocks_skipped": int(stats[1]), "device_ms": stats[2], "rho_ms": stats[3],
            "fock_ms": stats[4], "bf_squares": stats[5], "rho_bytes": stats[6], "fock_bytes": stats[7],
            "bf_squares_padded": stats[8], "product_ms": stats[9], "product_bytes": stats[10]}

    def close(self):
        if self._h:
            _lib.lib.xq_batch_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_memory():
    """(free, total) device bytes."""
    f, t = ctypes.c_longlong(), ctypes.c_longlong()
    _check(_lib.lib.xq_device_memory(ctypes.byref(f), ctypes.byref(t)), "device memory")
    return f.value, t.value


def estimated_points(accuracy, atomic_numbers):
    """DFTGrid_EstimatedPoints (DFTGrid.c:301-356): grid size before the weight cull."""
    z = _lib.i32(atomic_numbers)
    n = ctypes.c_longlong()
    _check(_lib.lib.xq_estimated_points(int(accuracy), z.size, _lib.iptr(z), ctypes.byref(n)), "estimated points")
    return n.value


def estimate_bytes(number_of_points, number_of_functions, mean_kept_functions, precision="fp64"):
    """Peak device bytes of a job of the given size, set-up included (admission control)."""
    b = ctypes.c_longlong()
    _check(_lib.lib.xq_estimate_bytes(int(number_of_points), int(number_of_functions), float(mean_kept_functions),
                                      PRECISIONS[precision], ctypes.byref(b)), "estimate bytes")
    return b.value


# ---------------------------------------------------------------------------------------------------------------------------------
# The plugin interface.
# ---------------------------------------------------------------------------------------------------------------------------------
class OrbitalBasis:
    """The orbital basis of a QC state: a basis label, the per-atom element specification and its shape."""

    def __init__(self, label, atomic_numbers, spherical=True):
        self.label = label
        self.spherical = bool(spherical)
        self._basis = load_basis(label)
        self.spec = element_spec(self._basis, atomic_numbers)
        self.n = number_of_functions(self._basis, atomic_numbers, spherical)

    def __len__(self):
        return self.n


class FunctionalModel:
    """Functional by name ('b3lyp', 'blyp', 'lda', 'pbe'): libxc IDs; order and sigma as DFTFunctionalModel sets them
    (DFTFunctionalModel.c:175-192: order 0 without sigma for LDAs, 1 with sigma for GGAs); exchangeScaling, the fraction of
    exact exchange the caller's Fock build adds (DFTFunctionalModel.c:82-103, read by QCModelDFT.py:343).  The library
    evaluates every model on its GGA path: for an LDA, v_sigma is zero, which leaves E_xc, V_xc and the gradients those of
    the LDA path."""

    def __init__(self, name):
        self.name = name
        self.ids = functional_ids(name)
        self.hasSigma = any(i not in (LDA_X, LDA_C_VWN, LDA_C_VWN_RPA) for i in self.ids)
        self.order = 1 if self.hasSigma else 0
        self.exchangeScaling = exact_exchange_scaling(self.ids)


class DFTGridIntegrator:
    """Drop-in counterpart of pMolecule/QCModel/DFTGridIntegrator.DFTGridIntegrator on the GPU.

    accuracy  : DFTGridAccuracy (default Medium, DFTGridIntegrator.py:22).
    inCore    : accepted for interface parity; basis values always stay on the device across calls.
    precision : 'fp64' or 'mixed' (basis-value storage and contraction precision).
    """

    def __init__(self, accuracy=DFTGridAccuracy.Medium, inCore=True, precision="fp64"):
        self.accuracy = DFTGridAccuracy(accuracy)
        self.inCore = inCore
        self.precision = precision

    def BuildGrid(self, target):
        """Build the grid (DFTGridIntegrator.py:25-29)."""
        grid = QuadratureGrid(target.qcState.atomicNumbers, target.scratch.qcCoordinates3AU, self.accuracy)
        target.scratch.dftGrid = grid
        target.scratch.qcEnergyReport["Quadrature Points"] = grid.number_of_points

    def EstimateMemory(self, qcModel, qcState):
        """Device bytes for this job (DFTGridIntegrator.py:31-39; sized for this library's layout and an unscreened basis)."""
        p = estimated_points(self.accuracy, qcState.atomicNumbers)
        n = len(qcState.orbitalBases)
        return estimate_bytes(p, n, float(n), self.precision)

    def Fock(self, target):
        """Add V_xc into the Fock matrix (or matrices) and return E_xc (DFTGridIntegrator.py:41-70).  Unrestricted: the
        total and spin densities (onePDMP, onePDMQ) become alpha = (T + S)/2 and beta = (T - S)/2, and the total and spin
        Fock matrices receive (F_a + F_b)/2 and (F_a - F_b)/2 (DFTIntegrator.pyx:82-107)."""
        scratch = target.scratch
        grid = scratch.dftGrid
        bases = target.qcState.orbitalBases
        if grid.number_of_functions == 0 or grid.precision != self.precision:
            grid.set_basis(bases.spec, bases.spherical, self.precision)
        ids = target.qcModel.functionalModel.ids
        if hasattr(scratch, "onePDMQ") and not target.electronicState.isSpinRestricted:
            T, S = scratch.onePDMP.density, scratch.onePDMQ.density
            Fa, Fb = np.zeros_like(T), np.zeros_like(T)
            eQuad, rhoQuad, stats = grid.integrate_uks(ids, 0.5 * (T + S), -0.5 * (S - T), Fa, Fb)
            fT, fS = scratch.onePDMP.fock, scratch.onePDMQ.fock
            fT += 0.5 * Fa
            fT += 0.5 * Fb
            fS += 0.5 * Fa
            fS += -0.5 * Fb
        elif not target.electronicState.isSpinRestricted:
            raise QCModelError("Error in DFT integrator (an unrestricted state needs the spin density onePDMQ).")
        else:
            eQuad, rhoQuad, stats = grid.integrate(ids, scratch.onePDMP.density, scratch.onePDMP.fock)
        scratch.qcEnergyReport["Quadrature Electron Count"] = (rhoQuad, "{:.3f}")
        scratch.qcEnergyReport["Quadrature Energy"] = eQuad
        n = stats["bf_values"]
        o = float(len(bases))
        scratch.qcEnergyReport["Quadrature BF Values"] = n
        scratch.qcEnergyReport["Quadrature BF Sparsity (%)"] = ((1.0 - n / (grid.number_of_points * o)) * 100.0, "{:.1f}")
        scratch.qcEnergyReport["Quadrature BF Storage (GB)"] = (grid.device_bytes() / 1.0e9, "{:.3f}")
        return eQuad

    def Gradients(self, target):
        """Add the XC contributions to the nuclear gradients (DFTGridIntegrator.py:72-88): scratch.qcGradients3AU
        (n_atoms x 3, float64) is incremented."""
        scratch = target.scratch
        if getattr(scratch, "doGradients", False):
            grid = scratch.dftGrid
            bases = target.qcState.orbitalBases
            if grid.number_of_functions == 0:
                grid.set_basis(bases.spec, bases.spherical, self.precision)
            ids = target.qcModel.functionalModel.ids
            if target.electronicState.isSpinRestricted:
                grid.gradients(ids, scratch.onePDMP.density, scratch.qcGradients3AU)
            elif hasattr(scratch, "onePDMQ"):
                T, S = scratch.onePDMP.density, scratch.onePDMQ.density
                grid.gradients_uks(ids, 0.5 * (T + S), -0.5 * (S - T), scratch.qcGradients3AU)
            else:
                raise QCModelError("Error in DFT integrator (an unrestricted state needs the spin density onePDMQ).")

    def SummaryItems(self):
        """Summary items (DFTGridIntegrator.py:90-93)."""
        return [("Grid Accuracy", "{:s}".format(self.accuracy.name)),
                ("Save Grid Function Data", "{:s}".format(repr(self.inCore))),
                ("Quadrature Precision", self.precision)]
