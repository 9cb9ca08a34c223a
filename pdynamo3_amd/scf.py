"""The SCF wrapper around the XC quadrature: restricted closed-shell Kohn-Sham DFT with a density-fitted Coulomb term, as
pDynamo3's QCModelDFT runs it with a fit basis (pMolecule/QCModel/QCModelDFT.py:130-252) under its DIIS SCF converger
(pMolecule/QCModel/DIISSCFConverger.py:168-543).

Per geometry (GPU, include/qcscf.h): overlap S, kinetic T and electron-nuclear V of Gaussian nuclei (f1KOf1i, f2Cm1V);
the electron-fit integrals (mu nu|P), kept in device memory (f1Xg2i); the constrained fit matrix (f1Xf1i_f1Oi), factorised
once on the host; the DFT grid and basis values (the quadrature library); the symmetric orthogonaliser X = S^-1/2
(QCModelBase.GetOrthogonalizer, QCModelBase.py:313-342, OrthogonalizationMethod.Symmetric).

Per SCF iteration, at density P (packed, total):
  F = H + J(P) + V_xc(P),   H = T + V                       (FockOne, QCModelBase.py:237-249)
  J: b_P = sum_mn P_mn (mn|P), [G q; q^T 0] [c; lambda] = [b; N], F_mn += sum_P c_P (mn|P), E_fit = (c.b + lambda N) / 2
                                                            (Fock_MakeFromFitIntegralsCoulomb, FockConstruction.c:31-108)
  E = Tr(P H) + E_fit + E_xc + E_nuclear
The first density is the diagonal guess P_ii = N / n (QCOnePDM.FromDiagonalGuess, QCOnePDM.py:52-67).  The converger
takes optimal-damping (ODA) steps while the commutator error X^T (F P S - S P F) X is above the DIIS onset, then DIIS
extrapolation over the last `history` Fock matrices (DIISSCFConverger.py: diisOnset 0.2, maximumHistory 10, useODA).
It stops when the energy changes by less than energy_tolerance and the RMS density change is below density_tolerance.

The SCF is a state machine so that a request loop can batch the XC quadratures of many jobs: `density()` is the density
whose Fock matrix is needed next, `advance(E_xc, V_xc)` takes that density's XC energy and matrix and moves on.  `run()`
drives it alone.
"""

import ctypes
import time

import numpy as np
import scipy.linalg

from . import _lib
from .basis import element_spec, load_basis, number_of_functions
from .definitions import functional_ids
from .integrator import QCModelError, QuadratureGrid


def _check(status, what):
    if status != 0:
        raise QCModelError("Error in SCF integrals ({:s}: status {:d}, {:s}).".format(what, status, _lib.last_error()))


_TRIL = {}


def _tril(n):
    """Row and column indices of the packed lower triangle of an n x n matrix and the positions of its diagonal in the
    packed vector (computed once per n: SCF steps and serving jobs reuse them)."""
    t = _TRIL.get(n)
    if t is None:
        i, j = np.tril_indices(n)
        t = (i, j, np.arange(n) * (np.arange(n) + 3) // 2)
        _TRIL[n] = t
    return t


def unpack(v, n):
    """Dense symmetric matrix from a packed lower triangle."""
    i, j, _ = _tril(n)
    M = np.empty((n, n))
    M[i, j] = v
    M[j, i] = v
    return M


def pack(M):
    i, j, _ = _tril(M.shape[0])
    return np.ascontiguousarray(M[i, j])


def trace_product(a, b, n):
    """Tr(A B) of two symmetric matrices given packed (off-diagonal terms twice)."""
    d = _tril(n)[2]
    return float(2.0 * np.dot(a, b) - np.dot(a[d], b[d]))


def nuclear_repulsion(Z, xyz):
    Z = np.asarray(Z, dtype=np.float64)
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    e = 0.0
    for i in range(1, len(Z)):
        r = np.sqrt(np.sum((xyz[:i] - xyz[i]) ** 2, axis=1))
        e += float(np.sum(Z[i] * Z[:i] / r))
    return e


class GPUIntegrals:
    """The integrals of one molecule on the GPU: one-electron matrices and the resident electron-fit integrals (none
    when fit_spec is None: one-electron matrices only)."""

    def __init__(self, Z, xyz_bohr, orbital_spec, spherical, fit_spec, fit_spherical=True):
        self.Z = _lib.i32(Z)
        self.xyz = _lib.f64(xyz_bohr).reshape(-1, 3)
        self._orb, self._fitb, self._fit = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        L = _lib.lib
        o = _lib.f64(orbital_spec)
        _check(L.qs_basis_create(self.Z.size, _lib.iptr(self.Z), _lib.dptr(self.xyz), o.size, _lib.dptr(o),
                                 int(bool(spherical)), ctypes.byref(self._orb)), "orbital basis")
        n = ctypes.c_int()
        _check(L.qs_basis_size(self._orb, ctypes.byref(n)), "basis size")
        self.n, self.n_fit = n.value, 0
        if fit_spec is None:
            return
        f = _lib.f64(fit_spec)
        _check(L.qs_basis_create(self.Z.size, _lib.iptr(self.Z), _lib.dptr(self.xyz), f.size, _lib.dptr(f),
                                 int(bool(fit_spherical)), ctypes.byref(self._fitb)), "fit basis")
        nf = ctypes.c_int()
        _check(L.qs_basis_size(self._fitb, ctypes.byref(nf)), "basis size")
        self.n_fit = nf.value
        _check(L.qs_fit_create(self._orb, self._fitb, ctypes.byref(self._fit)), "fit integrals")

    def one_electron(self):
        m = self.n * (self.n + 1) // 2
        S, T, V = np.zeros(m), np.zeros(m), np.zeros(m)
        _check(_lib.lib.qs_one_electron(self._orb, _lib.dptr(S), _lib.dptr(T), _lib.dptr(V)), "one-electron")
        return S, T, V

    def fit_matrix(self):
        M = np.zeros((self.n_fit + 1) * (self.n_fit + 2) // 2)
        _check(_lib.lib.qs_fit_matrix(self._fit, _lib.dptr(M)), "fit matrix")
        return M

    def fit_b(self, P):
        b = np.zeros(self.n_fit)
        _check(_lib.lib.qs_fit_b(self._fit, _lib.dptr(_lib.f64(P)), _lib.dptr(b)), "fit b")
        return b

    def fit_fock(self, c, F):
        _check(_lib.lib.qs_fit_fock(self._fit, _lib.dptr(_lib.f64(c)), _lib.dptr(F)), "fit Fock")

    def stats(self):
        r, p, b = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        _check(_lib.lib.qs_fit_stats(self._fit, ctypes.byref(r), ctypes.byref(p), ctypes.byref(b)), "fit stats")
        return {"fit_rows": r.value, "fit_shell_pairs": p.value, "fit_device_bytes": b.value}

    def close(self):
        if self._fit:
            _lib.lib.qs_fit_free(self._fit)
            self._fit = ctypes.c_void_p()
        for h in ("_orb", "_fitb"):
            if getattr(self, h):
                _lib.lib.qs_basis_free(getattr(self, h))
                setattr(self, h, ctypes.c_void_p())

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceDenseSteps:
    """The SCF's dense steps on the GPU (include/qcscf.h, qs_dense_*: rocBLAS and rocSOLVER): the orbitals and density
    of a Fock matrix and the DIIS error vector, with the orthogonaliser X and the overlap S resident in device memory."""

    def __init__(self, X, S_packed):
        self.n = int(X.shape[0])
        self._h = ctypes.c_void_p()
        X = np.ascontiguousarray(X, dtype=np.float64)
        _check(_lib.lib.qs_dense_create(self.n, _lib.dptr(X), _lib.dptr(_lib.f64(S_packed)), ctypes.byref(self._h)),
               "dense SCF steps")

    def aufbau(self, F, n_occupied):
        """Packed P = 2 C_occ C_occ^T of the packed Fock matrix F, and the orbital energies."""
        P, e = np.empty(self.n * (self.n + 1) // 2), np.empty(self.n)
        _check(_lib.lib.qs_dense_aufbau(self._h, _lib.dptr(_lib.f64(F)), int(n_occupied), _lib.dptr(P), _lib.dptr(e)),
               "aufbau")
        return P, e

    def commutator(self, F, P):
        """X^T (F P S - S P F) X (dense) of packed F and P."""
        err = np.empty((self.n, self.n))
        _check(_lib.lib.qs_dense_commutator(self._h, _lib.dptr(_lib.f64(F)), _lib.dptr(_lib.f64(P)), _lib.dptr(err)),
               "commutator")
        return err

    def close(self):
        if self._h:
            _lib.lib.qs_dense_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# Basis size from which the dense steps run on the GPU by default: below it one host core is as fast (measured on
# MI355X: eigenproblems of 304 functions 4.6 ms on one core against 6.9 ms on the device; of 1216 functions 126 ms
# against 29 ms, and a 1216^3 product 29 ms against 0.09 ms).
DEVICE_DENSE_MIN_FUNCTIONS = 512


class RKSDFT:
    """Restricted Kohn-Sham SCF with a fit basis for J; the XC quadrature on the GPU.

    Z, xyz_bohr      atoms
    basis            orbital basis label (6-31g_st, def2-svp) and spherical / Cartesian functions
    functional       name (blyp, pbe, lda, b3lyp's XC part is refused: it needs exact exchange)
    fit_basis        the J fitting basis (dgauss-a1-dftjfit, spherical functions as in its parameter files)
    precision        'fp64' or 'mixed' quadrature
    integrals, grid  injected providers (tests); by default the GPU integrals and a QuadratureGrid
    dense            where the orbitals, the density and the DIIS error are formed: 'host' (numpy / LAPACK), 'device'
                     (DeviceDenseSteps), or 'auto': the device from DEVICE_DENSE_MIN_FUNCTIONS functions when the
                     integrals are on the GPU
    """

    def __init__(self, Z, xyz_bohr, basis="6-31g_st", spherical=False, functional="blyp",
                 fit_basis="dgauss-a1-dftjfit", precision="fp64", accuracy=2, charge=0, energy_tolerance=1.0e-6,
                 density_tolerance=1.0e-6, max_iterations=100, diis_onset=0.2, history=10, integrals=None, grid=None,
                 guess=None, dense="auto"):
        if dense not in ("auto", "host", "device"):
            raise ValueError("dense must be 'auto', 'host' or 'device'")
        self.dense_mode = dense
        self.dense = None
        self.Z = _lib.i32(Z)
        self.xyz = _lib.f64(xyz_bohr).reshape(-1, 3)
        self.functional = functional
        self.ids = functional_ids(functional)
        if 402 in self.ids:
            raise QCModelError("Hybrid functionals need exact exchange, which this SCF does not build.")
        self.precision = precision
        self.accuracy = accuracy
        self.spherical = bool(spherical)
        self.energy_tolerance = energy_tolerance
        self.density_tolerance = density_tolerance
        self.max_iterations = max_iterations
        self.diis_onset = diis_onset
        self.history = history
        b = load_basis(basis)
        self.spec = element_spec(b, self.Z)
        self.n = number_of_functions(b, self.Z, self.spherical)
        self.n_electrons = float(np.sum(self.Z)) - charge
        if int(round(self.n_electrons)) % 2 != 0:
            raise QCModelError("A restricted closed-shell SCF needs an even number of electrons.")
        self.n_occupied = int(round(self.n_electrons)) // 2
        self.fit_spec = element_spec(load_basis(fit_basis), self.Z)
        self.integrals = integrals if integrals is not None else GPUIntegrals(self.Z, self.xyz, self.spec, self.spherical,
                                                                              self.fit_spec)
        self.grid = grid
        if self.grid is None:
            self.grid = QuadratureGrid(self.Z, self.xyz, accuracy)
            self.grid.set_basis(self.spec, self.spherical, precision)
        self._setup(guess)

    # ---------------------------------------------------------------------------------------------------------------
    def _setup(self, guess):
        self.timings = {"quadrature_s": 0.0, "coulomb_s": 0.0, "diagonalisation_s": 0.0, "diis_s": 0.0}
        S, T, V = self.integrals.one_electron()
        self.S_packed, self.H = S, T + V
        self.S = unpack(S, self.n)
        s, U = np.linalg.eigh(self.S)
        if s[0] <= 0.0:
            raise QCModelError("The overlap matrix is not positive definite.")
        self.X = (U / np.sqrt(s)) @ U.T            # S^-1/2 (symmetric orthogonalisation)
        if self.dense_mode == "device" or (self.dense_mode == "auto" and self.n >= DEVICE_DENSE_MIN_FUNCTIONS and
                                           isinstance(self.integrals, GPUIntegrals)):
            self.dense = DeviceDenseSteps(self.X, S)
        M = unpack(self.integrals.fit_matrix(), self.integrals.n_fit + 1)
        self.fit_lu = scipy.linalg.lu_factor(M)
        self.e_nuclear = nuclear_repulsion(self.Z, self.xyz)
        if guess is None:                          # diagonal guess
            P = np.zeros((self.n, self.n))
            np.fill_diagonal(P, self.n_electrons / self.n)
            guess = pack(P)
        self.P = _lib.f64(guess).copy()
        self.iteration = 0
        self.converged = False
        self.diis = False
        self.history_F, self.history_e = [], []
        self.diis_B = np.zeros((0, 0))             # inner products of the stored error vectors
        self.energies, self.records = [], []
        self.oda_F = self.oda_P = self.oda_E = None
        self.last_P = None
        self.last_energy = None

    def density(self):
        """The packed density whose Fock matrix is needed next."""
        return self.P

    def coulomb(self, P, F):
        """Add the fitted J(P) to F (packed); return E_fit."""
        t = time.perf_counter()
        b = self.integrals.fit_b(P)
        rhs = np.append(b, self.n_electrons)
        c = scipy.linalg.lu_solve(self.fit_lu, rhs)
        self.integrals.fit_fock(c[:-1], F)
        self.timings["coulomb_s"] += time.perf_counter() - t
        return 0.5 * float(np.dot(c, rhs))

    def fock(self, e_xc, v_xc):
        """F and the energy terms at the current density, given its XC energy and matrix."""
        P = self.P
        F = self.H.copy()
        e_fit = self.coulomb(P, F)
        F += v_xc
        e_one = trace_product(P, self.H, self.n)
        energy = e_one + e_fit + e_xc + self.e_nuclear
        return F, energy, {"one_electron": e_one, "coulomb_fit": e_fit, "xc": e_xc, "nuclear": self.e_nuclear}

    def _aufbau(self, F):
        """The density of the n_occupied lowest orbitals of F (in the orthonormal basis X)."""
        t = time.perf_counter()
        if self.dense is not None:
            P, self.orbital_energies = self.dense.aufbau(F, self.n_occupied)
        else:
            Fo = self.X.T @ unpack(F, self.n) @ self.X
            e, C = np.linalg.eigh(Fo)
            Cocc = self.X @ C[:, :self.n_occupied]
            self.orbital_energies = e
            P = pack(2.0 * Cocc @ Cocc.T)
        self.timings["diagonalisation_s"] += time.perf_counter() - t
        return P

    def commutator_error(self, F, P):
        if self.dense is not None:
            return self.dense.commutator(F, P)
        Fd, Pd = unpack(F, self.n), unpack(P, self.n)
        FPS = Fd @ Pd @ self.S
        return self.X.T @ (FPS - FPS.T) @ self.X

    def advance(self, e_xc, v_xc):
        """Take the current density's XC energy and matrix; build its Fock matrix, test convergence and choose the next
        density.  Returns True once the SCF has converged or stopped."""
        F, energy, parts = self.fock(e_xc, v_xc)
        t = time.perf_counter()
        err = self.commutator_error(F, self.P)
        self.timings["diis_s"] += time.perf_counter() - t
        err_max = float(np.max(np.abs(err)))
        dE = None if self.last_energy is None else energy - self.last_energy
        dP = None if self.last_P is None else float(np.sqrt(np.mean((self.P - self.last_P) ** 2)))
        self.iteration += 1
        self.energies.append(energy)
        self.records.append({"iteration": self.iteration, "energy": energy, "dE": dE, "rms_dP": dP,
                             "diis_error": err_max, "mode": "diis" if self.diis else "oda", **parts})
        self.parts, self.energy, self.fock_matrix = parts, energy, F
        if (self.diis and dE is not None and abs(dE) < self.energy_tolerance and dP is not None
                and dP < self.density_tolerance):
            self.converged = True
            return True
        if self.iteration >= self.max_iterations:
            return True
        self.last_energy, self.last_P = energy, self.P.copy()
        if not self.diis and err_max < self.diis_onset:
            self.diis = True
        if self.diis:
            t = time.perf_counter()
            # The new error vector's inner products with the stored ones extend the DIIS matrix by one row.
            flat = err.ravel()
            row = np.array([float(np.dot(flat, e.ravel())) for e in self.history_e] + [float(np.dot(flat, flat))])
            m = len(self.history_e)
            B = np.zeros((m + 1, m + 1))
            B[:m, :m] = self.diis_B
            B[m, :] = B[:, m] = row
            self.history_F.append(F)
            self.history_e.append(err)
            if len(self.history_e) > self.history:
                self.history_F, self.history_e, B = self.history_F[1:], self.history_e[1:], B[1:, 1:]
            self.diis_B = B
            self.timings["diis_s"] += time.perf_counter() - t
            t = time.perf_counter()
            F_next = self._diis_extrapolate()
            self.timings["diis_s"] += time.perf_counter() - t
            self.P = self._aufbau(F_next)
        else:
            self.P = self._oda_step(F, energy)
        return False

    def _oda_step(self, F, energy):
        """Optimal damping: the energy along the segment from the damped density to the new aufbau density is taken as
        the quadratic through both ends with the damped Fock matrix's slope; the damped density moves to its minimum."""
        if self.oda_P is None:
            Pd, Fd, Ed = self.P.copy(), F, energy
        else:
            P0, F0, E0 = self.oda_P, self.oda_F, self.oda_E
            dP = self.P - P0
            s = trace_product(F0, dP, self.n)
            c = energy - E0 - s
            lam = 1.0 if c <= -0.5 * s else -s / (2.0 * c)
            lam = min(1.0, max(0.0, lam))
            Pd = P0 + lam * dP
            Fd = F0 + lam * (F - F0)
            Ed = E0 + lam * s + lam * lam * c
        self.oda_P, self.oda_F, self.oda_E = Pd, Fd, Ed
        return self._aufbau(Fd)

    def _diis_extrapolate(self):
        m = len(self.history_F)
        if m == 1:
            return self.history_F[0]
        B = np.zeros((m + 1, m + 1))
        B[:m, :m] = self.diis_B
        B[m, :m] = B[:m, m] = -1.0
        rhs = np.zeros(m + 1)
        rhs[m] = -1.0
        try:
            w = np.linalg.solve(B, rhs)[:m]
        except np.linalg.LinAlgError:
            w = np.linalg.lstsq(B, rhs, rcond=None)[0][:m]
        return sum(wi * Fi for wi, Fi in zip(w, self.history_F))

    def xc(self, P):
        """The XC energy and matrix at P on the GPU (a batch of one)."""
        t = time.perf_counter()
        V = np.zeros_like(P)
        e, _, _ = self.grid.integrate(self.ids, P, V)
        self.timings["quadrature_s"] += time.perf_counter() - t
        return e, V

    def run(self, log=None):
        """Iterate to convergence; returns a summary."""
        while True:
            e, V = self.xc(self.P)
            done = self.advance(e, V)
            if log:
                r = self.records[-1]
                log("SCF {:3d} {:s} E {:.10f} dE {} err {:.2e}".format(r["iteration"], r["mode"], r["energy"],
                                                                       "-" if r["dE"] is None else "{:.2e}".format(r["dE"]),
                                                                       r["diis_error"]))
            if done:
                break
        return self.summary()

    def summary(self):
        return {"energy": self.energy, "converged": self.converged, "iterations": self.iteration, **self.parts,
                "timings": dict(self.timings)}

    def close(self):
        if self.dense is not None:
            self.dense.close()
            self.dense = None
        if isinstance(self.integrals, GPUIntegrals):
            self.integrals.close()
        if isinstance(self.grid, QuadratureGrid):
            self.grid.close()
