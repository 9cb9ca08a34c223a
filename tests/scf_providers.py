"""CPU providers for the SCF driver's tests (test infrastructure): the compiled reference's integrals (oracle/_ref) and the
oracle's quadrature stand in for the GPU integrals and grid, so that the SCF loop itself is checked on the CPU."""

import numpy as np

import oracle


class ReferenceIntegrals:
    """One-electron matrices, electron-fit integrals and the fit matrix from the compiled reference."""

    def __init__(self, Z, xyz, spec, spherical, fit_spec):
        self.orb = oracle.RefBasis(Z, xyz, spec, spherical)
        self.fitb = oracle.RefBasis(Z, xyz, fit_spec, True)
        self.n, self.n_fit = self.orb.n, self.fitb.n
        self.D, _ = self.orb.fit_integrals(self.fitb)
        n = self.n
        self.w = np.full(n * (n + 1) // 2, 2.0)
        self.w[np.arange(n) * (np.arange(n) + 3) // 2] = 1.0

    def one_electron(self):
        return self.orb.one_electron()

    def fit_matrix(self):
        return self.fitb.fit_matrix()

    def fit_b(self, P):
        return (self.w * P) @ self.D

    def fit_fock(self, c, F):
        F += self.D @ c


class OracleGrid:
    """The oracle's FP64 quadrature with the QuadratureGrid.integrate interface."""

    def __init__(self, Z, xyz, spec, spherical, ids, accuracy=2):
        self.h = oracle.oracle().setup(accuracy, Z, xyz, spec, spherical, ids)

    def integrate(self, ids, P, V):
        e, n, v, _ = self.h.integrate(P, 8)
        V += v
        return e, n, {}
