"""Writes tests/golden/reference_scf_h2o16.npz: a converged SCF density of the (H2O)16 PBE cluster and the compiled
reference's XC quadrature at it (test infrastructure; needs oracle/_ref/libxcref.so and oracle/liboracle.so).

The cluster is the benchmark one (16 waters, 9 A cube, seed 1, 6-31G* Cartesian, PBE, Medium grid).  The SCF is the
package's restricted Kohn-Sham loop (pdynamo3_amd/scf.py, the steps of DIISSCFConverger.py:168-543) driven on the CPU by
the compiled reference's integrals (one-electron matrices, electron-fit integrals and fit matrix, dgauss-a1-dftjfit) and
the oracle's FP64 quadrature, converged to 1e-10 Eh and an RMS density change of 1e-9.  Its density couples every pair
of waters, unlike the superposition of monomer densities.  Stored arrays:

  P          packed converged density (46,360 values)
  V          packed V_xc of the compiled reference's DFTIntegrator_Integrate at P
  E_xc, N_e  its XC energy and electron count
  E_scf      the converged total energy, and iterations, the SCF's iteration count

Run: make -C oracle && make -C oracle ref && python tests/golden/make_scf_golden.py
"""

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(HERE))

import oracle  # noqa: E402
import scf_providers as sp  # noqa: E402
from pdynamo3_amd import basis, molecules, scf  # noqa: E402

PBE = [101, 130]


def main():
    Z, xa, _, _ = molecules.water_cluster(16, 9.0, 1)
    xyz = molecules.to_bohr(xa)
    spec = basis.element_spec(basis.load_basis(), Z)
    fspec = basis.element_spec(basis.load_basis("dgauss-a1-dftjfit"), Z)
    t0 = time.perf_counter()
    ints = sp.ReferenceIntegrals(Z, xyz, spec, False, fspec)
    print("reference integrals {:.1f} s".format(time.perf_counter() - t0), flush=True)
    grid = sp.OracleGrid(Z, xyz, spec, False, PBE)
    s = scf.RKSDFT(Z, xyz, functional="pbe", integrals=ints, grid=grid, energy_tolerance=1e-10, density_tolerance=1e-9)
    r = s.run(log=lambda m: print(m, flush=True))
    if not r["converged"]:
        raise RuntimeError("SCF did not converge")
    P = s.P.copy()
    h = oracle.reference().setup(2, Z, xyz, spec, False, PBE)
    e, ne, V, sec = h.integrate(P, 0)
    print("reference quadrature at the converged density: E_xc {:.12f} N_e {:.10f} ({:.1f} s)".format(e, ne, sec))
    np.savez_compressed(os.path.join(HERE, "reference_scf_h2o16.npz"), P=P, V=V, E_xc=np.float64(e), N_e=np.float64(ne),
                        E_scf=np.float64(r["energy"]), iterations=np.int64(r["iterations"]))


if __name__ == "__main__":
    main()
