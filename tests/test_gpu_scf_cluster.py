"""The (H2O)16 PBE cluster (16 waters, 9 A cube, seed 1, 6-31G* Cartesian, Medium grid) at a converged SCF density, whose
density matrix couples every pair of waters (tests/golden/reference_scf_h2o16.npz, written by make_scf_golden.py from the
compiled reference's integrals and quadrature).

- The GPU quadrature at that density equals the compiled reference's entry by entry: FP64 to 1e-9 in E_xc and every V_xc
  entry, mixed precision to 1e-6 Eh and 1e-7.
- The GPU SCF (integrals and quadrature on the device, FP64) converges to the same total energy as the SCF driven by the
  compiled reference's integrals and the oracle's quadrature, within 1e-8 Eh; the mixed-precision SCF within 1e-6 Eh.
"""

import os

import numpy as np
import pytest

from helpers import cluster_system

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_scf_h2o16.npz")


def _golden():
    with np.load(GOLDEN) as d:
        return {k: d[k] for k in d.files}


@pytest.mark.parametrize("precision,tol_e,tol_v", [("fp64", 1e-9, 1e-9), ("mixed", 1e-6, 1e-7)])
def test_quadrature_at_scf_density_equals_reference(precision, tol_e, tol_v):
    from pdynamo3_amd import integrator
    ref = _golden()
    Z, xyz, spec, _ = cluster_system(16, 9.0, 1)
    g = integrator.QuadratureGrid(Z, xyz, 2)
    g.set_basis(spec, False, precision)
    V = np.zeros_like(ref["P"])
    E, N, _ = g.integrate([101, 130], ref["P"], V)
    g.close()
    dv = float(np.max(np.abs(V - ref["V"])))
    print(precision, "dE", E - float(ref["E_xc"]), "dN", N - float(ref["N_e"]), "max|dV|", dv)
    assert abs(E - float(ref["E_xc"])) <= tol_e
    assert abs(N - float(ref["N_e"])) <= tol_e
    assert dv <= tol_v


def test_gpu_scf_reaches_the_reference_driven_energy():
    from pdynamo3_amd import scf
    ref = _golden()
    Z, xyz, _, _ = cluster_system(16, 9.0, 1)
    out = {}
    for precision in ("fp64", "mixed"):
        s = scf.RKSDFT(Z, xyz, functional="pbe", precision=precision, energy_tolerance=1e-10, density_tolerance=1e-9)
        r = s.run()
        P = s.P.copy()
        s.close()
        out[precision] = r
        print(precision, r["energy"], "iterations", r["iterations"], "dE vs reference-driven",
              r["energy"] - float(ref["E_scf"]), "max|dP|", float(np.max(np.abs(P - ref["P"]))))
        assert r["converged"]
    assert abs(out["fp64"]["energy"] - float(ref["E_scf"])) < 1e-8
    assert abs(out["mixed"]["energy"] - float(ref["E_scf"])) < 1e-6
