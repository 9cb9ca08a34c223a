"""The precision claims of the mixed-precision path, reproduced on the CPU (tests/precision_emulator.py) on (H2O)4 PBE
(6 A cube, seed 5, spherical 6-31G*) at its whole-cluster extended-Hueckel density:

- with no operand rounded the emulator is the oracle's quadrature (E_xc and N_e to 1e-11);
- Phi rounded to 4 base-128 digits per point lowers rho (N_e falls, E_xc rises); 5 digits shrink that bias more than
  30-fold;
- grad Phi as 24-bit fixed point raises E_xc through sigma and leaves N_e unchanged; 32 bits shrink the shift more than
  100-fold;
- dropping the digit orders above 4 (5 P digits x 5 Phi digits) lowers N_e;
- the arithmetic as built (P 5 digits per row, Phi 5 digits per point, orders <= 4, 32-bit grad Phi) stays within
  1e-6 / 128 Eh per water of FP64, so that a (H2O)128 cluster stays within the 1e-6 Eh bar.
"""

import numpy as np
import pytest

import oracle
from helpers import cluster_system
from precision_emulator import Emulator

PBE = [101, 130]


@pytest.fixture(scope="module")
def case():
    if not oracle.reference_available():
        pytest.skip("needs the compiled reference (make -C oracle ref) for the overlap matrix")
    from pdynamo3_amd import densities
    Z, xyz, spec, _ = cluster_system(4, 6.0, 5)
    S, _, _ = oracle.RefBasis(Z, xyz, spec, True).one_electron()
    P = densities.eht_density(Z, S, spherical=True)
    h = oracle.oracle().setup(2, Z, xyz, spec, True, PBE)
    e0, n0, _, _ = h.integrate(P, 4)
    return Emulator(h, oracle.oracle(), PBE), P, e0, n0


def test_unrounded_emulator_is_the_oracle(case):
    em, P, e0, n0 = case
    e, n = em.quadrature(P)
    assert abs(e - e0) < 1e-11 and abs(n - n0) < 1e-11


def test_rounding_biases_have_the_stated_signs_and_scales(case):
    em, P, _, _ = case
    e0, n0 = em.quadrature(P)      # the emulator's own FP64 result: the differences below are the roundings alone
    rows = {}
    for name, kw in (("phi 4 digits", dict(phi_digits=4)), ("phi 5 digits", dict(phi_digits=5)),
                     ("P 5 digits", dict(p_digits=5)), ("grad 24 bit", dict(grad_bits=23)),
                     ("grad 32 bit", dict(grad_bits=31)),
                     ("orders <= 4 (5 x 5 digits)", dict(phi_digits=5, p_digits=5, kept_orders=4)),
                     ("as built", dict(phi_digits=5, p_digits=5, kept_orders=4, grad_bits=31))):
        e, n = em.quadrature(P, **kw)
        rows[name] = (e - e0, n - n0)
        print("{:28s} dE_xc {:+.2e}  dN_e {:+.2e}".format(name, e - e0, n - n0))
    dE4, dN4 = rows["phi 4 digits"]
    dE5, dN5 = rows["phi 5 digits"]
    assert dN4 < 0.0 and dE4 > 0.0
    assert abs(dE5) < abs(dE4) / 30.0 and abs(dN5) < abs(dN4) / 30.0
    dEg24, dNg24 = rows["grad 24 bit"]
    dEg32, dNg32 = rows["grad 32 bit"]
    assert dEg24 > 0.0 and dNg24 == 0.0 and dNg32 == 0.0
    assert abs(dEg32) < abs(dEg24) / 100.0
    assert rows["orders <= 4 (5 x 5 digits)"][1] < rows["phi 5 digits"][1] + rows["P 5 digits"][1]
    dE, dN = rows["as built"]
    assert abs(dE) <= 4 * 1e-6 / 128 and abs(dN) <= 4 * 1e-6 / 128
