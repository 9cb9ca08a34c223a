"""The oracle against mathematics the reference does not hold as data: closed forms, limits, invariants and brute force.

Each check is chosen so that a plausible slip (a dropped term, a wrong sign, index or factor) fails it.
"""

import math

import numpy as np
import pytest

import oracle
from helpers import packed, unpacked, water_system
from pdynamo3_amd import densities

RULES = [6, 14, 26, 38, 50, 74, 86, 110, 146, 170, 194, 230, 266, 302, 350, 434, 590, 770, 974, 1202]
DEGREE = {6: 3, 14: 5, 26: 7, 38: 9, 50: 11, 74: 13, 86: 15, 110: 17, 146: 19, 170: 21, 194: 23, 230: 25, 266: 27,
          302: 29, 350: 31, 434: 35, 590: 41, 770: 47, 974: 53, 1202: 59}


def sphere_monomial(a, b, c):
    """(1/4pi) int x^a y^b z^c dOmega = (a-1)!!(b-1)!!(c-1)!!/(a+b+c+1)!! for even a, b, c, else 0."""
    if a % 2 or b % 2 or c % 2:
        return 0.0
    df = lambda n: math.prod(range(n, 0, -2)) if n > 0 else 1
    return df(a - 1) * df(b - 1) * df(c - 1) / df(a + b + c + 1)


@pytest.mark.parametrize("n", RULES)
def test_lebedev_rule_is_exact(n):
    u, w = oracle.lebedev_points(n)
    assert np.max(np.abs(np.sum(u * u, axis=1) - 1.0)) < 1e-14
    assert abs(np.sum(w) - 1.0) < 1e-13
    deg = DEGREE[n]
    worst = 0.0
    for a in range(0, deg + 1):
        for b in range(0, deg + 1 - a):
            for c in range(0, min(deg + 1 - a - b, 8)):
                q = np.sum(w * u[:, 0] ** a * u[:, 1] ** b * u[:, 2] ** c)
                worst = max(worst, abs(q - sphere_monomial(a, b, c)))
    assert worst < 1e-13


def test_radial_and_angular_quadrature_of_a_gaussian(oracle_lib):
    """A single atom's grid integrates exp(-a r^2) to (pi/a)^(3/2) (radial rule and angular rule together)."""
    Z = np.array([8], dtype=np.int32)
    xyz = np.zeros((1, 3))
    from pdynamo3_amd import basis
    spec = basis.element_spec(basis.load_basis(), [8])
    h = oracle_lib.setup(2, Z, xyz, spec, True, [106, 131])
    p, _, _ = h.grid()
    r2 = np.sum(p[:, :3] ** 2, axis=1)
    for a in (0.3, 1.0, 5.0, 40.0):
        assert np.sum(p[:, 3] * np.exp(-a * r2)) == pytest.approx((math.pi / a) ** 1.5, rel=2e-8)
    # A lone atom's Becke weights are all 1: no point is culled (O: 13,450 points on the Medium grid).  Some weights are
    # negative because some Lebedev rules (74, 230, 266 points) carry negative weights.
    assert h.npoints == 13450
    assert np.any(p[:, 3] < 0)


def test_partition_of_unity_via_electron_count(oracle_lib):
    """sum_p w_p rho_p = tr(P S) = 10 for the water density (S by the same quadrature in the fixture generator is not
    used here: the count is set by the Becke partition summing to one over atoms)."""
    Z, xyz, spec, n = water_system(True)
    P1, meta = densities.load_monomer_density(spherical=True)
    h = oracle_lib.setup(2, Z, xyz, spec, True, [106, 131])
    _, ne, _, _ = h.integrate(packed(P1))
    assert ne == pytest.approx(10.0, abs=2e-6)


@pytest.mark.parametrize("spherical", [True, False])
def test_basis_functions_are_normalised(oracle_lib, spherical):
    """<phi|phi> = 1 for every contracted function (primitive unnormalisation and the Cartesian factors; the 6-31G*
    contractions are normalised)."""
    Z, xyz, spec, n = water_system(spherical)
    h = oracle_lib.setup(2, Z, xyz, spec, spherical, [106, 131])
    p, _, _ = h.grid()
    f, _, _, _ = oracle.basis_at(h, p[:, :3])
    norms = np.sum(f * f * p[:, 3], axis=1)
    assert np.max(np.abs(norms - 1.0)) < 5e-6


def test_basis_gradients_are_derivatives(oracle_lib):
    Z, xyz, spec, n = water_system(True)
    h = oracle_lib.setup(2, Z, xyz, spec, True, [106, 131])
    rng = np.random.default_rng(4)
    pts = xyz[0] + rng.normal(size=(20, 3)) * 1.2
    f, fx, fy, fz = oracle.basis_at(h, pts)
    d = 1e-5
    for k, g in enumerate((fx, fy, fz)):
        e = np.zeros(3)
        e[k] = d
        fp = oracle.basis_at(h, pts + e)[0]
        fm = oracle.basis_at(h, pts - e)[0]
        assert np.max(np.abs((fp - fm) / (2 * d) - g)) < 1e-7 * max(1.0, np.max(np.abs(g)))


@pytest.mark.parametrize("ids", [[106], [101], [131], [130]])
def test_potentials_are_derivatives_of_the_energy_density(oracle_lib, ids):
    """vrho = d(rho eps)/d rho and vsigma = d(rho eps)/d sigma, central differences."""
    rho = np.array([1e-3, 0.02, 0.3, 2.0, 40.0])
    sigma = np.array([1e-7, 1e-4, 0.05, 3.0, 800.0])
    e, vr, vs = oracle_lib.functional(ids, rho, sigma)
    h = 1e-6
    ep, _, _ = oracle_lib.functional(ids, rho * (1 + h), sigma)
    em, _, _ = oracle_lib.functional(ids, rho * (1 - h), sigma)
    fd = ((rho * (1 + h)) * ep - (rho * (1 - h)) * em) / (2 * h * rho)
    assert np.allclose(vr, fd, rtol=2e-6, atol=1e-10)
    ep, _, _ = oracle_lib.functional(ids, rho, sigma * (1 + h))
    em, _, _ = oracle_lib.functional(ids, rho, sigma * (1 - h))
    fd = rho * (ep - em) / (2 * h * sigma)
    assert np.allclose(vs, fd, rtol=2e-5, atol=1e-12)


def test_exchange_reduces_to_dirac_at_zero_gradient(oracle_lib):
    rho = np.array([1e-4, 0.01, 1.0, 100.0])
    dirac = -0.75 * (3.0 / math.pi) ** (1.0 / 3.0) * rho ** (1.0 / 3.0)
    for ids in ([106], [101]):
        e, vr, _ = oracle_lib.functional(ids, rho, np.zeros(4))
        assert np.allclose(e, dirac, rtol=1e-12)
        assert np.allclose(vr, 4.0 / 3.0 * dirac, rtol=1e-12)


def test_pbe_correlation_reduces_to_pw92_at_zero_gradient(oracle_lib):
    """PBE correlation with t = 0 is the PW92 correlation energy (Perdew & Wang 1992, eq. 10, PBE constants)."""
    rho = np.array([1e-4, 0.01, 1.0, 100.0])
    rs = (3.0 / (4.0 * math.pi * rho)) ** (1.0 / 3.0)
    a, alpha, b1, b2, b3, b4 = 0.0310907, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294
    G = -2 * a * (1 + alpha * rs) * np.log(1 + 1 / (2 * a * (b1 * rs ** 0.5 + b2 * rs + b3 * rs ** 1.5 + b4 * rs ** 2)))
    e, _, _ = oracle_lib.functional([130], rho, np.zeros(4))
    assert np.allclose(e, G, rtol=1e-9)


def test_vxc_is_the_derivative_of_exc_with_respect_to_p(oracle_lib):
    """dE_xc/dP_mn = (2 - delta_mn) V_mn for the symmetric P (finite differences on the water density)."""
    Z, xyz, spec, n = water_system(True)
    P1, _ = densities.load_monomer_density(spherical=True)
    h = oracle_lib.setup(2, Z, xyz, spec, True, [106, 131])
    _, _, V, _ = h.integrate(packed(P1))
    Vd = unpacked(V, n)
    d = 1e-6
    for (m, k) in ((0, 0), (3, 1), (7, 7), (14, 2), (17, 9)):
        D = np.zeros((n, n))
        D[m, k] = D[k, m] = d
        ep, _, _, _ = h.integrate(packed(P1 + D))
        em, _, _, _ = h.integrate(packed(P1 - D))
        fd = (ep - em) / (2 * d)
        expect = Vd[m, k] * (1.0 if m == k else 2.0)
        assert fd == pytest.approx(expect, rel=1e-5, abs=1e-8), (m, k)
