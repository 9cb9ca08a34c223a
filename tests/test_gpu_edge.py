"""Edge cases of the quadrature on the GPU, checked against the oracle or against exact answers.

- A zero density: every block is skipped by the density tolerance, E_xc = N_e = 0 and V is left exactly as it was.
- One atom (O, and H with two functions): fewer kept functions than one chunk of 32 rows, a single superblock per run
  of blocks.
- Two waters 12 Angstrom apart: blocks far from the other molecule keep few functions, and the union rows of a
  superblock mix blocks with very different kept sets.
- A handle reused with different densities gives what a fresh handle gives (no state leaks between calls).
Tolerances as in test_gpu_parity.py: FP64 1e-9 (E, N, V); mixed 1e-6 (E, N) and 1e-7 (V).
"""

import numpy as np
import pytest

from helpers import ids, random_density

pytestmark = pytest.mark.gpu

TOL = {"fp64": (1e-9, 1e-9), "mixed": (1e-6, 1e-7)}


def _system(Z, xyz_bohr, spherical=False):
    from pdynamo3_amd import basis
    b = basis.load_basis("6-31g_st")
    Z = np.asarray(Z, dtype=np.int32)
    return Z, np.asarray(xyz_bohr, dtype=np.float64), basis.element_spec(b, Z), basis.number_of_functions(b, Z, spherical)


def _check_against_oracle(oracle_lib, Z, xyz, spec, n, precision, functional, seed):
    from pdynamo3_amd import integrator
    P = random_density(n, max(1, min(5, n // 2)), seed=seed)
    h = oracle_lib.setup(2, Z, xyz, spec, False, ids(functional))
    e0, n0, V0, _ = h.integrate(P)
    g = integrator.QuadratureGrid(Z, xyz, 2)
    g.set_basis(spec, False, precision)
    V = np.zeros_like(P)
    e, ne, _ = g.integrate(ids(functional), P, V)
    g.close()
    tolE, tolV = TOL[precision]
    print(precision, "dE", e - e0, "dN", ne - n0, "max dV", np.max(np.abs(V - V0)))
    assert abs(e - e0) <= tolE
    assert abs(ne - n0) <= tolE
    assert np.max(np.abs(V - V0)) <= tolV


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_zero_density_skips_every_block(precision):
    from pdynamo3_amd import integrator, molecules
    Z, xyz = molecules.water()
    Z, xyz, spec, n = _system(Z, xyz)
    g = integrator.QuadratureGrid(Z, xyz, 2)
    g.set_basis(spec, False, precision)
    P = np.zeros(n * (n + 1) // 2)
    V = np.linspace(-1.0, 1.0, P.size)
    V_in = V.copy()
    e, ne, stats = g.integrate(ids("blyp"), P, V)
    g.close()
    assert e == 0.0 and ne == 0.0
    assert np.array_equal(V, V_in)
    assert stats["blocks_skipped"] > 0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("element", [8, 1])
def test_single_atom_matches_oracle(oracle_lib, precision, element):
    Z, xyz, spec, n = _system([element], [[0.0, 0.0, 0.0]])
    _check_against_oracle(oracle_lib, Z, xyz, spec, n, precision, "pbe", seed=21 + element)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_separated_waters_match_oracle(oracle_lib, precision):
    from pdynamo3_amd import molecules
    Z1, x1 = molecules.water()
    shift = molecules.to_bohr(np.array([[12.0, 0.0, 0.0]]))[0]
    Z, xyz, spec, n = _system(np.concatenate([Z1, Z1]), np.vstack([x1, x1 + shift]))
    _check_against_oracle(oracle_lib, Z, xyz, spec, n, precision, "blyp", seed=31)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_handle_reuse_matches_fresh_handles(precision):
    from pdynamo3_amd import integrator
    from helpers import cluster_system
    Z, xyz, spec, n = cluster_system(3, 5.5, 9, False)
    P1, P2 = random_density(n, 12, seed=41), random_density(n, 12, seed=42)

    def fresh(P):
        g = integrator.QuadratureGrid(Z, xyz, 2)
        g.set_basis(spec, False, precision)
        V = np.zeros_like(P)
        e, ne, _ = g.integrate(ids("blyp"), P, V)
        g.close()
        return e, ne, V

    g = integrator.QuadratureGrid(Z, xyz, 2)
    g.set_basis(spec, False, precision)
    results = []
    for P in (P1, P2, P1):
        V = np.zeros_like(P)
        e, ne, _ = g.integrate(ids("blyp"), P, V)
        results.append((e, ne, V))
    g.close()
    for (e, ne, V), P in zip(results, (P1, P2, P1)):
        e0, n0, V0 = fresh(P)
        assert e == e0 and ne == n0
        assert np.max(np.abs(V - V0)) <= 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_out_of_memory_leaves_the_handle_usable(precision):
    """An allocation that fails inside a quadrature returns status 7 (OutOfMemory); the next call on the same handle
    allocates its work arrays again and gives the fresh handle's result."""
    from pdynamo3_amd import _lib, basis, integrator, molecules
    Z, xyz = molecules.water()
    spec = basis.element_spec(basis.load_basis(), Z)
    n = basis.number_of_functions(basis.load_basis(), Z, False)
    P = random_density(n, 5, seed=3)
    g0 = integrator.QuadratureGrid(Z, xyz, 2)
    g0.set_basis(spec, False, precision)
    V0 = np.zeros_like(P)
    e0, n0, _ = g0.integrate(ids("blyp"), P, V0)
    g0.close()
    failures = 0
    for k in range(8):            # the k-th allocation of the first call fails: each work array and the job table
        g = integrator.QuadratureGrid(Z, xyz, 2)
        g.set_basis(spec, False, precision)
        _lib.lib.xq_test_fail_allocations(k)
        try:
            g.integrate(ids("blyp"), P, np.zeros_like(P))
        except integrator.QCModelError as exc:
            assert "status 7" in str(exc)
            failures += 1
        finally:
            _lib.lib.xq_test_fail_allocations(-1)
        V = np.zeros_like(P)
        e, ne, _ = g.integrate(ids("blyp"), P, V)
        g.close()
        print(precision, "fail at", k, "then E", e, "dE", e - e0)
        assert e == e0 and ne == n0
        assert np.max(np.abs(V - V0)) <= 1e-12
    assert failures >= 6          # d_job, d_P, d_V, packed, block sums, skip flags


def test_device_bytes_follow_the_precision():
    """The handle's device byte count falls when the FP64 planes are released for the mixed ones and returns to the
    same value when FP64 is chosen again."""
    from pdynamo3_amd import basis, integrator, molecules
    Z, xa, _, _ = molecules.water_cluster(4, 9.0 * (4 / 16.0) ** (1.0 / 3.0), 1)
    xyz = molecules.to_bohr(xa)
    spec = basis.element_spec(basis.load_basis(), Z)
    n = basis.number_of_functions(basis.load_basis(), Z, False)
    P = random_density(n, 20, seed=4)
    g = integrator.QuadratureGrid(Z, xyz, 2)
    grid_only = g.device_bytes()
    g.set_basis(spec, False, "fp64")
    g.integrate(ids("pbe"), P)
    b64 = g.device_bytes()
    g.set_basis(spec, False, "mixed")
    g.integrate(ids("pbe"), P)
    bmix = g.device_bytes()
    g.set_basis(spec, False, "fp64")
    g.integrate(ids("pbe"), P)
    b64again = g.device_bytes()
    g.close()
    print("grid", grid_only, "fp64", b64, "mixed", bmix, "fp64 again", b64again)
    assert grid_only > 0
    assert bmix < b64
    assert b64again == b64


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_density_product_statistics(precision):
    """stats[9] (the density product kernel's device time) lies inside the density stage's time (stats[3]); stats[10]
    (its algorithmic bytes) is the stage's figure on the FP64 path (one kernel) and, on the mixed path, 17 bytes per
    kept value, P digits and row exponents, 4 bytes per point and a whole number of 4 KB row-chunk shares, at least
    one per 32 x 128 kept values."""
    from pdynamo3_amd import basis, integrator, molecules
    Z, xa, _, _ = molecules.water_cluster(4, 9.0 * (4 / 16.0) ** (1.0 / 3.0), 1)
    xyz = molecules.to_bohr(xa)
    spec = basis.element_spec(basis.load_basis(), Z)
    n = basis.number_of_functions(basis.load_basis(), Z, False)
    P = random_density(n, 20, seed=5)
    g = integrator.QuadratureGrid(Z, xyz, 2)
    g.set_basis(spec, False, precision)
    _, _, st = g.integrate(ids("pbe"), P)
    G = g.number_of_points
    g.close()
    assert 0.0 < st["product_ms"] <= st["rho_ms"]
    if precision == "fp64":
        assert st["product_bytes"] == st["rho_bytes"]
    else:
        shares = st["product_bytes"] - (17.0 * st["bf_values"] + 8.0 * n * n + 4.0 * n + 4.0 * G)
        chunks = shares / 4096.0
        assert chunks == round(chunks)
        assert chunks >= st["bf_values"] / (32.0 * 128.0)
