"""Batched quadratures (xq_integrate_batch): every job's result is the same alone and in the mix.

E_xc and N_e are per-block FP64 sums added on the host in block order, so they must agree bit for bit; V_xc is
accumulated with FP64 atomics, whose order differs between launches, so it must agree to rounding (1e-12 absolute).
"""

import numpy as np
import pytest

from helpers import cluster_system, ids, packed, random_density, water_system

pytestmark = pytest.mark.gpu


def _jobs(precision):
    from pdynamo3_amd import densities, integrator, molecules
    from pdynamo3_amd import basis as B
    out = []
    Z, xyz, spec, n = water_system(False)
    g = integrator.QuadratureGrid(Z, xyz, 2)
    g.set_basis(spec, False, precision)
    P1, _ = densities.load_monomer_density("6-31g_st", spherical=False)
    out.append((g, ids("blyp"), packed(P1)))
    for n_w, cube, seed, fn in ((4, 6.0, 5, "pbe"), (8, 7.5, 6, "blyp")):
        Z, xa, rots, _ = molecules.water_cluster(n_w, cube, seed)
        g = integrator.QuadratureGrid(Z, molecules.to_bohr(xa), 2)
        g.set_basis(B.element_spec(B.load_basis("6-31g_st"), Z), False, precision)
        out.append((g, ids(fn), densities.cluster_density(rots, spherical=False)))
    return out


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_batch_matches_alone(precision):
    from pdynamo3_amd import integrator
    jobs = _jobs(precision)
    alone = []
    for g, fid, P in jobs:
        V = np.zeros_like(P)
        E, N, _ = g.integrate(fid, P, V)
        alone.append((E, N, V))
    batch = integrator.QuadratureBatch(8)
    for order in ([0, 1, 2], [2, 0, 1]):
        Vs = [np.zeros_like(jobs[k][2]) for k in order]
        res, stats = batch.integrate([(jobs[k][0], jobs[k][1], jobs[k][2], V) for k, V in zip(order, Vs)])
        assert stats["device_ms"] > 0.0
        for (E, N), V, k in zip(res, Vs, order):
            assert E == alone[k][0]
            assert N == alone[k][1]
            assert np.max(np.abs(V - alone[k][2])) <= 1e-12
    # The same job twice in one batch is refused (its work arrays would be shared).
    g, fid, P = jobs[0]
    with pytest.raises(integrator.QCModelError):
        batch.integrate([(g, fid, P, None), (g, fid, P, None)])
    batch.close()


def test_batch_of_one_matches_single_call():
    from pdynamo3_amd import integrator
    Z, xyz, spec, n = cluster_system(2, 4.5, 3, True)
    g = integrator.QuadratureGrid(Z, xyz, 2)
    g.set_basis(spec, True, "mixed")
    P = random_density(n, 10, 7)
    V1, V2 = np.zeros_like(P), np.zeros_like(P)
    E1, N1, _ = g.integrate(ids("pbe"), P, V1)
    batch = integrator.QuadratureBatch(1)
    [(E2, N2)], _ = batch.integrate([(g, ids("pbe"), P, V2)])
    assert E1 == E2 and N1 == N2
    assert np.max(np.abs(V1 - V2)) <= 1e-12
    free, total = integrator.device_memory()
    assert 0 < free <= total
