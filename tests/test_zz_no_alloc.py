"""xq_integrate_noalloc, the quadrature entry point that allocates nothing, against xq_integrate through the C-ABI.

Two calls in a row in each precision share the one scratch array: first (H2O)8 (152 Cartesian functions), then the
smaller (H2O)3 on a handle that has not yet allocated its own work arrays, so the second call runs over what the first
left in the array.  Each call's E_xc and N_e equal
xq_integrate's bit for bit on the same handle and density, its V_xc (added to a non-zero V, as the reference's Fock
matrix is incremented) agrees within 1e-12, and so do its statistics.  A molecule whose work arrays exceed the array is
refused with status 3 before any work is queued, and the handle still answers xq_integrate afterwards.
"""

import ctypes

import numpy as np
import pytest

from helpers import cluster_system

pytestmark = pytest.mark.gpu


def _call(name, grid, ids, P, V0):
    from pdynamo3_amd import _lib, integrator
    V = V0.copy()
    e, n = ctypes.c_double(), ctypes.c_double()
    st = np.zeros(integrator.NSTATS)
    ids = _lib.i32(ids)
    status = getattr(_lib.lib, name)(grid._h, ids.size, _lib.iptr(ids), _lib.dptr(P), ctypes.byref(e), ctypes.byref(n),
                                     _lib.dptr(V), _lib.dptr(st))
    return status, e.value, n.value, V, st


def _system(n_waters, seed):
    from pdynamo3_amd import densities, molecules
    Z, xyz, spec, nf = cluster_system(n_waters, 9.0 * (n_waters / 16.0) ** (1.0 / 3.0), seed, spherical=False)
    _, _, rots, _ = molecules.water_cluster(n_waters, 9.0 * (n_waters / 16.0) ** (1.0 / 3.0), seed)
    P = densities.cluster_density(rots, spherical=False)
    assert P.size == nf * (nf + 1) // 2
    return Z, xyz, spec, P


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_noalloc_equals_integrate_on_two_calls(precision):
    from pdynamo3_amd import _lib, integrator
    ids = [106, 131]
    for n_waters, seed in ((8, 11), (3, 12)):   # the second molecule is smaller
        Z, xyz, spec, P = _system(n_waters, seed)
        g = integrator.QuadratureGrid(Z, xyz, 2)
        g.set_basis(spec, False, precision)
        need, limit = ctypes.c_longlong(), ctypes.c_longlong()
        assert _lib.lib.xq_noalloc_scratch(g._h, ctypes.byref(need), ctypes.byref(limit)) == 0
        assert 0 < need.value <= limit.value == 64 << 20
        V0 = np.random.default_rng(seed).uniform(-1.0, 1.0, P.size)
        if n_waters == 8:
            s1, e1, n1, V1, st1 = _call("xq_integrate", g, ids, P, V0)
            s2, e2, n2, V2, st2 = _call("xq_integrate_noalloc", g, ids, P, V0)
        else:                                        # on a handle that has never allocated its work arrays
            s2, e2, n2, V2, st2 = _call("xq_integrate_noalloc", g, ids, P, V0)
            s1, e1, n1, V1, st1 = _call("xq_integrate", g, ids, P, V0)
        g.close()
        assert s1 == 0 and s2 == 0, _lib.last_error()
        print(precision, n_waters, "E", e1, e2, "N", n1, n2, "max|dV|", float(np.max(np.abs(V2 - V1))),
              "scratch bytes", need.value, flush=True)
        assert e2 == e1 and n2 == n1
        assert np.max(np.abs(V2 - V1)) <= 1e-12
        assert np.max(np.abs(V2 - V0)) > 1e-3          # V_xc was added
        np.testing.assert_array_equal(st2[[0, 1, 5, 6, 7, 8]], st1[[0, 1, 5, 6, 7, 8]])


def test_noalloc_refuses_a_molecule_beyond_the_scratch():
    from pdynamo3_amd import _lib, integrator
    Z, xyz, spec, P = _system(96, 13)            # 1,824 Cartesian functions
    g = integrator.QuadratureGrid(Z, xyz, 2)
    g.set_basis(spec, False, "mixed")
    need, limit = ctypes.c_longlong(), ctypes.c_longlong()
    assert _lib.lib.xq_noalloc_scratch(g._h, ctypes.byref(need), ctypes.byref(limit)) == 0
    assert need.value > limit.value
    V0 = np.zeros(P.size)
    status, e, n, V, _ = _call("xq_integrate_noalloc", g, [101, 130], P, V0)
    msg = _lib.last_error()
    assert status == 3 and "scratch" in msg and e == 0.0 and n == 0.0
    assert np.all(V == 0.0)
    status, e, n, V, _ = _call("xq_integrate", g, [101, 130], P, V0)
    g.close()
    assert status == 0 and abs(n - 960.0) < 1e-2
