"""The oracle against the golden vectors written from the compiled reference path (tests/golden/reference_*.json).

Bars: grid bit-exact (same point order, block structure; coordinates and weights equal to the last bit for the water
molecule), functional values to 1e-13 relative, basis values to 1e-13, E_xc / N_e to 1e-11 Eh and V_xc to 1e-12 (different
summation order only).
"""

import json
import os

import numpy as np
import pytest

from helpers import packed, water_system
from pdynamo3_amd import basis, densities, molecules

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def water_ref():
    return load("reference_water.json")


def test_water_grid_is_the_reference_grid(oracle_lib, water_ref):
    Z, xyz, spec, _ = water_system()
    assert np.array_equal(np.asarray(water_ref["xyz_bohr"]), xyz)
    h = oracle_lib.setup(2, Z, xyz, spec, True, [106, 131])
    xyzw, ba, bc = h.grid()
    assert h.npoints == water_ref["npoints"] == 42078
    assert h.nblocks == water_ref["nblocks"] == 330
    assert ba.tolist() == water_ref["block_atom"]
    assert bc.tolist() == water_ref["block_count"]
    assert np.array_equal(xyzw[:64], np.asarray(water_ref["first_points"]))
    for k, v in water_ref["checksums"].items():
        col = {"sum_x": xyzw[:, 0], "sum_y": xyzw[:, 1], "sum_z": xyzw[:, 2], "sum_w": xyzw[:, 3],
               "sum_abs_w": np.abs(xyzw[:, 3]), "sum_w_r2": xyzw[:, 3] * np.sum(xyzw[:, :3] ** 2, axis=1)}[k]
        assert np.sum(col) == pytest.approx(v, rel=1e-14, abs=1e-12)


@pytest.mark.parametrize("case", range(4))
def test_water_quadrature_matches_reference(oracle_lib, water_ref, case):
    c = water_ref["cases"][case]
    Z, xyz, spec, n = water_system(c["spherical"])
    P1, _ = densities.load_monomer_density(spherical=c["spherical"])
    ids = {"blyp": [106, 131], "pbe": [101, 130]}[c["functional"]]
    h = oracle_lib.setup(2, Z, xyz, spec, c["spherical"], ids)
    e, ne, V, _ = h.integrate(packed(P1))
    assert abs(e - c["E_xc"]) <= 1e-11
    assert abs(ne - c["N_e"]) <= 1e-11
    assert np.max(np.abs(V - np.asarray(c["V"]))) <= 1e-12
    for blk in c["blocks"]:
        idx, f, fx, fy, fz = h.block_values(blk["block"])
        assert idx.tolist() == blk["indices"]
        for mine, ref in ((f, "f"), (fx, "fx"), (fy, "fy"), (fz, "fz")):
            r = np.asarray(blk[ref]).reshape(mine.shape)
            assert np.max(np.abs(mine - r)) <= 1e-13 * max(1.0, np.max(np.abs(r)))


def test_cluster4_matches_reference(oracle_lib):
    c = load("reference_cluster.json")["h2o4"]
    Z, xa, rots, _ = molecules.water_cluster(c["waters"], c["cube"], c["seed"])
    xyz = molecules.to_bohr(xa)
    spec = basis.element_spec(basis.load_basis(), Z)
    P = densities.cluster_density(rots, spherical=c["spherical"])
    h = oracle_lib.setup(2, Z, xyz, spec, c["spherical"], [101, 130])
    assert h.npoints == c["npoints"] and h.nblocks == c["nblocks"]
    xyzw, _, _ = h.grid()
    assert np.sum(xyzw[:, 3]) == pytest.approx(c["checksums"]["sum_w"], rel=1e-13)
    e, ne, V, _ = h.integrate(P, threads=4)
    assert abs(e - c["E_xc"]) <= 1e-10
    assert abs(ne - c["N_e"]) <= 1e-10
    assert np.max(np.abs(V - np.asarray(c["V"]))) <= 1e-11


def test_cluster16_grid_matches_reference_count():
    """Grid size of the benchmark cluster (H2O)16 (the quadrature itself is pinned on the GPU against this E_xc)."""
    c = load("reference_cluster.json")["h2o16"]
    assert c["npoints"] == 625927 and c["nblocks"] == 4913


@pytest.mark.parametrize("name", ["blyp", "pbe", "b88", "lyp", "pbex", "pbec", "slater", "vwn", "vwnrpa", "lda", "b3lyp"])
def test_functionals_match_reference(oracle_lib, name):
    d = load("reference_functionals.json")
    ref = d[name]
    e, vr, vs = oracle_lib.functional(ref["ids"], np.asarray(d["rho"]), np.asarray(d["sigma"]))
    for mine, key in ((e, "eps"), (vr, "vrho"), (vs, "vsigma")):
        r = np.asarray(ref[key])
        assert np.allclose(mine, r, rtol=1e-13, atol=1e-300), (key, mine, r)


@pytest.mark.parametrize("l", range(7))
def test_c2s_matches_reference_fixture(oracle_lib, l):
    d = load("reference_c2s.json")
    mine = oracle_lib.c2s(l, l)
    fixture = np.asarray(d["yaml"][str(l)], dtype=float)
    assert np.max(np.abs(mine - fixture)) <= 1e-10          # the reference's own test bar
    assert np.max(np.abs(mine - np.asarray(d["reference_driver_l{:d}".format(l)]))) <= 1e-14
