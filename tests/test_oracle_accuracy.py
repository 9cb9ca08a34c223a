"""The oracle at the five grid accuracies (DFTGridAccuracy VeryLow .. VeryHigh, DFTGrid.c:475-510) against the compiled
reference's outputs committed in tests/golden/reference_accuracy.json (water, BLYP, spherical 6-31G*, EHT monomer
density): grid size, block count and checksums, E_xc, N_e, packed V_xc and the XC nuclear gradients; at Medium also the
unrestricted quadrature and gradients (PBE, alpha = 0.6 P, beta = 0.4 P).  The point count grows with the level and
E_xc converges (VeryHigh and High agree to 1e-6 Eh), which a wrong radial or angular parameter per level would break.
"""

import json
import os

import numpy as np
import pytest

from pdynamo3_amd import basis, densities, molecules
from pdynamo3_amd.definitions import functional_ids

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def golden():
    with open(os.path.join(HERE, "golden", "reference_accuracy.json")) as f:
        return json.load(f)


def _system():
    Z, xyz = molecules.water()
    spec = basis.element_spec(basis.load_basis(), Z)
    P1, _ = densities.load_monomer_density(spherical=True)
    return Z, xyz, spec, P1[np.tril_indices(P1.shape[0])]


def _sums(xyzw):
    return {"sum_x": np.sum(xyzw[:, 0]), "sum_y": np.sum(xyzw[:, 1]), "sum_z": np.sum(xyzw[:, 2]),
            "sum_w": np.sum(xyzw[:, 3]), "sum_abs_w": np.sum(np.abs(xyzw[:, 3])),
            "sum_w_r2": np.sum(xyzw[:, 3] * np.sum(xyzw[:, :3] ** 2, axis=1))}


@pytest.mark.parametrize("accuracy", range(5))
def test_oracle_matches_reference_at_every_accuracy(oracle_lib, golden, accuracy):
    ref = golden["levels"][accuracy]
    assert ref["accuracy"] == accuracy
    Z, xyz, spec, P = _system()
    h = oracle_lib.setup(accuracy, Z, xyz, spec, True, list(functional_ids("blyp")))
    assert h.npoints == ref["npoints"] and h.nblocks == ref["nblocks"]
    sums = _sums(h.grid()[0])
    for k, v in ref["checksums"].items():
        assert sums[k] == pytest.approx(v, rel=1e-13, abs=1e-11), k
    e, ne, V, _ = h.integrate(P, 4)
    assert abs(e - ref["E_xc"]) <= 1e-11 and abs(ne - ref["N_e"]) <= 1e-11
    assert np.max(np.abs(V - np.asarray(ref["V"]))) <= 1e-12
    g = h.gradients(P, 4)
    assert np.max(np.abs(g - np.asarray(ref["gradients"]))) <= 1e-11


def test_levels_refine_the_quadrature(golden):
    levels = golden["levels"]
    counts = [lv["npoints"] for lv in levels]
    assert counts == sorted(counts) and len(set(counts)) == 5
    e = [lv["E_xc"] for lv in levels]
    assert abs(e[4] - e[3]) < 1e-6 < abs(e[0] - e[4])
    assert all(abs(lv["N_e"] - 10.0) < 1e-2 for lv in levels[1:])


def test_oracle_uks_matches_reference_fixture(oracle_lib, golden):
    ref = golden["uks"]
    Z, xyz, spec, P = _system()
    h = oracle_lib.setup(ref["accuracy"], Z, xyz, spec, True, list(functional_ids(ref["functional"])))
    Pa, Pb = ref["alpha"] * P, ref["beta"] * P
    e, ne, Va, Vb, _ = h.integrate_uks(Pa, Pb, 4)
    assert abs(e - ref["E_xc"]) <= 1e-11 and abs(ne - ref["N_e"]) <= 1e-11
    assert np.max(np.abs(Va - np.asarray(ref["Va"]))) <= 1e-12
    assert np.max(np.abs(Vb - np.asarray(ref["Vb"]))) <= 1e-12
    g = h.gradients_uks(Pa, Pb, 4)
    assert np.max(np.abs(g - np.asarray(ref["gradients"]))) <= 1e-11
