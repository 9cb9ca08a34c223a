"""Writes the golden vectors of tests/golden/ from the compiled reference path (oracle/_ref/libxcref.so).

The reference's C sources are compiled where they lie under the reference tree by `make -C oracle ref`; this script calls
them through oracle/ref_driver.c and stores inputs and outputs only (numbers, no reference code):

  reference_water.json      water: grid size, block structure, (x, y, z, w) checksums, the first 64 grid points; per
                            basis flavour (spherical / Cartesian) and functional (BLYP / PBE): E_xc, N_e and packed V_xc for
                            the EHT monomer density, and the kept functions and values of three blocks.
  reference_cluster.json    (H2O)4 (6 A cube, seed 5): grid size and checksums, E_xc, N_e and packed V_xc (PBE, spherical);
                            (H2O)16 (9 A cube, seed 1, the benchmark cluster): grid size, E_xc and N_e (PBE, Cartesian).
  reference_functionals.json  libxc-1.2.0 (as vendored) eps, vrho, vsigma of BLYP, PBE, their parts, Slater exchange, VWN
                            (both fits), LDA and B3LYP on bulk and threshold (rho, sigma).
  reference_functionals_pol.json  the same functionals spin-polarised (rho_a, rho_b; s_aa, s_ab, s_bb).
  reference_accuracy.json   water at the five grid accuracies (VeryLow .. VeryHigh): grid size, block count and checksums,
                            E_xc, N_e, packed V_xc and the XC nuclear gradients (BLYP, spherical, EHT monomer density);
                            and at Medium the unrestricted E_xc, N_e, V_xc of both spins and gradients for alpha = 0.6 P,
                            beta = 0.4 P (PBE).
  reference_cluster_h2o64.json  (H2O)64 BLYP (14 A cube, seed 2, Cartesian) at the whole-cluster extended-Hueckel
                            density: grid size, E_xc, N_e, every 97th packed V_xc entry, and the same sample of P.
  reference_c2s.json        the reference's own Cartesian -> spherical fixture (examples/pMolecule/data/yaml/
                            C2STransformations_0to6.yaml, read with SafeLoader) converted to JSON.

Run: make -C oracle ref && python tests/golden/make_golden.py [names of the functions below; default all]
"""

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402
from pdynamo3_amd import basis, densities, molecules  # noqa: E402

REFERENCE = os.environ.get("PDYNAMO3_REFERENCE", "/root/reference")
BLYP, PBE = [106, 131], [101, 130]


def checksums(xyzw):
    return {"sum_x": float(np.sum(xyzw[:, 0])), "sum_y": float(np.sum(xyzw[:, 1])), "sum_z": float(np.sum(xyzw[:, 2])),
            "sum_w": float(np.sum(xyzw[:, 3])), "sum_abs_w": float(np.sum(np.abs(xyzw[:, 3]))),
            "sum_w_r2": float(np.sum(xyzw[:, 3] * np.sum(xyzw[:, :3] ** 2, axis=1)))}


def water():
    ref = oracle.reference()
    b = basis.load_basis()
    Z, xyz = molecules.water()
    spec = basis.element_spec(b, Z)
    out = {"Z": Z.tolist(), "xyz_bohr": xyz.tolist(), "accuracy": 2, "cases": []}
    h = ref.setup(2, Z, xyz, spec, True, BLYP)
    xyzw, ba, bc = h.grid()
    out["npoints"] = h.npoints
    out["nblocks"] = h.nblocks
    out["block_atom"] = ba.tolist()
    out["block_count"] = bc.tolist()
    out["checksums"] = checksums(xyzw)
    out["first_points"] = xyzw[:64].tolist()
    for spherical in (True, False):
        P1, _ = densities.load_monomer_density(spherical=spherical)
        P = P1[np.tril_indices(P1.shape[0])]
        for name, ids in (("blyp", BLYP), ("pbe", PBE)):
            h = ref.setup(2, Z, xyz, spec, spherical, ids)
            e, ne, V, _ = h.integrate(P, 0)
            case = {"spherical": spherical, "functional": name, "E_xc": e, "N_e": ne, "V": V.tolist(), "blocks": []}
            if name == "blyp":
                for blk in (0, 40, 200):
                    idx, f, fx, fy, fz = h.block_values(blk)
                    case["blocks"].append({"block": blk, "indices": idx.tolist(), "f": f.tolist(), "fx": fx.tolist(),
                                           "fy": fy.tolist(), "fz": fz.tolist()})
            out["cases"].append(case)
    with open(os.path.join(HERE, "reference_water.json"), "w") as f:
        json.dump(out, f)


def cluster():
    ref = oracle.reference()
    b = basis.load_basis()
    out = {}
    Z, xa, rots, _ = molecules.water_cluster(4, 6.0, 5)
    xyz = molecules.to_bohr(xa)
    spec = basis.element_spec(b, Z)
    P = densities.cluster_density(rots, spherical=True)
    h = ref.setup(2, Z, xyz, spec, True, PBE)
    xyzw, ba, bc = h.grid()
    e, ne, V, _ = h.integrate(P, 0)
    out["h2o4"] = {"waters": 4, "cube": 6.0, "seed": 5, "spherical": True, "functional": "pbe", "npoints": h.npoints,
                   "nblocks": h.nblocks, "checksums": checksums(xyzw), "E_xc": e, "N_e": ne, "V": V.tolist()}
    print("h2o4", h.npoints, e, ne, flush=True)
    Z, xa, rots, _ = molecules.water_cluster(16, 9.0, 1)
    xyz = molecules.to_bohr(xa)
    spec = basis.element_spec(b, Z)
    P = densities.cluster_density(rots, spherical=False)
    h = ref.setup(2, Z, xyz, spec, False, PBE)
    xyzw, _, _ = h.grid()
    e, ne, V, sec = h.integrate(P, 0)
    out["h2o16"] = {"waters": 16, "cube": 9.0, "seed": 1, "spherical": False, "functional": "pbe", "npoints": h.npoints,
                    "nblocks": h.nblocks, "checksums": checksums(xyzw), "E_xc": e, "N_e": ne,
                    "V_norm": float(np.linalg.norm(V)), "V_trace_like": float(np.sum(V)), "seconds": sec}
    print("h2o16", h.npoints, e, ne, sec, flush=True)
    with open(os.path.join(HERE, "reference_cluster.json"), "w") as f:
        json.dump(out, f)


def eht_cluster_density(Z, xyz, spec, spherical):
    """The whole-cluster extended-Hueckel density (every pair of waters coupled), with the compiled reference's overlap."""
    S, _, _ = oracle.RefBasis(Z, xyz, spec, spherical).one_electron()
    return densities.eht_density(Z, S, spherical=spherical)


def cluster_large():
    """(H2O)64 BLYP (14 A cube, seed 2, Cartesian), the mixed-precision benchmark cluster, at its whole-cluster
    extended-Hueckel density: E_xc, N_e and every 97th packed V_xc entry from the compiled reference, with checksums and
    the same sample of P so that a test can confirm it rebuilt the same density."""
    Z, xa, _, _ = molecules.water_cluster(64, 14.0, 2)
    xyz = molecules.to_bohr(xa)
    spec = basis.element_spec(basis.load_basis(), Z)
    P = eht_cluster_density(Z, xyz, spec, False)
    h = oracle.reference().setup(2, Z, xyz, spec, False, BLYP)
    print("h2o64 grid", h.npoints, flush=True)
    e, ne, V, sec = h.integrate(P, 0)
    stride = 97
    out = {"waters": 64, "cube": 14.0, "seed": 2, "spherical": False, "functional": "blyp",
           "density": "whole-cluster extended-Hueckel (pdynamo3_amd.densities.eht_density, reference overlap)",
           "npoints": h.npoints, "nblocks": h.nblocks, "E_xc": e, "N_e": ne, "stride": stride,
           "V_sample": V[::stride].tolist(), "P_sample": P[::stride].tolist(), "V_norm": float(np.linalg.norm(V)),
           "V_sum": float(np.sum(V)), "P_norm": float(np.linalg.norm(P)), "seconds": sec}
    print("h2o64", h.npoints, e, ne, sec, flush=True)
    with open(os.path.join(HERE, "reference_cluster_h2o64.json"), "w") as f:
        json.dump(out, f)


def functionals():
    ref = oracle.reference()
    rho = [5.0e-13, 4.9e-13, 1.0e-12, 1.0e-10, 1.0e-6, 1.0e-3, 0.1, 1.0, 10.0, 300.0, 1.0e-3, 1.0, 1.0, 0.5, 2.0e-5]
    sigma = [0.0, 1.0e-20, 1.0e-30, 1.0e-18, 1.0e-9, 1.0e-5, 0.01, 1.0, 100.0, 1.0e5, 0.0, 2.5e-25, 1.0e-26, 7.3, 4.0e-8]
    out = {"rho": rho, "sigma": sigma}
    for name, ids in (("blyp", BLYP), ("pbe", PBE), ("b88", [106]), ("lyp", [131]), ("pbex", [101]), ("pbec", [130]),
                      ("slater", [1]), ("vwn", [7]), ("vwnrpa", [8]), ("lda", [1, 7]), ("b3lyp", [402])):
        e, vr, vs = ref.functional(ids, np.asarray(rho), np.asarray(sigma))
        out[name] = {"ids": ids, "eps": e.tolist(), "vrho": vr.tolist(), "vsigma": vs.tolist()}
    with open(os.path.join(HERE, "reference_functionals.json"), "w") as f:
        json.dump(out, f)


def functionals_pol():
    """Spin-polarised values (libxc with nspin = 2): bulk, one spin empty or below threshold, fully polarised, negative
    s_ab."""
    ref = oracle.reference()
    rho = [[0.3, 0.1], [1e-3, 2e-3], [2.0, 2.0], [0.5, 0.0], [1e-13, 0.3], [0.7, 1e-14], [0.2, 0.2], [5e-13, 5e-13],
           [0.4, 0.4], [10.0, 3.0], [1e-6, 3e-6], [0.25, 0.25]]
    sigma = [[0.01, 0.004, 0.003], [1e-8, -1e-8, 3e-8], [3.0, 2.5, 3.0], [0.2, 0.0, 0.0], [0.0, 0.0, 0.1],
             [0.3, 0.0, 0.0], [1e-30, 0.0, 1e-30], [1e-20, 0.0, 1e-20], [0.5, 0.5, 0.5], [40.0, 15.0, 9.0],
             [1e-12, 1e-12, 2e-12], [0.0, 0.0, 0.0]]
    out = {"rho": rho, "sigma": sigma}
    for name, ids in (("blyp", BLYP), ("pbe", PBE), ("b88", [106]), ("lyp", [131]), ("pbex", [101]), ("pbec", [130]),
                      ("slater", [1]), ("vwn", [7]), ("vwnrpa", [8]), ("lda", [1, 7]), ("b3lyp", [402])):
        e, vr, vs = ref.functional_pol(ids, np.asarray(rho), np.asarray(sigma))
        out[name] = {"ids": ids, "eps": e.tolist(), "vrho": vr.tolist(), "vsigma": vs.tolist()}
    with open(os.path.join(HERE, "reference_functionals_pol.json"), "w") as f:
        json.dump(out, f)


def c2s():
    import yaml
    path = os.path.join(REFERENCE, "examples", "pMolecule", "data", "yaml", "C2STransformations_0to6.yaml")
    with open(path) as f:
        data = yaml.load(f, Loader=yaml.SafeLoader)
    out = {"source": "examples/pMolecule/data/yaml/C2STransformations_0to6.yaml"}
    ref = oracle.reference()
    for l in range(0, 7):
        out["reference_driver_l{:d}".format(l)] = ref.c2s(l, l).tolist()
    out["yaml"] = json.loads(json.dumps(data, default=str))
    with open(os.path.join(HERE, "reference_c2s.json"), "w") as f:
        json.dump(out, f)


def accuracies():
    ref = oracle.reference()
    b = basis.load_basis()
    Z, xyz = molecules.water()
    spec = basis.element_spec(b, Z)
    P1, _ = densities.load_monomer_density(spherical=True)
    P = P1[np.tril_indices(P1.shape[0])]
    out = {"Z": Z.tolist(), "xyz_bohr": xyz.tolist(), "functional": "blyp", "spherical": True, "levels": []}
    for accuracy in range(5):
        h = ref.setup(accuracy, Z, xyz, spec, True, BLYP)
        xyzw, ba, bc = h.grid()
        e, ne, V, _ = h.integrate(P, 0)
        g = h.gradients(P, 0)
        out["levels"].append({"accuracy": accuracy, "npoints": h.npoints, "nblocks": h.nblocks,
                              "checksums": checksums(xyzw), "E_xc": e, "N_e": ne, "V": V.tolist(),
                              "gradients": g.tolist()})
    h = ref.setup(2, Z, xyz, spec, True, PBE)
    e, ne, Va, Vb, _ = h.integrate_uks(0.6 * P, 0.4 * P, 0)
    g = h.gradients_uks(0.6 * P, 0.4 * P, 0)
    out["uks"] = {"accuracy": 2, "functional": "pbe", "alpha": 0.6, "beta": 0.4, "E_xc": e, "N_e": ne,
                  "Va": Va.tolist(), "Vb": Vb.tolist(), "gradients": g.tolist()}
    with open(os.path.join(HERE, "reference_accuracy.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    steps = {"functionals": functionals, "functionals_pol": functionals_pol, "c2s": c2s, "water": water,
             "cluster": cluster, "cluster_large": cluster_large, "accuracies": accuracies}
    for name in sys.argv[1:] or list(steps):
        steps[name]()
