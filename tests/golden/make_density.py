"""Writes the water monomer densities pdynamo3_amd/data/water_density_<basis>_<sph|cart>.json (synthetic input data of
the benchmarks and tests; the script lives here because it takes the overlap matrix from the oracle).

The reference's converged SCF densities cannot be produced here (its SCF needs the full Python package), so the monomer
density is an extended-Hueckel density in the reference's own basis: H_ii from valence-orbital ionisation energies
(O 1s -20.6 Eh, O 2s -1.187 Eh, O 2p -0.544 Eh, H 1s -0.4998 Eh, O d -0.05 Eh), H_ij = 1.75 S_ij (H_ii + H_jj) / 2, the
overlap S evaluated on the Medium grid with the FP64 oracle (S_mn = sum_p w_p phi_m phi_n), five doubly occupied orbitals
from H C = S C e.  It is a test input only: realistic magnitudes, symmetry and positivity, N_e = 10.

Run: python tests/golden/make_density.py   (needs oracle/liboracle.so)
"""

import json
import os
import sys

import numpy as np
import scipy.linalg

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402
from pdynamo3_amd import basis, molecules  # noqa: E402

HII = {(8, 0, 0): -20.6, (8, "s"): -1.187, (8, "p"): -0.544, (8, "d"): -0.05, (1, "s"): -0.4998}


def diagonal_energies(b, Z, spherical):
    h = []
    for z in Z:
        for k, shell in enumerate(b["elements"][str(z)]["shells"]):
            for l in range(shell["lLow"], shell["lHigh"] + 1):
                n = (2 * l + 1) if spherical else (l + 1) * (l + 2) // 2
                if z == 8 and k == 0:
                    e = HII[(8, 0, 0)]
                else:
                    e = HII[(z, "spd"[l])]
                h += [e] * n
    return np.asarray(h)


def overlap(Z, xyz, spec, spherical):
    h = oracle.oracle().setup(2, Z, xyz, spec, spherical, [106, 131])
    n = h.nfunctions
    S = np.zeros((n, n))
    xyzw, _, counts = h.grid()
    start = 0
    for b, c in enumerate(counts):
        idx, f, _, _, _ = h.block_values(b)
        w = xyzw[start:start + c, 3]
        start += c
        if len(idx):
            S[np.ix_(idx, idx)] += (f * w) @ f.T
    return 0.5 * (S + S.T)


def main():
    b = basis.load_basis("6-31g_st")
    Z, xyz = molecules.water()
    for spherical in (True, False):
        spec = basis.element_spec(b, Z)
        S = overlap(Z, xyz, spec, spherical)
        hd = diagonal_energies(b, Z, spherical)
        H = 1.75 * S * 0.5 * (hd[:, None] + hd[None, :])
        np.fill_diagonal(H, hd * np.diag(S))
        e, C = scipy.linalg.eigh(H, S)
        Cocc = C[:, :5]
        P = 2.0 * Cocc @ Cocc.T
        out = {"description": "Extended-Hueckel water density (tests/golden/make_density.py); O at origin, H at "
                              "(+-0.757, 0.586, 0) A; basis 6-31g_st, {:s} functions.".format(
                                  "spherical" if spherical else "Cartesian"),
               "orbital_energies": e[:8].tolist(), "electrons_tr_PS": float(np.sum(P * S)), "P": P.tolist()}
        name = os.path.join(os.path.dirname(os.path.dirname(HERE)), "pdynamo3_amd", "data",
                            "water_density_6-31g_st_{:s}.json".format("sph" if spherical else "cart"))
        with open(name, "w") as f:
            json.dump(out, f)
        print(name, "Tr(PS) =", out["electrons_tr_PS"], "e =", np.round(e[:6], 4))


if __name__ == "__main__":
    main()
