"""Gaussian basis sets for the quadrature: the committed JSON copies of the reference's basis parameter files, flattened into
the element specification the C-ABI takes.

Element specification (flat float64 array), per element:
    Z, nShells, then per shell: lLow, lHigh, nPrimitives, then per primitive: exponent, one coefficient per l in lLow..lHigh.
Coefficients are the primitive-normalised input values of the parameter files; the library unnormalises them exactly as
GaussianBasis_UnnormalizePrimitives and GaussianBasis_NormalizePrimitiveCCBF do
(pMolecule/QCModel/GaussianBases/__extensions__/csource/GaussianBasis.c:448-501,728-751).
"""

import json
import os

import numpy as np

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def load_basis(label="6-31g_st"):
    """The basis set with the given label, as stored in data/basis_<label>.json."""
    with open(os.path.join(_DATA, "basis_{:s}.json".format(label))) as f:
        return json.load(f)


def element_spec(basis, atomic_numbers):
    """Flat float64 element specification for the unique elements among atomic_numbers (sorted by Z)."""
    out = []
    for z in sorted(set(int(a) for a in atomic_numbers)):
        entry = basis["elements"].get(str(z))
        if entry is None:
            raise ValueError("Basis {:s} has no entry for Z = {:d}.".format(basis["label"], z))
        out += [float(z), float(len(entry["shells"]))]
        for shell in entry["shells"]:
            out += [float(shell["lLow"]), float(shell["lHigh"]), float(len(shell["exponents"]))]
            for exponent, coefficients in zip(shell["exponents"], shell["coefficients"]):
                out.append(float(exponent))
                out += [float(c) for c in coefficients]
    return np.asarray(out, dtype=np.float64)


def number_of_functions(basis, atomic_numbers, spherical=True):
    """Number of basis functions for the given atoms."""
    n = 0
    for z in atomic_numbers:
        for shell in basis["elements"][str(int(z))]["shells"]:
            for l in range(shell["lLow"], shell["lHigh"] + 1):
                n += (2 * l + 1) if spherical else ((l + 1) * (l + 2)) // 2
    return n
