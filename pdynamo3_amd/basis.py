"""Gaussian basis sets for the quadrature: the committed JSON copies of the reference's basis parameter files, or the
parameter files themselves (load_basis_directory), flattened into the element specification the C-ABI takes.

Parameter files: one YAML mapping per element, <directory>/<Symbol>.yaml, with the keys "Atomic Number", "Basis Type",
"Is Spherical", "Normalized Primitives" and "Shells" (each shell a "Type" label and "Primitives": [exponent,
[coefficient per angular momentum]]), as GaussianBasis.__setstate__ reads them
(pMolecule/QCModel/GaussianBases/__extensions__/pyrex/pMolecule.QCModel.GaussianBases.GaussianBasis.pyx:109-137).

Element specification (flat float64 array), per element:
    Z, nShells, then per shell: lLow, lHigh, nPrimitives, then per primitive: exponent, one coefficient per l in lLow..lHigh.
Coefficients are the primitive-normalised input values of the parameter files; the library unnormalises them exactly as
GaussianBasis_UnnormalizePrimitives and GaussianBasis_NormalizePrimitiveCCBF do
(pMolecule/QCModel/GaussianBases/__extensions__/csource/GaussianBasis.c:448-501,728-751).
"""

import json
import os

import numpy as np

from .formats import symbol

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


class BasisError(ValueError):
    pass


def shell_label_decode(label):
    """(lLow, lHigh) of a shell label: letters of consecutive angular momenta, S = 0, P = 1, D = 2, then F = 3, G = 4, ...
    (ShellLabelDecode and AMLabelDecode, pMolecule/QCModel/GaussianBases/GaussianBasisUtilities.py:8-38)."""
    lLow = lHigh = None
    for i, c in enumerate(label):
        u = c.upper()
        l = {"S": 0, "P": 1, "D": 2}.get(u, ord(u) - ord("C"))
        if i == 0:
            lLow = lHigh = l
        elif l == lHigh + 1:
            lHigh = l
        else:
            lLow = None
            break
    if lLow is None or lLow < 0:
        raise BasisError("Invalid shell label {!r}.".format(label))
    return lLow, lHigh


def element_from_mapping(mapping):
    """One element's entry (the committed JSON form) from the mapping of its parameter file."""
    shells = []
    for shell in mapping["Shells"]:
        lLow, lHigh = shell_label_decode(shell["Type"])
        exponents, coefficients = [], []
        for exponent, cs in shell["Primitives"]:
            if len(cs) != lHigh - lLow + 1:
                raise BasisError("Shell {:s} needs {:d} coefficients per primitive.".format(shell["Type"], lHigh - lLow + 1))
            exponents.append(float(exponent))
            coefficients.append([float(c) for c in cs])
        shells.append({"type": shell["Type"], "lLow": lLow, "lHigh": lHigh, "exponents": exponents,
                       "coefficients": coefficients})
    z = int(mapping["Atomic Number"])
    return {"symbol": symbol(z), "isSpherical": bool(mapping["Is Spherical"]),
            "normalizedPrimitives": bool(mapping["Normalized Primitives"]), "shells": shells}


def load_basis_directory(path, atomic_numbers, label=None):
    """The basis set of a parameter directory for the unique elements among atomic_numbers, reading <Symbol>.yaml for
    each (GaussianBasisContainer.FromParameterDirectory, pMolecule/QCModel/GaussianBases/__extensions__/pyrex/
    pMolecule.QCModel.GaussianBases.GaussianBasisContainer.pyx:122-141).  The files are YAML mappings, read with
    yaml.SafeLoader; elements without a file are named together in one BasisError."""
    import yaml
    label = label if label is not None else os.path.basename(os.path.normpath(path))
    out, missing = {"label": label, "elements": {}}, []
    for z in sorted(set(int(a) for a in atomic_numbers)):
        entry = os.path.join(path, symbol(z) + ".yaml")
        if not os.path.exists(entry):
            missing.append(symbol(z))
            continue
        with open(entry) as f:
            mapping = yaml.load(f, Loader=yaml.SafeLoader)
        if int(mapping["Atomic Number"]) != z:
            raise BasisError("{:s} holds atomic number {:d}, not {:d}.".format(entry, int(mapping["Atomic Number"]), z))
        out["elements"][str(z)] = element_from_mapping(mapping)
    if missing:
        raise BasisError("There are missing {:s} basis sets: {:s}.".format(label, ", ".join(missing)))
    return out


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
            raise BasisError("Basis {:s} has no entry for Z = {:d}.".format(basis["label"], z))
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
