"""Definitions shared by the quadrature API: grid accuracies, functional names and IDs, unit constants.

Mirrors pMolecule/QCModel/DFTDefinitions.py (accuracy enum :11-17; functional table :49,54,78-79,212-222) and the unit
constants of pScientific (Constants.h:15, Units.h:35, Units.pyx:22,47).
"""

import enum


class DFTGridAccuracy(enum.IntEnum):
    """Grid accuracy classes (pMolecule/QCModel/DFTDefinitions.py:11-17; DFTGrid.h:16-25)."""
    VeryLow = 0
    Low = 1
    Medium = 2
    High = 3
    VeryHigh = 4


# . libxc functional IDs used by the supported functionals (DFTDefinitions.py:22,28-29,49,54,78-79,173).
LDA_X = 1
LDA_C_VWN = 7
LDA_C_VWN_RPA = 8
GGA_X_PBE = 101
GGA_X_B88 = 106
GGA_C_PBE = 130
GGA_C_LYP = 131
HYB_GGA_XC_B3LYP = 402

# . Functional name -> libxc IDs (DFTDefinitions.py:212-222): the named models whose functionals the library evaluates.
FUNCTIONAL_IDS = {"b3lyp": (HYB_GGA_XC_B3LYP,), "blyp": (GGA_X_B88, GGA_C_LYP), "lda": (LDA_X, LDA_C_VWN),
                  "pbe": (GGA_X_PBE, GGA_C_PBE)}

# . Fraction of exact (Hartree-Fock) exchange of the hybrid functionals (hyb_gga_xc_b3lyp.c:56); the caller's Fock build
#   adds it, as QCModelDFT does from DFTFunctionalModel's scaling (DFTFunctionalModel.c:82-103, QCModelDFT.py:225-243).
EXACT_EXCHANGE = {HYB_GGA_XC_B3LYP: 0.20}


def exact_exchange_scaling(ids):
    """The largest exact-exchange fraction among the functional IDs (0 for pure functionals)."""
    return max([EXACT_EXCHANGE.get(int(i), 0.0) for i in ids] + [0.0])

# . Bohr radius in metres (pScientific/__extensions__/cinclude/Constants.h:15).
BOHR_RADIUS_M = 5.291772083e-11
# . Angstrom -> bohr as computed in C (Units.h:35).
ANGSTROM_TO_BOHR_C = 1.0e-10 / BOHR_RADIUS_M
# . Angstrom -> bohr as the reference's Python layer applies it to atom coordinates (QCModel.py:174): the Cython extern
#   declares the constant as C float (pScientific/__extensions__/pyrex/pScientific.Units.pyx:22,47), so the value seen by
#   Python is the float32 rounding of the C constant.
ANGSTROM_TO_BOHR = float(__import__("numpy").float32(ANGSTROM_TO_BOHR_C))

# . Status codes of the C-ABI (pCore/__extensions__/cinclude/Status.h:24-33).
STATUS_OK = 0
STATUS_INVALID_ARGUMENT = 3
STATUS_NON_CONFORMABLE_ARRAYS = 6
STATUS_OUT_OF_MEMORY = 7


def functional_ids(name):
    """libxc IDs for a functional name (case-insensitive); raises ValueError for unsupported names."""
    try:
        return FUNCTIONAL_IDS[name.lower()]
    except KeyError:
        raise ValueError("Unsupported functional: {!r} (supported: {:s}).".format(name, ", ".join(sorted(FUNCTIONAL_IDS))))
