"""Input formats beside the quadrature: element symbols and XYZ coordinate files.

atomic_number follows PeriodicTable.AtomicNumber (pScientific/PeriodicTable.py:104-121): a token that starts with a letter
is a symbol (first letter upper case, a second letter lower case when it is a letter, so "C1" is carbon and "CA" calcium);
a token that starts with a digit is the integer its leading digits spell; anything else is -1.

read_xyz follows XYZFileReader.Parse (pBabel/XYZFileReader.py:19-47): the atom count on the first line, a title line, then
one "element x y z" line per atom (Angstrom).  The reference stops quietly at the end of a short file; here a short or
malformed file raises FormatError, since a partly read geometry would be integrated as if it were whole.
"""

import numpy as np

SYMBOLS = ("",
           "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si", "P", "S", "Cl", "Ar", "K", "Ca",
           "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr", "Rb", "Sr", "Y",
           "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn", "Sb", "Te", "I", "Xe", "Cs", "Ba", "La", "Ce",
           "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W", "Re", "Os", "Ir",
           "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra", "Ac", "Th", "Pa", "U", "Np", "Pu", "Am", "Cm",
           "Bk", "Cf", "Es", "Fm", "Md", "No", "Lr", "Rf", "Db", "Sg", "Bh", "Hs", "Mt", "Ds", "Rg", "Cn", "Nh", "Fl", "Mc",
           "Lv", "Ts", "Og")
_NUMBER = {s: z for z, s in enumerate(SYMBOLS) if s}
UNKNOWN_ELEMENT = -1


class FormatError(Exception):
    pass


def symbol(z):
    """The element symbol of atomic number z."""
    if not 0 < int(z) < len(SYMBOLS):
        raise FormatError("No element with atomic number {:d}.".format(int(z)))
    return SYMBOLS[int(z)]


def atomic_number(token):
    """The atomic number a token names (PeriodicTable.AtomicNumber), or -1."""
    if len(token) == 0:
        return UNKNOWN_ELEMENT
    if token[0].isalpha():
        name = token[0].upper()
        if len(token) > 1 and token[1].isalpha():
            name += token[1].lower()
        return _NUMBER.get(name, UNKNOWN_ELEMENT)
    digits = 0
    for c in token:
        if not c.isdigit():
            break
        digits += 1
    return int(token[:digits]) if digits > 0 else UNKNOWN_ELEMENT


def read_xyz(path):
    """(atomic numbers int32 [n], coordinates float64 [n, 3] in Angstrom, title) of an XYZ file."""
    with open(path) as f:
        lines = f.read().splitlines()
    try:
        n = int(lines[0].split()[0])
    except (IndexError, ValueError):
        raise FormatError("{:s}: the first line must give the number of atoms.".format(path))
    if n < 0 or len(lines) < n + 2:
        raise FormatError("{:s}: {:d} atoms announced, {:d} atom lines present.".format(path, n, max(len(lines) - 2, 0)))
    title = lines[1]
    Z = np.zeros(n, dtype=np.int32)
    xyz = np.zeros((n, 3), dtype=np.float64)
    for i in range(n):
        tokens = lines[2 + i].split()
        if len(tokens) < 4:
            raise FormatError("{:s}, line {:d}: expected an element and three coordinates.".format(path, i + 3))
        Z[i] = atomic_number(tokens[0])
        if Z[i] <= 0:
            raise FormatError("{:s}, line {:d}: unknown element {!r}.".format(path, i + 3, tokens[0]))
        try:
            xyz[i] = [float(t) for t in tokens[1:4]]
        except ValueError:
            raise FormatError("{:s}, line {:d}: the coordinates are not numbers.".format(path, i + 3))
    return Z, xyz, title


def write_xyz(path, atomic_numbers, xyz_angstrom, title=""):
    """Write an XYZ file (the format read_xyz reads)."""
    xyz = np.asarray(xyz_angstrom, dtype=np.float64).reshape(-1, 3)
    with open(path, "w") as f:
        f.write("{:d}\n{:s}\n".format(len(xyz), title))
        for z, (x, y, w) in zip(atomic_numbers, xyz):
            f.write("{:<2s} {:20.12f} {:20.12f} {:20.12f}\n".format(symbol(z), x, y, w))
