"""Input formats (CPU): element symbols, XYZ files, and basis parameter files.

- atomic_number against the cases PeriodicTable.AtomicNumber (pScientific/PeriodicTable.py:104-121) distinguishes;
- XYZ files written and read back, and the errors of short or malformed files;
- shell labels against ShellLabelDecode (GaussianBasisUtilities.py:23-38), including the labels it rejects;
- the parameter-file loader against the committed JSON copies (6-31G* and def2-SVP, H C N O), when the reference tree is
  present;
- a def2-SVP water quadrature (spherical d on O, p on H): the oracle against the compiled reference (E_xc, N_e, V_xc),
  and methane read from an XYZ file, which puts carbon through the same comparison.
"""

import os

import numpy as np
import pytest

from helpers import ids, methane_angstrom, random_density, water_system
from pdynamo3_amd import basis, formats, molecules

REFERENCE_BASES = "/root/reference/parameters/gaussianBasisSets"


@pytest.mark.parametrize("token,z", [("O", 8), ("o", 8), ("Cl", 17), ("CL", 17), ("C1", 6), ("CA", 20), ("Ca2", 20),
                                     ("8", 8), ("12abc", 12), ("Xx", -1), ("", -1), ("-3", -1)])
def test_atomic_number(token, z):
    assert formats.atomic_number(token) == z


def test_symbols():
    assert formats.symbol(1) == "H" and formats.symbol(8) == "O" and formats.symbol(118) == "Og"
    assert all(formats.atomic_number(formats.symbol(z)) == z for z in range(1, 119))
    with pytest.raises(formats.FormatError):
        formats.symbol(0)


def test_xyz_round_trip(tmp_path):
    Z, xa, _, _ = molecules.water_cluster(4, 6.0, 5)
    path = str(tmp_path / "w4.xyz")
    formats.write_xyz(path, Z, xa, "water tetramer")
    Z2, x2, title = formats.read_xyz(path)
    assert title == "water tetramer"
    assert np.array_equal(Z2, np.asarray(Z)) and np.max(np.abs(x2 - xa)) <= 1e-12


def test_xyz_errors(tmp_path):
    path = str(tmp_path / "bad.xyz")
    for text in ["", "x\n", "3\ntitle\nO 0 0 0\nH 0 0 1\n", "1\nt\nQq 0 0 0\n", "1\nt\nO 0 zero 0\n", "1\nt\nO 0 0\n"]:
        with open(path, "w") as f:
            f.write(text)
        with pytest.raises(formats.FormatError):
            formats.read_xyz(path)
    with open(path, "w") as f:
        f.write("2\n\n8 0.0 0.0 0.0\nh 0.0 0.0 0.96\n")       # element numbers and lower case symbols are accepted
    Z, xyz, _ = formats.read_xyz(path)
    assert list(Z) == [8, 1] and xyz[1, 2] == 0.96


@pytest.mark.parametrize("label,l", [("S", (0, 0)), ("P", (1, 1)), ("D", (2, 2)), ("F", (3, 3)), ("G", (4, 4)),
                                     ("SP", (0, 1)), ("SPD", (0, 2)), ("sp", (0, 1)), ("DF", (2, 3))])
def test_shell_labels(label, l):
    assert basis.shell_label_decode(label) == l


@pytest.mark.parametrize("label", ["", "PS", "SD", "SPP", "A"])
def test_shell_labels_rejected(label):
    with pytest.raises(basis.BasisError):
        basis.shell_label_decode(label)


@pytest.mark.parametrize("label", ["6-31g_st", "def2-svp"])
def test_parameter_files_match_committed_copies(label):
    path = os.path.join(REFERENCE_BASES, label)
    if not os.path.isdir(path):
        pytest.skip("reference parameter files not present")
    loaded = basis.load_basis_directory(path, [8, 1, 6, 7, 1])
    assert loaded == basis.load_basis(label)


def test_missing_parameter_files_are_named():
    path = os.path.join(REFERENCE_BASES, "6-31g_st")
    if not os.path.isdir(path):
        pytest.skip("reference parameter files not present")
    with pytest.raises(basis.BasisError, match="U"):
        basis.load_basis_directory(path, [1, 8, 92])


def test_committed_basis_needs_every_element():
    with pytest.raises(basis.BasisError):
        basis.element_spec(basis.load_basis("def2-svp"), [1, 8, 16])


def _reference_or_skip():
    import oracle
    if not oracle.reference_available():
        pytest.skip("compiled reference not built (make -C oracle ref)")
    return oracle


@pytest.mark.parametrize("functional", ["blyp", "pbe"])
def test_def2svp_water_oracle_matches_reference(functional):
    oracle = _reference_or_skip()
    Z, xyz, spec, n = water_system(True, "def2-svp")
    assert n == 24                                    # O: 3 s, 2 p, 1 d (spherical) = 14; H: 2 s, 1 p = 5 each
    P = random_density(n, 5, seed=21)
    e0, n0, V0, _ = oracle.reference().setup(2, Z, xyz, spec, True, ids(functional)).integrate(P, 1)
    e, ne, V, _ = oracle.oracle().setup(2, Z, xyz, spec, True, ids(functional)).integrate(P, 4)
    assert abs(e - e0) <= 1e-11 and abs(ne - n0) <= 1e-11
    assert np.max(np.abs(V - V0)) <= 1e-12


def test_methane_from_xyz_oracle_matches_reference(tmp_path):
    oracle = _reference_or_skip()
    Zm, xa = methane_angstrom()
    path = str(tmp_path / "ch4.xyz")
    formats.write_xyz(path, Zm, xa, "methane")
    Z, xa2, _ = formats.read_xyz(path)
    xyz = molecules.to_bohr(xa2)
    for label, spherical in (("def2-svp", True), ("6-31g_st", False)):
        b = basis.load_basis(label)
        spec, n = basis.element_spec(b, Z), basis.number_of_functions(b, Z, spherical)
        P = random_density(n, 5, seed=22)
        e0, n0, V0, _ = oracle.reference().setup(2, Z, xyz, spec, spherical, ids("pbe")).integrate(P, 1)
        e, ne, V, _ = oracle.oracle().setup(2, Z, xyz, spec, spherical, ids("pbe")).integrate(P, 4)
        assert abs(e - e0) <= 1e-11 and abs(ne - n0) <= 1e-11, label
        assert np.max(np.abs(V - V0)) <= 1e-12, label
