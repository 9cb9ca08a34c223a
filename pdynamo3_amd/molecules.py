"""Synthetic inputs: the rigid water molecule and random water clusters.

Water: O at the origin, H at (+-0.757, 0.586, 0) Angstrom.  Clusters: each water is rotated by a uniformly random rotation
(random unit quaternion) and centred (its O atom) uniformly in a cube; a placement is rejected if any O-O distance is below
the minimum.  Random numbers come from numpy.random.default_rng(seed).  The reference has no generator for these systems.
"""

import numpy as np

from .definitions import ANGSTROM_TO_BOHR

WATER_Z = np.array([8, 1, 1], dtype=np.int32)
WATER_XYZ_ANGSTROM = np.array([[0.0, 0.0, 0.0], [0.757, 0.586, 0.0], [-0.757, 0.586, 0.0]], dtype=np.float64)


def random_rotation(rng):
    """Rotation matrix of a uniformly random unit quaternion."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def water_cluster(n_waters, cube_angstrom, seed, min_oo_angstrom=2.6, max_tries=1000000):
    """Atomic numbers (int32), coordinates in Angstrom and the per-water rotations of a random water cluster."""
    rng = np.random.default_rng(seed)
    centres, rotations = [], []
    tries = 0
    while len(centres) < n_waters:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("Could not place {:d} waters in a {:.1f} A cube.".format(n_waters, cube_angstrom))
        rot = random_rotation(rng)
        c = rng.uniform(0.0, cube_angstrom, size=3)
        if centres and np.min(np.linalg.norm(np.asarray(centres) - c, axis=1)) < min_oo_angstrom:
            continue
        centres.append(c)
        rotations.append(rot)
    z = np.tile(WATER_Z, n_waters)
    xyz = np.concatenate([WATER_XYZ_ANGSTROM @ r.T + c for r, c in zip(rotations, centres)])
    return z, xyz, np.asarray(rotations), np.asarray(centres)


def to_bohr(xyz_angstrom):
    """Angstrom -> bohr with the constant the reference applies to QC coordinates."""
    return np.ascontiguousarray(np.asarray(xyz_angstrom, dtype=np.float64) * ANGSTROM_TO_BOHR)


def water():
    """A single water: atomic numbers and bohr coordinates."""
    return WATER_Z.copy(), to_bohr(WATER_XYZ_ANGSTROM)
