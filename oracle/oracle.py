"""ctypes access to the test oracle (liboracle.so, the FP64 C restatement) and to the compiled reference (_ref/libxcref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the product
package.  Both libraries expose the same operations:

    setup(accuracy, Z, xyz_bohr, spec, spherical, ids) -> handle
    npoints / nblocks / nfunctions
    grid()            -> (xyzw [G,4], block_atom [B], block_count [B])
    integrate(P)      -> (E_xc, N_e, V_packed, seconds)          (V starts at zero here)
    gradients(P)      -> XC nuclear gradients [nAtoms, 3] (from zero)
    gradients_uks(Pa, Pb) -> the same, spin unrestricted
    block_values(b)   -> (indices, f, fx, fy, fz)
    functional(ids, rho, sigma) -> (eps, vrho, vsigma)
    c2s(lLow, lHigh)  -> matrix
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libxcref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


class _Library:
    """One of the two CPU libraries; prefix is 'xqo_' (oracle) or 'xcref_' (reference)."""

    def __init__(self, path, prefix):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = ctypes.CDLL(path)
        self.prefix = prefix
        f = self._fn
        f("setup").restype = ctypes.c_void_p
        f("setup").argtypes = [ctypes.c_int, ctypes.c_int, _ip, _dp, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _ip, _ip]
        for name in ("npoints", "nblocks", "nfunctions"):
            f(name).restype = ctypes.c_int
            f(name).argtypes = [ctypes.c_void_p]
        f("grid").restype = None
        f("grid").argtypes = [ctypes.c_void_p, _dp, _ip, _ip]
        f("integrate").restype = ctypes.c_int
        f("integrate").argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, _dp, _dp, _dp, _dp]
        f("gradients").restype = ctypes.c_int
        f("gradients").argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, _dp]
        f("gradients_uks").restype = ctypes.c_int
        f("gradients_uks").argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, _dp]
        f("block_values").restype = ctypes.c_int
        f("block_values").argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _dp, _dp, _dp, _dp]
        f("functional").restype = ctypes.c_int
        f("functional").argtypes = [ctypes.c_int, _ip, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        f("integrate_uks").restype = ctypes.c_int
        f("integrate_uks").argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        f("functional_pol").restype = ctypes.c_int
        f("functional_pol").argtypes = [ctypes.c_int, _ip, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        f("c2s").restype = None
        f("c2s").argtypes = [ctypes.c_int, ctypes.c_int, _dp]
        f("free").restype = None
        f("free").argtypes = [ctypes.c_void_p]

    def _fn(self, name):
        return getattr(self.lib, self.prefix + name)

    def functional(self, ids, rho, sigma):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        sigma = np.ascontiguousarray(sigma, dtype=np.float64)
        n = rho.size
        e, vr, vs = np.zeros(n), np.zeros(n), np.zeros(n)
        st = self._fn("functional")(len(ids), _i(ids), n, _d(rho), _d(sigma), _d(e), _d(vr), _d(vs))
        if st != 0:
            raise ValueError("functional status {:d}".format(st))
        return e, vr, vs

    def functional_pol(self, ids, rho, sigma):
        """Spin-polarised: rho (n, 2), sigma (n, 3) -> eps (n), vrho (n, 2), vsigma (n, 3)."""
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1, 2)
        sigma = np.ascontiguousarray(sigma, dtype=np.float64).reshape(-1, 3)
        n = rho.shape[0]
        e, vr, vs = np.zeros(n), np.zeros((n, 2)), np.zeros((n, 3))
        st = self._fn("functional_pol")(len(ids), _i(ids), n, _d(rho), _d(sigma), _d(e), _d(vr), _d(vs))
        if st != 0:
            raise ValueError("functional status {:d}".format(st))
        return e, vr, vs

    def c2s(self, lLow, lHigh):
        nc = sum((l + 1) * (l + 2) // 2 for l in range(lLow, lHigh + 1))
        ns = sum(2 * l + 1 for l in range(lLow, lHigh + 1))
        out = np.zeros((nc, ns))
        self._fn("c2s")(lLow, lHigh, _d(out))
        return out

    def setup(self, accuracy, Z, xyz_bohr, spec, spherical, ids):
        return _Handle(self, accuracy, Z, xyz_bohr, spec, spherical, ids)


class _Handle:
    def __init__(self, lib, accuracy, Z, xyz_bohr, spec, spherical, ids):
        self.lib = lib
        self._Z = np.ascontiguousarray(Z, dtype=np.int32)
        self._xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1, 3)
        self._spec = np.ascontiguousarray(spec, dtype=np.float64)
        self._ids = np.ascontiguousarray(ids, dtype=np.int32)
        st = np.zeros(1, dtype=np.int32)
        self.h = lib._fn("setup")(int(accuracy), len(self._Z), _i(self._Z), _d(self._xyz), self._spec.size, _d(self._spec),
                                  int(bool(spherical)), len(self._ids), _i(self._ids), _i(st))
        if not self.h or st[0] != 0:
            raise ValueError("setup status {:d}".format(int(st[0])))
        self.npoints = lib._fn("npoints")(self.h)
        self.nblocks = lib._fn("nblocks")(self.h)
        self.nfunctions = lib._fn("nfunctions")(self.h)
        self.natoms = len(self._Z)

    def grid(self):
        xyzw = np.zeros((self.npoints, 4))
        ba = np.zeros(self.nblocks, dtype=np.int32)
        bc = np.zeros(self.nblocks, dtype=np.int32)
        self.lib._fn("grid")(self.h, _d(xyzw), _i(ba), _i(bc))
        return xyzw, ba, bc

    def integrate(self, P_packed, threads=1, V=None):
        """E_xc, N_e, V (packed; incremented from V if given, else from zero) and the call's wall time."""
        P = np.ascontiguousarray(P_packed, dtype=np.float64)
        n = self.nfunctions
        if P.size != n * (n + 1) // 2:
            raise ValueError("P has {:d} entries, expected {:d}".format(P.size, n * (n + 1) // 2))
        V = np.zeros_like(P) if V is None else np.ascontiguousarray(V, dtype=np.float64).copy()
        e, ne, sec = np.zeros(1), np.zeros(1), np.zeros(1)
        st = self.lib._fn("integrate")(self.h, _d(P), int(threads), _d(e), _d(ne), _d(V), _d(sec))
        if st != 0:
            raise ValueError("integrate status {:d}".format(st))
        return float(e[0]), float(ne[0]), V, float(sec[0])

    def integrate_uks(self, Pa_packed, Pb_packed, threads=1):
        """Unrestricted: E_xc, N_e, V_alpha, V_beta (packed, from zero) and the call's wall time."""
        n = self.nfunctions
        Pa = np.ascontiguousarray(Pa_packed, dtype=np.float64)
        Pb = np.ascontiguousarray(Pb_packed, dtype=np.float64)
        if Pa.size != n * (n + 1) // 2 or Pb.size != Pa.size:
            raise ValueError("P has {:d} entries, expected {:d}".format(Pa.size, n * (n + 1) // 2))
        Va, Vb = np.zeros_like(Pa), np.zeros_like(Pb)
        e, ne, sec = np.zeros(1), np.zeros(1), np.zeros(1)
        st = self.lib._fn("integrate_uks")(self.h, _d(Pa), _d(Pb), int(threads), _d(e), _d(ne), _d(Va), _d(Vb), _d(sec))
        if st != 0:
            raise ValueError("integrate_uks status {:d}".format(st))
        return float(e[0]), float(ne[0]), Va, Vb, float(sec[0])

    def gradients(self, P_packed, threads=1):
        """XC nuclear gradients (nAtoms x 3, hartree/bohr) of E_xc at the density P (the oracle's threads, or the reference's
        in-core flag, as the third argument)."""
        P = np.ascontiguousarray(P_packed, dtype=np.float64)
        n = self.nfunctions
        if P.size != n * (n + 1) // 2:
            raise ValueError("P has {:d} entries, expected {:d}".format(P.size, n * (n + 1) // 2))
        g = np.zeros((self.natoms, 3))
        st = self.lib._fn("gradients")(self.h, _d(P), int(threads), _d(g))
        if st != 0:
            raise ValueError("gradients status {:d}".format(st))
        return g

    def gradients_uks(self, Pa_packed, Pb_packed, threads=1):
        """Unrestricted XC nuclear gradients (nAtoms x 3) at the alpha and beta densities."""
        n = self.nfunctions
        Pa = np.ascontiguousarray(Pa_packed, dtype=np.float64)
        Pb = np.ascontiguousarray(Pb_packed, dtype=np.float64)
        if Pa.size != n * (n + 1) // 2 or Pb.size != Pa.size:
            raise ValueError("P has {:d} entries, expected {:d}".format(Pa.size, n * (n + 1) // 2))
        g = np.zeros((self.natoms, 3))
        st = self.lib._fn("gradients_uks")(self.h, _d(Pa), _d(Pb), int(threads), _d(g))
        if st != 0:
            raise ValueError("gradients_uks status {:d}".format(st))
        return g

    def block_values(self, block):
        n = self.nfunctions
        xyzw, ba, bc = self.grid() if not hasattr(self, "_bc") else (None, None, self._bc)
        self._bc = bc
        npb = int(bc[block])
        idx = np.zeros(n, dtype=np.int32)
        f, fx, fy, fz = (np.zeros(n * npb) for _ in range(4))
        nb = self.lib._fn("block_values")(self.h, int(block), _i(idx), _d(f), _d(fx), _d(fy), _d(fz))
        shape = (nb, npb)
        return idx[:nb].copy(), f[:nb * npb].reshape(shape), fx[:nb * npb].reshape(shape), \
            fy[:nb * npb].reshape(shape), fz[:nb * npb].reshape(shape)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib._fn("free")(self.h)
            self.h = None


def lebedev_points(n):
    """Unit vectors and weights of the oracle's Lebedev-Laikov rule with n points ([n, 3], [n])."""
    lib = oracle().lib
    fn = lib.xqo_lebedev_points
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
    x, y, z, w = (np.zeros(n) for _ in range(4))
    m = fn(n, _d(x), _d(y), _d(z), _d(w))
    if m != n:
        raise ValueError("no Lebedev rule with {:d} points".format(n))
    return np.stack([x, y, z], axis=1), w


def basis_at(handle, xyz):
    """All basis functions and gradients of an oracle handle at arbitrary points xyz [K, 3]: four [nF, K] arrays."""
    fn = handle.lib.lib.xqo_basis_at
    fn.restype = None
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    k, n = xyz.shape[0], handle.nfunctions
    out = [np.zeros(n * k) for _ in range(4)]
    fn(handle.h, k, _d(xyz), *[_d(o) for o in out])
    return [o.reshape(n, k) for o in out]


_oracle = None
_reference = None


def oracle():
    """The C restatement (built by `make -C oracle`)."""
    global _oracle
    if _oracle is None:
        _oracle = _Library(ORACLE_PATH, "xqo_")
    return _oracle


def reference():
    """The compiled reference path (built by `make -C oracle ref`); raises FileNotFoundError when absent."""
    global _reference
    if _reference is None:
        _reference = _Library(REF_PATH, "xcref_")
    return _reference


def reference_available():
    return os.path.exists(REF_PATH)


# . The compiled reference names its integrate "inCore" flag where the oracle takes a thread count; both are ints.


class RefBasis:
    """A basis-only handle of the compiled reference (orbital or fit basis on the given atoms), for the integrals of the
    SCF wrapper: one-electron matrices (f1KOf1i + f2Cm1V), electron-fit Coulomb integrals (f1Xg2i) and the constrained fit
    matrix (f1Xf1i_f1Oi)."""

    def __init__(self, Z, xyz_bohr, spec, spherical):
        lib = reference().lib
        self.lib = lib
        f = lib.xcref_basis_setup
        f.restype = ctypes.c_void_p
        f.argtypes = [ctypes.c_int, _ip, _dp, ctypes.c_int, _dp, ctypes.c_int, _ip]
        for name, args in (("xcref_one_electron", [ctypes.c_void_p, _dp, _dp, _dp]),
                           ("xcref_fit_integrals", [ctypes.c_void_p, ctypes.c_void_p, _dp, _ip]),
                           ("xcref_fit_matrix", [ctypes.c_void_p, _dp])):
            getattr(lib, name).restype = ctypes.c_int
            getattr(lib, name).argtypes = args
        lib.xcref_nfunctions.restype = ctypes.c_int
        lib.xcref_nfunctions.argtypes = [ctypes.c_void_p]
        lib.xcref_free.restype = None
        lib.xcref_free.argtypes = [ctypes.c_void_p]
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
        spec = np.ascontiguousarray(spec, dtype=np.float64)
        st = np.zeros(1, dtype=np.int32)
        self.h = f(Z.size, _i(Z), _d(xyz), spec.size, _d(spec), int(bool(spherical)), _i(st))
        if not self.h or st[0] != 0:
            raise ValueError("reference basis set-up failed (status {:d})".format(int(st[0])))
        self.n = lib.xcref_nfunctions(self.h)

    def one_electron(self):
        """(S, T, V) packed: overlap, kinetic energy and electron-nuclear attraction (point nuclei of charge Z)."""
        m = self.n * (self.n + 1) // 2
        S, T, V = np.zeros(m), np.zeros(m), np.zeros(m)
        if self.lib.xcref_one_electron(self.h, _d(S), _d(T), _d(V)) != 0:
            raise ValueError("reference one-electron integrals failed")
        return S, T, V

    def fit_integrals(self, fit):
        """Dense (mu nu|P) [packed orbital pair, fit function] as stored by the reference, and the stored count."""
        dense = np.zeros((self.n * (self.n + 1) // 2, fit.n))
        count = np.zeros(1, dtype=np.int32)
        if self.lib.xcref_fit_integrals(self.h, fit.h, _d(dense), _i(count)) != 0:
            raise ValueError("reference fit integrals failed")
        return dense, int(count[0])

    def fit_matrix(self):
        """The constrained fit matrix, packed, extent n + 1 (this handle is the fit basis)."""
        M = np.zeros((self.n + 1) * (self.n + 2) // 2)
        if self.lib.xcref_fit_matrix(self.h, _d(M)) != 0:
            raise ValueError("reference fit matrix failed")
        return M

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.xcref_free(self.h)
            self.h = None
