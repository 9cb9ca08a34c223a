"""The SCF's dense steps on the GPU (include/qcscf.h, qs_dense_*; scf.DeviceDenseSteps) against numpy / LAPACK:

- on a random positive definite S (600 functions), its symmetric orthogonaliser X = S^-1/2 and a random symmetric F: the
  orbital energies equal numpy's eigenvalues of X^T F X to 1e-10 relative, P = 2 C_occ C_occ^T to 1e-10 (P does not
  depend on the eigenvectors' signs), and the commutator X^T (F P S - S P F) X at a random symmetric P (the aufbau
  density of F itself commutes with F) to 1e-10 relative;
- the closed forms: no occupied orbital gives P = 0, all n occupied give P = 2 S^-1 (C C^T = X X^T = S^-1);
- a restricted SCF on (H2O)32 (608 functions) with the dense steps on the device converges to the energy of the same SCF
  with them on the host within 1e-8 Eh, in the same number of iterations.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(n, seed):
    from pdynamo3_amd import scf
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) / np.sqrt(n)
    S = A @ A.T + np.eye(n)
    s, U = np.linalg.eigh(S)
    X = (U / np.sqrt(s)) @ U.T
    F = rng.standard_normal((n, n))
    F = 0.5 * (F + F.T)
    return S, X, F, scf


def test_dense_steps_match_numpy():
    n, nocc = 600, 137
    S, X, F, scf = _system(n, 5)
    d = scf.DeviceDenseSteps(X, scf.pack(S))
    P, e = d.aufbau(scf.pack(F), nocc)
    e_ref, C = np.linalg.eigh(X.T @ F @ X)
    Cocc = X @ C[:, :nocc]
    P_ref = 2.0 * Cocc @ Cocc.T
    assert np.max(np.abs(e - e_ref)) <= 1e-10 * np.max(np.abs(e_ref))
    assert np.max(np.abs(scf.unpack(P, n) - P_ref)) <= 1e-10 * np.max(np.abs(P_ref))
    # The commutator at a density that is not F's own (which commutes with F): a random symmetric matrix.
    R = np.random.default_rng(7).standard_normal((n, n))
    Pd = 0.5 * (R + R.T)
    err = d.commutator(scf.pack(F), scf.pack(Pd))
    FPS = F @ Pd @ S
    err_ref = X.T @ (FPS - FPS.T) @ X
    print("max|de|", np.max(np.abs(e - e_ref)), "max|dP|", np.max(np.abs(scf.unpack(P, n) - P_ref)), "max|d err|",
          np.max(np.abs(err - err_ref)), "max|err|", np.max(np.abs(err_ref)))
    assert np.max(np.abs(err - err_ref)) <= 1e-10 * np.max(np.abs(err_ref))
    d.close()


def test_dense_steps_closed_forms():
    n = 257
    S, X, F, scf = _system(n, 6)
    d = scf.DeviceDenseSteps(X, scf.pack(S))
    P0, _ = d.aufbau(scf.pack(F), 0)
    assert np.max(np.abs(P0)) == 0.0
    Pn, _ = d.aufbau(scf.pack(F), n)
    two_inv = 2.0 * np.linalg.inv(S)
    assert np.max(np.abs(scf.unpack(Pn, n) - two_inv)) <= 1e-10 * np.max(np.abs(two_inv))
    d.close()


def test_scf_with_device_dense_steps_equals_host():
    from pdynamo3_amd import molecules, scf
    Z, xa, _, _ = molecules.water_cluster(32, 11.0, 3)
    xyz = molecules.to_bohr(xa)
    out = {}
    for mode in ("device", "host"):
        job = scf.RKSDFT(Z, xyz, functional="blyp", precision="mixed", dense=mode)
        assert (job.dense is not None) == (mode == "device")
        res = job.run()
        job.close()
        out[mode] = res
        print(mode, res["energy"], res["iterations"], res["timings"])
    assert out["device"]["converged"] and out["host"]["converged"]
    assert abs(out["device"]["energy"] - out["host"]["energy"]) <= 1e-8
    assert out["device"]["iterations"] == out["host"]["iterations"]
