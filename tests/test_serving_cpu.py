"""Host-side parts of the request loop: the request stream, the synthetic SCF sequence and the admission estimate."""

import numpy as np


def test_request_stream_is_deterministic_and_in_range():
    from pdynamo3_amd import serving
    a = serving.request_stream(n_requests=12, seed=3)
    b = serving.request_stream(n_requests=12, seed=3)
    assert [r.n_waters for r in a] == [r.n_waters for r in b]
    assert all(4 <= r.n_waters <= 64 for r in a)
    assert all(r.functional in ("blyp", "pbe") for r in a)
    t = [r.arrival for r in a]
    assert t == sorted(t) and t[0] > 0.0
    for r in a[:3]:
        assert r.Z.size == 3 * r.n_waters
        assert r.density.size == r.n_functions * (r.n_functions + 1) // 2
        assert np.array_equal(r.density, b[r.rid].density)


def test_synthetic_scf_converges_geometrically():
    from pdynamo3_amd import serving
    P = np.linspace(0.1, 1.0, 55)
    scf = serving.SyntheticSCF(P, seed=5, rate=0.4, tolerance=1e-6)
    d0 = scf.density()
    assert not np.allclose(d0, P)
    # Feed an energy that is linear in the density error: the loop stops once two successive values differ < 1e-6.
    k = 0
    while not scf.record(float(np.sum(scf.density() - P))):
        k += 1
    assert 2 <= scf.iteration < scf.max_iterations
    assert abs(scf.energies[-1] - scf.energies[-2]) < 1e-6
    assert np.max(np.abs(scf.density() - P)) < np.max(np.abs(d0 - P)) * 0.4 ** (scf.iteration - 1)


def test_admission_estimate_grows_with_the_job():
    from pdynamo3_amd import serving
    reqs = serving.request_stream(n_requests=20, seed=3)
    small = min(reqs, key=lambda r: r.n_waters)
    large = max(reqs, key=lambda r: r.n_waters)
    assert serving.estimated_peak_bytes(small, "mixed") < serving.estimated_peak_bytes(large, "mixed")


def test_fit_integral_estimate_counts_every_pair_of_a_water():
    """qs_fit_estimate_bytes on a single water (no shell pair is screened out at these distances): the integrals hold one
    row per function pair of every shell pair A >= B (an SP shell counts as an S and a P shell), times the fit functions, plus the row maps and work arrays; a
    larger cluster needs more."""
    from pdynamo3_amd import basis, molecules, serving
    Z, xyz = molecules.water()
    b = basis.load_basis("6-31g_st")
    spec = basis.element_spec(b, Z)
    nf = []   # the integral code takes an SP shell as an S and a P shell
    for z in Z:
        for shell in b["elements"][str(int(z))]["shells"]:
            nf += [(l + 1) * (l + 2) // 2 for l in range(shell["lLow"], shell["lHigh"] + 1)]
    rows = sum(nf[a] * nf[c] for a in range(len(nf)) for c in range(a + 1))
    fb = basis.load_basis(serving.FIT_BASIS)
    n_fit = basis.number_of_functions(fb, Z, True)
    n = basis.number_of_functions(b, Z, False)
    expected = 8 * rows * n_fit + 9 * rows + 16 * (n * (n + 1) // 2) + 16 * n_fit + 8 * ((rows + 1023) // 1024) * n_fit
    assert serving.fit_integral_bytes(Z, xyz, spec, False) == expected
    reqs = serving.request_stream(n_requests=20, seed=3)
    small = min(reqs, key=lambda r: r.n_waters)
    large = max(reqs, key=lambda r: r.n_waters)
    assert serving.fit_integral_bytes(small.Z, small.xyz_bohr, small.spec, False) < \
        serving.fit_integral_bytes(large.Z, large.xyz_bohr, large.spec, False)
