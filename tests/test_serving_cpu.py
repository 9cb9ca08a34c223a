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
