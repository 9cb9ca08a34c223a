"""The serving configuration at full size: 256 SCF requests on (H2O)N clusters, N log-uniform in [4, 64] (seed 3),
BLYP or PBE by coin flip, Poisson arrivals faster than the loop serves them, through the request loop in each precision.

Every request's sequence of E_xc values served in the mix equals the sequence it gives alone on a fresh handle, bit for
bit, and so does its last N_e; its last V_xc agrees within 1e-12 (the FP64 atomic additions of the Fock stage come in
another order in a batch).  The FP64 loop's final energies agree with the mixed loop's within 1e-6 Eh per request."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FINAL = {}


def _alone(request, precision, quadratures):
    from pdynamo3_amd import integrator, serving
    from pdynamo3_amd.definitions import functional_ids
    g = integrator.QuadratureGrid(request.Z, request.xyz_bohr, 2)
    g.set_basis(request.spec, False, precision)
    scf = serving.SyntheticSCF(request.density, request.seed)
    energies, N, V = [], None, None
    for _ in range(quadratures):
        V = np.zeros_like(request.density)
        E, N, _ = g.integrate(functional_ids(request.functional), scf.density(), V)
        scf.record(E)
        energies.append(E)
    g.close()
    return energies, N, V


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_stream_results_equal_alone(precision):
    from pdynamo3_amd import serving
    reqs = serving.request_stream(n_requests=256, seed=3, rate=200.0)
    srv = serving.Server(precision=precision, max_jobs=16, keep_fock=True, scf="synthetic")
    finished, loop = srv.run(reqs)
    srv.close()
    summary = serving.summarize(finished, loop)
    print(precision, {k: summary[k] for k in ("requests", "latency_s_p50", "latency_s_p99", "requests_per_s",
                                              "mean_batch", "max_batch", "failed")})
    assert len(finished) == 256 and summary["failed"] == 0
    assert loop["max_batch"] >= 8                  # requests overlapped on the device
    worst = 0.0
    for job in sorted(finished, key=lambda j: j.request.rid):
        energies, N, V = _alone(job.request, precision, job.quadratures)
        assert energies == job.energies, job.request.rid
        assert N == job.electrons
        worst = max(worst, float(np.max(np.abs(V - job.fock))))
        assert job.scf.iteration == job.quadratures
        assert abs(job.energies[-1] - job.energies[-2]) < 1e-6 or job.quadratures == job.scf.max_iterations
    print(precision, "max |dV| alone vs mix", worst)
    assert worst <= 1e-12
    FINAL[precision] = {j.request.rid: j.energies[-1] for j in finished}


def test_stream_fp64_against_mixed():
    if len(FINAL) < 2:
        pytest.skip("needs both stream runs of this module")
    d = max(abs(FINAL["fp64"][r] - FINAL["mixed"][r]) for r in FINAL["fp64"])
    print("max |E_xc(fp64) - E_xc(mixed)| over the final energies", d)
    assert d <= 1e-6
