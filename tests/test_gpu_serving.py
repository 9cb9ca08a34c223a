"""The request loop: every job's SCF energies are the same served in the mix as run alone, admission respects the HBM
budget, and jobs that arrive while others run are admitted and batched with them."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _stream(n, seed):
    from pdynamo3_amd import serving
    return serving.request_stream(n_requests=n, seed=seed, min_waters=2, max_waters=6, rate=2000.0)


def test_served_energies_equal_alone():
    from pdynamo3_amd import integrator, serving
    from pdynamo3_amd.definitions import functional_ids
    reqs = _stream(6, 11)
    srv = serving.Server(precision="mixed", max_jobs=4, hbm_budget=16 << 30, scf="synthetic")
    finished, loop = srv.run(reqs)
    srv.close()
    assert len(finished) == len(reqs)
    assert loop["max_batch"] >= 2          # jobs overlapped in one launch
    for job in finished:
        r = job.request
        g = integrator.QuadratureGrid(r.Z, r.xyz_bohr, 2)
        g.set_basis(r.spec, False, "mixed")
        scf = serving.SyntheticSCF(r.density, r.seed)
        alone = []
        for _ in range(job.quadratures):
            E, _, _ = g.integrate(functional_ids(r.functional), scf.density())
            scf.record(E)
            alone.append(E)
        g.close()
        assert alone == job.energies
        assert abs(job.energies[-1] - job.energies[-2]) < 1.0e-6 or job.quadratures == scf.max_iterations


def test_admission_defers_for_memory():
    from pdynamo3_amd import serving
    # Four equal jobs (3 waters each) and room for one and a half of them: they must run one at a time.
    reqs = serving.request_stream(n_requests=4, seed=12, min_waters=3, max_waters=3, rate=2000.0)
    need = max(serving.estimated_peak_bytes(r, "mixed") for r in reqs)
    srv = serving.Server(precision="mixed", max_jobs=4, hbm_budget=int(1.5 * need))
    finished, loop = srv.run(reqs)
    srv.close()
    assert len(finished) == len(reqs)
    assert loop["deferred_for_memory"] > 0
    assert loop["max_batch"] == 1


def test_request_larger_than_budget_is_refused():
    from pdynamo3_amd import integrator, serving
    reqs = _stream(1, 13)
    srv = serving.Server(precision="mixed", max_jobs=2, hbm_budget=1024)
    with pytest.raises(integrator.QCModelError):
        srv.run(reqs)
    srv.close()
