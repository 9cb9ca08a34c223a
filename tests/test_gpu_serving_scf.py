"""The request loop with real SCF jobs (scf.RKSDFT: integrals, J fit and XC quadrature on the GPU, ODA/DIIS on the host).

- A stream of 16 requests, (H2O)N with N log-uniform in [4, 16], BLYP or PBE, served with the quadratures of up to 16 jobs
  batched: every SCF converges, and its energy equals the same SCF run alone within 1e-7 Eh (the SCF is iterated to
  1e-8 Eh and an RMS density change of 1e-7, so the two runs end at the same fixed point; their V_xc differ only by the
  order of FP64 atomic additions).
- The serving configuration at full size: the 256 requests of seed 3 ((H2O)4 to (H2O)64), each SCF to 1e-6 Eh, in both
  precisions: every request converges, and each request's mixed-precision energy agrees with its FP64 energy within
  1e-6 Eh.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_served_scf_equals_alone():
    from pdynamo3_amd import scf, serving
    reqs = serving.request_stream(n_requests=16, seed=21, min_waters=4, max_waters=16, rate=1.0e4)
    srv = serving.Server(precision="mixed", max_jobs=16, energy_tolerance=1e-8, density_tolerance=1e-7)
    finished, loop = srv.run(reqs)
    srv.close()
    summary = serving.summarize(finished, loop)
    print({k: summary[k] for k in ("requests", "converged", "scf_iterations_mean", "latency_s_p50", "latency_s_p99",
                                   "mean_batch", "max_batch", "failed")})
    assert len(finished) == 16 and summary["failed"] == 0 and summary["converged"] == 16
    assert loop["max_batch"] >= 4                  # SCF jobs shared quadrature launches
    worst = 0.0
    for job in sorted(finished, key=lambda j: j.request.rid):
        r = job.request
        alone = scf.RKSDFT(r.Z, r.xyz_bohr, functional=r.functional, precision="mixed", energy_tolerance=1e-8,
                           density_tolerance=1e-7)
        res = alone.run()
        alone.close()
        assert res["converged"]
        d = abs(res["energy"] - job.result["energy"])
        worst = max(worst, d)
        assert d <= 1e-7, (r.rid, res["energy"], job.result["energy"])
    print("max |E(served) - E(alone)|", worst)


ENERGIES = {}


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_serving_stream_scf(precision):
    from pdynamo3_amd import serving
    reqs = serving.request_stream(n_requests=256, seed=3, rate=1.0e4)
    srv = serving.Server(precision=precision, max_jobs=16)
    finished, loop = srv.run(reqs, log=_progress(precision))
    srv.close()
    summary = serving.summarize(finished, loop)
    print(precision, {k: summary[k] for k in ("requests", "converged", "scf_iterations_mean", "requests_per_s",
                                              "latency_s_p50", "latency_s_p99", "mean_batch", "failed")}, flush=True)
    assert len(finished) == 256 and summary["failed"] == 0 and summary["converged"] == 256
    ENERGIES[precision] = {j.request.rid: j.result["energy"] for j in finished}


def test_serving_stream_mixed_against_fp64():
    if len(ENERGIES) < 2:
        pytest.skip("needs both stream runs of this module")
    d = np.array([abs(ENERGIES["mixed"][r] - ENERGIES["fp64"][r]) for r in ENERGIES["fp64"]])
    print("per-request |E(mixed) - E(fp64)|: max", float(d.max()), "median", float(np.median(d)))
    assert d.max() <= 1e-6


def _progress(precision):
    """A log function for the request loop that reports every 32nd finished request on stderr."""
    import sys
    count = [0]

    def log(message):
        if "latency" in message:
            count[0] += 1
            if count[0] % 32 == 0:
                print("{:s}: {:d} requests done".format(precision, count[0]), file=sys.stderr, flush=True)
    return log
