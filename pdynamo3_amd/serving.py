"""Request loop for the XC quadratures of SCF jobs on one GPU: continuous batching with admission by free HBM.

A request is an SCF job on one molecule.  Its grid, screened basis values and density stay on the device for the whole
job; each SCF iteration needs one XC quadrature (the rest of an SCF step -- integrals, J, diagonalisation -- is outside
this library and is represented by the density sequence the job supplies).  The loop:

  * takes requests as they arrive (arrival times are part of the request stream);
  * admits a pending request when a slot is free and its estimated peak device bytes fit under the HBM budget
    (xq_estimate_bytes against the budget, less what admitted jobs hold);
  * builds an admitted job's grid and basis values on a set-up thread, on the job's own stream, while the GPU runs the
    quadratures of the jobs already set up;
  * runs one batched quadrature per tick over every job that is set up (xq_integrate_batch: one launch per stage for all
    of them), then advances each job's SCF;
  * retires finished jobs at once, which frees their slot and their device memory for the next pending request.

Latency is completion minus arrival.  Each job's result does not depend on which other jobs share its launches
(tests/test_gpu_batch.py, tests/test_gpu_serving.py).

The density sequence of a job (SyntheticSCF) converges geometrically from a perturbed start to the job's density,
P_k = P* + r^k (P_0 - P*), and the job stops when two successive quadratures differ by less than the convergence
tolerance in E_xc, as an SCF stops on its energy change.
"""

import concurrent.futures
import time
from dataclasses import dataclass, field

import numpy as np

from . import basis as _basis
from . import densities, integrator, molecules
from .definitions import functional_ids


@dataclass
class Request:
    """One SCF job: a water cluster, its functional and its density sequence."""
    rid: int
    arrival: float                  # seconds after the start of the stream
    n_waters: int
    functional: str
    Z: np.ndarray
    xyz_bohr: np.ndarray
    density: np.ndarray             # converged packed density P*
    spec: np.ndarray
    n_functions: int
    seed: int


class SyntheticSCF:
    """Density sequence P_k = P* + r^k (P_0 - P*) with P_0 = P* scaled entrywise by (1 + a u), u symmetric in [-1, 1]."""

    def __init__(self, density, seed, rate=0.4, amplitude=0.05, tolerance=1.0e-6, max_iterations=30):
        rng = np.random.default_rng(seed)
        self.target = density
        self.delta = density * amplitude * rng.uniform(-1.0, 1.0, size=density.size)
        self.rate = rate
        self.tolerance = tolerance
        self.max_iterations = max_iterations
        self.iteration = 0
        self.energies = []

    def density(self):
        return self.target + (self.rate ** self.iteration) * self.delta

    def record(self, energy):
        """Record this iteration's E_xc; True when the job has converged (or reached its iteration limit)."""
        self.energies.append(energy)
        self.iteration += 1
        if self.iteration >= self.max_iterations:
            return True
        return len(self.energies) >= 2 and abs(self.energies[-1] - self.energies[-2]) < self.tolerance


@dataclass
class Job:
    request: Request
    scf: SyntheticSCF
    reserved: int
    grid: object = None
    admitted: float = 0.0
    ready: float = 0.0
    done: float = 0.0
    points: int = 0
    quadratures: int = 0
    energies: list = field(default_factory=list)
    error: str = ""
    electrons: float = 0.0          # the last quadrature's N_e
    fock: object = None             # the last quadrature's V_xc (packed), kept when the server is asked to


def request_stream(n_requests=256, seed=3, min_waters=4, max_waters=64, rate=50.0, spherical=False):
    """Water-cluster SCF requests: N waters log-uniform in [min, max] (rounded), BLYP or PBE by coin flip, uniform random
    orientations and centres in a cube of side 9 (N/16)^(1/3) angstrom, Poisson arrivals at `rate` per second; all draws
    from numpy default_rng(seed)."""
    rng = np.random.default_rng(seed)
    b = _basis.load_basis("6-31g_st")
    t = 0.0
    out = []
    for rid in range(n_requests):
        t += rng.exponential(1.0 / rate)
        n_w = int(np.rint(np.exp(rng.uniform(np.log(min_waters), np.log(max_waters)))))
        functional = "blyp" if rng.uniform() < 0.5 else "pbe"
        cube = 9.0 * (n_w / 16.0) ** (1.0 / 3.0)
        cseed = int(rng.integers(1 << 30))
        Z, xa, rots, _ = molecules.water_cluster(n_w, cube, cseed)
        out.append(Request(rid=rid, arrival=t, n_waters=n_w, functional=functional, Z=Z,
                           xyz_bohr=molecules.to_bohr(xa), density=densities.cluster_density(rots, spherical=spherical),
                           spec=_basis.element_spec(b, Z), n_functions=_basis.number_of_functions(b, Z, spherical),
                           seed=cseed))
    return out


def estimated_peak_bytes(request, precision):
    """Peak device bytes of a request, set-up included: the grid size before the weight cull and at most 256 kept
    functions per block (an upper bound for water clusters in 6-31G*)."""
    points = integrator.estimated_points(2, request.Z)
    return integrator.estimate_bytes(points, request.n_functions, float(min(request.n_functions, 256)), precision)


class Server:
    """The request loop.  max_jobs: jobs set up or running at once (the batch size); hbm_budget: device bytes the loop may
    reserve (default 90% of the free memory when it starts)."""

    def __init__(self, precision="mixed", max_jobs=16, hbm_budget=None, spherical=False, keep_fock=False):
        self.precision = precision
        self.keep_fock = keep_fock          # keep each job's last V_xc (packed) on the finished job
        self.max_jobs = max_jobs
        self.spherical = spherical
        if hbm_budget is None:
            free, _ = integrator.device_memory()
            hbm_budget = int(0.9 * free)
        self.hbm_budget = int(hbm_budget)
        self.batch = integrator.QuadratureBatch(max_jobs)
        self.setup_pool = concurrent.futures.ThreadPoolExecutor(max_workers=1)

    def _set_up(self, job):
        r = job.request
        g = integrator.QuadratureGrid(r.Z, r.xyz_bohr, 2)
        try:
            g.set_basis(r.spec, self.spherical, self.precision)
        except Exception:
            g.close()
            raise
        return g

    def run(self, requests, log=None):
        """Serve the requests (sorted by arrival); returns the finished jobs and loop statistics."""
        pending = sorted(requests, key=lambda r: r.arrival)
        queue, setting_up, active, finished, failed = [], {}, [], [], []
        reserved = 0
        deferred_for_memory = 0
        batch_sizes = []
        t_batch = t_device = t_idle = 0.0
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        while pending or queue or setting_up or active:
            t = now()
            while pending and pending[0].arrival <= t:
                queue.append(pending.pop(0))
            # Admission: free slot and room under the HBM budget, in arrival order.
            while queue and len(setting_up) + len(active) < self.max_jobs:
                need = estimated_peak_bytes(queue[0], self.precision)
                if reserved + need > self.hbm_budget:
                    if setting_up or active:
                        deferred_for_memory += 1
                        break
                    if need > self.hbm_budget:
                        raise integrator.QCModelError("Request {:d} needs {:d} bytes; the HBM budget is {:d}.".format(
                            queue[0].rid, need, self.hbm_budget))
                r = queue.pop(0)
                job = Job(request=r, scf=SyntheticSCF(r.density, r.seed), reserved=need, admitted=now())
                reserved += need
                setting_up[self.setup_pool.submit(self._set_up, job)] = job
            for fut in [f for f in setting_up if f.done()]:
                job = setting_up.pop(fut)
                try:
                    job.grid = fut.result()
                except Exception as exc:
                    # A set-up that fails (e.g. out of device memory) ends its request only: its reservation and slot
                    # are released and the loop goes on.
                    reserved -= job.reserved
                    job.error = repr(exc)
                    job.done = now()
                    failed.append(job)
                    if log:
                        log("request {:d}: set-up failed: {:s}".format(job.request.rid, job.error))
                    continue
                job.points = job.grid.number_of_points
                job.ready = now()
                active.append(job)
            if not active:
                # Nothing to run: wait for a set-up to finish or the next arrival.
                t_wait = now()
                if setting_up:
                    concurrent.futures.wait(list(setting_up), timeout=0.05,
                                            return_when=concurrent.futures.FIRST_COMPLETED)
                elif pending:
                    time.sleep(max(0.0, min(0.05, pending[0].arrival - now())))
                t_idle += now() - t_wait
                continue
            # One batched quadrature over every job that is set up.
            batch_sizes.append(len(active))
            focks = [np.zeros(j.request.n_functions * (j.request.n_functions + 1) // 2) if self.keep_fock else None
                     for j in active]
            work = [(j.grid, functional_ids(j.request.functional), j.scf.density(), V) for j, V in zip(active, focks)]
            t_call = now()
            results, stats = self.batch.integrate(work)
            t_batch += now() - t_call
            t_device += 1e-3 * stats["device_ms"]
            still = []
            for job, (E, N), V in zip(active, results, focks):
                job.quadratures += 1
                job.energies.append(E)
                job.electrons = N
                job.fock = V
                if job.scf.record(E):
                    job.done = now()
                    job.grid.close()
                    job.grid = None
                    reserved -= job.reserved
                    finished.append(job)
                    if log:
                        log("request {:d} ({:d} waters, {:s}): {:d} quadratures, latency {:.3f} s".format(
                            job.request.rid, job.request.n_waters, job.request.functional, job.quadratures,
                            job.done - job.request.arrival))
                else:
                    still.append(job)
            active = still
        wall = now()
        return finished, {"wall_s": wall, "batches": len(batch_sizes), "failed": failed,
                          "mean_batch": float(np.mean(batch_sizes)) if batch_sizes else 0.0,
                          "max_batch": int(max(batch_sizes)) if batch_sizes else 0,
                          "deferred_for_memory": deferred_for_memory, "batch_call_s": t_batch,
                          "batch_device_s": t_device, "idle_s": t_idle}

    def close(self):
        self.setup_pool.shutdown(wait=True)
        self.batch.close()


def summarize(finished, loop):
    loop = dict(loop)
    failed = loop.pop("failed", [])
    lat = np.array([j.done - j.request.arrival for j in finished])
    wait = np.array([j.admitted - j.request.arrival for j in finished])
    points = sum(j.points * j.quadratures for j in finished)
    return {"requests": len(finished), "latency_s_p50": float(np.percentile(lat, 50)),
            "latency_s_p99": float(np.percentile(lat, 99)), "queue_wait_s_p50": float(np.percentile(wait, 50)),
            "requests_per_s": len(finished) / loop["wall_s"], "grid_points_per_s": points / loop["wall_s"],
            "quadratures": int(sum(j.quadratures for j in finished)), "failed": len(failed), **loop}
