"""Request loop for SCF jobs on one GPU: continuous batching of their XC quadratures with admission by free HBM.

A request is a Kohn-Sham SCF job on one molecule (scf.RKSDFT: one-electron and density-fitted Coulomb integrals on the
GPU, the XC quadrature on the GPU, ODA then DIIS on the host).  Its grid, screened basis values, electron-fit integrals and
density stay on the device for the whole job; each SCF iteration needs one XC quadrature.  The loop:

  * takes requests as they arrive (arrival times are part of the request stream);
  * admits a pending request when a slot is free and its estimated peak device bytes fit under the HBM budget
    (xq_estimate_bytes and the fit integrals' bound against the budget, less what admitted jobs hold);
  * sets up an admitted job (grid, basis values, integrals, orthogonaliser) on set-up threads, on the job's own streams,
    while the GPU runs the quadratures of the jobs already set up;
  * runs one batched quadrature (xq_integrate_batch: one launch per stage for all of them) over every job whose next
    density is ready, and hands each result to a host worker, which builds the job's Fock matrix, tests convergence and
    chooses its next density while the GPU serves the other jobs; a job joins the next batch as soon as its worker is done;
  * retires a job as soon as its SCF has converged, which frees its slot and its device memory for the next request.

Latency is completion minus arrival.  A job's quadratures do not depend on which other jobs share their launches
(tests/test_gpu_batch.py): E_xc and N_e are bit for bit those of the job alone, V_xc equal up to the order of FP64 atomic
additions, so a served SCF ends at its energy alone within its convergence tolerance (tests/test_gpu_serving_scf.py).

scf="synthetic" replaces the SCF by a fixed density sequence (SyntheticSCF) that does not feed V_xc back: a job's E_xc
sequence is then the same bit for bit alone and served, which tests/test_gpu_serving_stream.py checks on the full stream.
"""

import concurrent.futures
import os
import time
from dataclasses import dataclass, field

import numpy as np

from . import basis as _basis
from . import densities, integrator, molecules
from . import scf as scf_module
from .definitions import functional_ids

FIT_BASIS = "dgauss-a1-dftjfit"


@dataclass
class Request:
    """One SCF job: a water cluster and its functional (and, for the synthetic sequence, a target density)."""
    rid: int
    arrival: float                  # seconds after the start of the stream
    n_waters: int
    functional: str
    Z: np.ndarray
    xyz_bohr: np.ndarray
    density: np.ndarray             # packed target density P* of the synthetic sequence
    spec: np.ndarray
    n_functions: int
    seed: int
    peak_bytes: dict = field(default_factory=dict)   # admission estimates by (precision, scf), computed once


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
        self.converged = len(self.energies) >= 2 and abs(self.energies[-1] - self.energies[-2]) < self.tolerance
        return self.converged or self.iteration >= self.max_iterations

    def advance(self, e_xc, v_xc):
        """The SCF state machine's step (scf.RKSDFT.advance): the sequence does not use V_xc."""
        return self.record(e_xc)

    def summary(self):
        return {"energy": self.energies[-1] if self.energies else None, "converged": self.converged,
                "iterations": self.iteration}


@dataclass
class Job:
    request: Request
    reserved: int
    scf: object = None              # scf.RKSDFT, or SyntheticSCF
    grid: object = None
    admitted: float = 0.0
    ready: float = 0.0
    done: float = 0.0
    points: int = 0
    quadratures: int = 0
    energies: list = field(default_factory=list)      # E_xc of every quadrature
    error: str = ""
    electrons: float = 0.0          # the last quadrature's N_e
    fock: object = None             # the last quadrature's V_xc (packed), kept when the server is asked to
    result: dict = field(default_factory=dict)        # the SCF's summary when it ends
    host_s: float = 0.0             # host seconds spent in the job's SCF steps


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


_FIT_SPEC = {}


def fit_integral_bytes(Z, xyz_bohr, spec, spherical):
    """Device bytes of a job's electron-fit integrals and their work arrays (qs_fit_estimate_bytes: the integral set-up's
    own pair screen, evaluated on the host)."""
    import ctypes
    from . import _lib
    key = tuple(sorted(set(int(z) for z in Z)))
    if key not in _FIT_SPEC:
        _FIT_SPEC[key] = _lib.f64(_basis.element_spec(_basis.load_basis(FIT_BASIS), Z))
    fspec = _FIT_SPEC[key]
    z, x, o = _lib.i32(Z), _lib.f64(xyz_bohr), _lib.f64(spec)
    out = ctypes.c_longlong()
    status = _lib.lib.qs_fit_estimate_bytes(z.size, _lib.iptr(z), _lib.dptr(x), o.size, _lib.dptr(o), int(bool(spherical)),
                                            fspec.size, _lib.dptr(fspec), 1, ctypes.byref(out))
    if status != 0:
        raise integrator.QCModelError("Error in the fit estimate (status {:d}, {:s}).".format(status, _lib.last_error()))
    return out.value


def estimated_peak_bytes(request, precision, scf="rks", spherical=False):
    """Peak device bytes of a request, set-up included: the quadrature's (the grid size before the weight cull and at
    most 256 kept functions per block, an upper bound for water clusters in 6-31G*) and, for an SCF, its electron-fit
    integrals and their work arrays.  Computed once per request."""
    key = (precision, scf, bool(spherical))
    if key not in request.peak_bytes:
        points = integrator.estimated_points(2, request.Z)
        b = integrator.estimate_bytes(points, request.n_functions, float(min(request.n_functions, 256)), precision)
        if scf == "rks":
            b += fit_integral_bytes(request.Z, request.xyz_bohr, request.spec, spherical)
            n = request.n_functions
            if n >= scf_module.DEVICE_DENSE_MIN_FUNCTIONS:   # the SCF's dense steps on the device: 6 n^2 + n^2 / 2 doubles
                b += 8 * (6 * n * n + n * (n + 1) // 2 + 2 * n)
        request.peak_bytes[key] = b
    return request.peak_bytes[key]


class Server:
    """The request loop.

    precision      'mixed' or 'fp64' quadratures
    max_jobs       jobs set up or running at once (the largest batch)
    hbm_budget     device bytes the loop may reserve (default 90% of the free memory when it starts, less 16 GiB)
    scf            'rks' (each request runs scf.RKSDFT to convergence) or 'synthetic' (SyntheticSCF)
    host_workers   threads for the jobs' SCF steps (default min(16, cores)); BLAS runs one thread per step meanwhile
    setup_workers  threads for the jobs' set-up
    keep_fock      keep each job's last V_xc (packed) on the finished job
    tolerances     the SCF's energy and RMS density tolerances (Eh; 1e-6 and 1e-6)
    """

    def __init__(self, precision="mixed", max_jobs=16, hbm_budget=None, spherical=False, keep_fock=False, scf="rks",
                 host_workers=None, setup_workers=4, energy_tolerance=1.0e-6, density_tolerance=1.0e-6):
        if scf not in ("rks", "synthetic"):
            raise ValueError("scf must be 'rks' or 'synthetic'")
        self.precision = precision
        self.keep_fock = keep_fock
        self.max_jobs = max_jobs
        self.spherical = spherical
        self.scf_kind = scf
        self.energy_tolerance = energy_tolerance
        self.density_tolerance = density_tolerance
        if hbm_budget is None:
            # 90% of the free memory, less room for the runtime's own allocations: the private memory of the SCF
            # integral kernels (a few GB per launch) is allocated by the runtime at dispatch.
            free, _ = integrator.device_memory()
            hbm_budget = max(int(0.9 * free) - (16 << 30), int(0.5 * free))
        self.hbm_budget = int(hbm_budget)
        self.batch = integrator.QuadratureBatch(max_jobs)
        self.setup_pool = concurrent.futures.ThreadPoolExecutor(max_workers=setup_workers)
        self.host_workers = host_workers or min(16, os.cpu_count() or 1)
        self.host_pool = concurrent.futures.ThreadPoolExecutor(max_workers=self.host_workers)

    def _set_up(self, job):
        r = job.request
        g = integrator.QuadratureGrid(r.Z, r.xyz_bohr, 2)
        try:
            g.set_basis(r.spec, self.spherical, self.precision)
            if self.scf_kind == "synthetic":
                return g, SyntheticSCF(r.density, r.seed)
            from .scf import RKSDFT
            return g, RKSDFT(r.Z, r.xyz_bohr, spherical=self.spherical, functional=r.functional,
                             precision=self.precision, grid=g, energy_tolerance=self.energy_tolerance,
                             density_tolerance=self.density_tolerance)
        except Exception:
            g.close()
            raise

    def _step(self, job, e_xc, v_xc):
        """One SCF step of a job on a host worker; True when its SCF has ended."""
        t = time.perf_counter()
        done = job.scf.advance(e_xc, v_xc)
        job.host_s += time.perf_counter() - t
        return done

    def _retire(self, job):
        if job.scf is not None and hasattr(job.scf, "close"):
            job.scf.close()           # closes the grid and the integrals
        elif job.grid is not None:
            job.grid.close()
        job.grid = None

    def run(self, requests, log=None):
        """Serve the requests (sorted by arrival); returns the finished jobs and loop statistics."""
        from threadpoolctl import threadpool_limits
        with threadpool_limits(limits=1 if self.host_workers > 1 else None):
            return self._run(requests, log)

    def _run(self, requests, log):
        pending = sorted(requests, key=lambda r: r.arrival)
        queue, setting_up, ready, stepping, finished, failed = [], {}, [], {}, [], []
        reserved = 0
        deferred_for_memory = 0
        batch_sizes = []
        t_batch = t_device = t_idle = 0.0
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0

        def fail(job, exc, what):
            nonlocal reserved
            reserved -= job.reserved
            job.error = repr(exc)
            job.done = now()
            failed.append(job)
            if log:
                log("request {:d}: {:s} failed: {:s}".format(job.request.rid, what, job.error))

        while pending or queue or setting_up or ready or stepping:
            t = now()
            while pending and pending[0].arrival <= t:
                queue.append(pending.pop(0))
            # Admission: free slot and room under the HBM budget, in arrival order.
            while queue and len(setting_up) + len(ready) + len(stepping) < self.max_jobs:
                need = estimated_peak_bytes(queue[0], self.precision, self.scf_kind, self.spherical)
                if reserved + need > self.hbm_budget:
                    if setting_up or ready or stepping:
                        deferred_for_memory += 1
                        break
                    if need > self.hbm_budget:
                        raise integrator.QCModelError("Request {:d} needs {:d} bytes; the HBM budget is {:d}.".format(
                            queue[0].rid, need, self.hbm_budget))
                r = queue.pop(0)
                job = Job(request=r, reserved=need, admitted=now())
                reserved += need
                setting_up[self.setup_pool.submit(self._set_up, job)] = job
            for fut in [f for f in setting_up if f.done()]:
                job = setting_up.pop(fut)
                try:
                    job.grid, job.scf = fut.result()
                except Exception as exc:
                    # A set-up that fails (e.g. out of device memory) ends its request only: its reservation and slot
                    # are released and the loop goes on.
                    fail(job, exc, "set-up")
                    continue
                job.points = job.grid.number_of_points
                job.ready = now()
                ready.append(job)
            for fut in [f for f in stepping if f.done()]:
                job = stepping.pop(fut)
                try:
                    ended = fut.result()
                except Exception as exc:
                    self._retire(job)
                    fail(job, exc, "SCF step")
                    continue
                if ended:
                    job.result = job.scf.summary()
                    job.done = now()
                    self._retire(job)
                    reserved -= job.reserved
                    finished.append(job)
                    if log:
                        log("request {:d} ({:d} waters, {:s}): {:d} SCF iterations, latency {:.3f} s".format(
                            job.request.rid, job.request.n_waters, job.request.functional, job.quadratures,
                            job.done - job.request.arrival))
                else:
                    ready.append(job)
            if not ready:
                # Nothing to run: wait for a set-up or an SCF step to finish, or for the next arrival.
                t_wait = now()
                busy = list(setting_up) + list(stepping)
                if busy:
                    concurrent.futures.wait(busy, timeout=0.05, return_when=concurrent.futures.FIRST_COMPLETED)
                elif pending:
                    time.sleep(max(0.0, min(0.05, pending[0].arrival - now())))
                t_idle += now() - t_wait
                continue
            # One batched quadrature over every job whose next density is ready.
            batch, ready = ready, []
            batch_sizes.append(len(batch))
            keep_v = self.keep_fock or self.scf_kind == "rks"
            focks = [np.zeros(j.request.n_functions * (j.request.n_functions + 1) // 2) if keep_v else None
                     for j in batch]
            work = [(j.grid, functional_ids(j.request.functional), j.scf.density(), V) for j, V in zip(batch, focks)]
            t_call = now()
            results, stats = self.batch.integrate(work)
            t_batch += now() - t_call
            t_device += 1e-3 * stats["device_ms"]
            for job, (E, N), V in zip(batch, results, focks):
                job.quadratures += 1
                job.energies.append(E)
                job.electrons = N
                job.fock = V if self.keep_fock else None
                if self.scf_kind == "synthetic":
                    # The sequence's step is a comparison: run it here, so that the job rejoins the next batch.
                    stepping[_done_future(self._step(job, E, V))] = job
                else:
                    stepping[self.host_pool.submit(self._step, job, E, V)] = job
        wall = now()
        return finished, {"wall_s": wall, "batches": len(batch_sizes), "failed": failed,
                          "mean_batch": float(np.mean(batch_sizes)) if batch_sizes else 0.0,
                          "max_batch": int(max(batch_sizes)) if batch_sizes else 0,
                          "deferred_for_memory": deferred_for_memory, "batch_call_s": t_batch,
                          "batch_device_s": t_device, "idle_s": t_idle, "scf": self.scf_kind,
                          "host_workers": self.host_workers}

    def close(self):
        self.setup_pool.shutdown(wait=True)
        self.host_pool.shutdown(wait=True)
        self.batch.close()


def _done_future(value):
    f = concurrent.futures.Future()
    f.set_result(value)
    return f


def summarize(finished, loop):
    loop = dict(loop)
    failed = loop.pop("failed", [])
    lat = np.array([j.done - j.request.arrival for j in finished])
    wait = np.array([j.admitted - j.request.arrival for j in finished])
    points = sum(j.points * j.quadratures for j in finished)
    iters = np.array([j.quadratures for j in finished])
    return {"requests": len(finished), "converged": int(sum(bool(j.result.get("converged")) for j in finished)),
            "scf_iterations_mean": float(np.mean(iters)) if len(iters) else 0.0,
            "scf_iterations_max": int(np.max(iters)) if len(iters) else 0,
            "host_s_per_request_mean": float(np.mean([j.host_s for j in finished])) if finished else 0.0,
            "latency_s_p50": float(np.percentile(lat, 50)),
            "latency_s_p99": float(np.percentile(lat, 99)), "queue_wait_s_p50": float(np.percentile(wait, 50)),
            "requests_per_s": len(finished) / loop["wall_s"], "grid_points_per_s": points / loop["wall_s"],
            "quadratures": int(sum(j.quadratures for j in finished)), "failed": len(failed), **loop}
