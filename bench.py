#!/usr/bin/env python
"""XC quadrature throughput on one MI355X: grid points per second of the mixed-precision path and of the FP64 path.

One step = one quadrature (xq_integrate: density and gradient, GGA functional, E_xc, N_e and V_xc) over every grid point of
the workload, with the grid, the screened basis values and the density already on the device.  `value` is grid points per
second of the mixed-precision path from device time; the FP64 path, the mixed path's E_xc / V_xc error against it, the
host wall time per call (P upload and V download over PCIe included) and the per-call latency percentiles are reported beside
it.  The CPU baseline (rank 0 only) times the reference's own C quadrature (compiled from its sources into
oracle/_ref/libxcref.so), single-threaded, on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--workload h2o16|h2o64|h2o128] [--cpu-baseline ...]
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOADS = {
    # name: (waters, cube in Angstrom, seed, functional): the (H2O)16 PBE, (H2O)64 BLYP and (H2O)128 PBE clusters.
    "h2o16": (16, 9.0, 1, "pbe"),
    "h2o64": (64, 14.0, 2, "blyp"),
    "h2o128": (128, 18.0, 4, "pbe"),
}
METRIC = "XC grid pts/s mixed vs FP64 at |dExc|<=1e-6 Eh; p50/p99 SCF-request latency"
PEAK_HBM_GBS = 8000.0
PEAK_INT8_TOPS = 5000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--workload", default="h2o64", choices=sorted(WORKLOADS))
    p.add_argument("--spherical", action="store_true", help="spherical d (reference default); default Cartesian d")
    p.add_argument("--cpu-baseline", default="reference", choices=["reference", "port", "none"])
    p.add_argument("--cpu-sample", default="h2o16", choices=sorted(WORKLOADS),
                   help="workload the CPU baseline is timed on (default the (H2O)16 PBE cluster, the configuration whose "
                        "quadrature is compared with the reference's CPU path)")
    p.add_argument("--cpu-sample-waters", type=int, default=8,
                   help="waters of the extra, smaller CPU sample and of the gradient leg's CPU sample (0 skips the extra)")
    p.add_argument("--skip-fp64", action="store_true")
    p.add_argument("--uks-steps", type=int, default=3,
                   help="timed unrestricted quadratures (alpha 0.55 P, beta 0.45 P) on the workload, each precision "
                        "(0 skips them)")
    p.add_argument("--grad-steps", type=int, default=2,
                   help="timed XC nuclear-gradient calls on the workload after the quadrature legs (0 skips them)")
    p.add_argument("--scf", default="h2o16,h2o64",
                   help="SCF legs, comma-separated workloads: h2o16 runs one full SCF in each precision, timed; h2o64 runs "
                        "the FP64 SCF with the mixed-precision quadrature's E_xc / V_xc errors at every iteration; "
                        "'none' skips them")
    p.add_argument("--serve-requests", type=int, default=256,
                   help="requests in the serving stream (0 skips it)")
    p.add_argument("--serve-slots", type=int, default=16, help="jobs set up or running at once in the serving loop")
    p.add_argument("--serve-load", type=float, default=1.0,
                   help="arrival rate as a multiple of the serving loop's measured capacity")
    p.add_argument("--copy-calibration", action="store_true",
                   help="after the measurement, copy 1 GiB device to device once (a plain copy kernel of known traffic "
                        "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    return p.parse_args()


def copy_calibration():
    """One device-to-device copy of exactly 1 GiB (read 1 GiB, write 1 GiB) by the HIP runtime's copy kernel."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    size = 1 << 30
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(size)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(size)):
        raise RuntimeError("hipMalloc failed for the copy calibration")
    hip.hipMemset(a, 1, ctypes.c_size_t(size))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(b, a, ctypes.c_size_t(size), 3)    # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def measured_traffic(kernel, workload):
    """HBM bytes per launch of a kernel from its committed rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE runs, each
    scaled by a plain copy's reading) and the commit the counters were taken at, or (None, None)."""
    path = os.path.join(ROOT, "profiles", "{:s}_{:s}_pmc.txt".format(kernel, workload))
    if not os.path.exists(path):
        return None, None
    value, commit = None, None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                value = float(line.split(":", 1)[1])
            elif line.startswith("profiled_at_commit:"):
                commit = line.split(":", 1)[1].strip()
    return value, commit


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_leg(grid, ids, P, precision, steps, warmup, spec, spherical):
    t0 = time.perf_counter()
    grid.set_basis(spec, spherical, precision)
    basis_s = time.perf_counter() - t0
    V = np.zeros_like(P)
    for _ in range(warmup):
        V[:] = 0.0
        grid.integrate(ids, P, V)
    dev, rho, product, fock, wall = [], [], [], [], []
    E = N = None
    for _ in range(steps):
        V[:] = 0.0
        t = time.perf_counter()
        E, N, st = grid.integrate(ids, P, V)
        wall.append((time.perf_counter() - t) * 1e3)
        dev.append(st["device_ms"])
        rho.append(st["rho_ms"])
        product.append(st["product_ms"])
        fock.append(st["fock_ms"])
    return {"E": E, "N": N, "V": V.copy(), "device_ms": np.asarray(dev), "rho_ms": np.asarray(rho),
            "product_ms": np.asarray(product), "fock_ms": np.asarray(fock), "wall_ms": np.asarray(wall), "stats": st,
            "basis_s": basis_s, "device_bytes": grid.device_bytes()}


def cpu_baseline(kind, n_waters, cube, seed, functional, spherical, threads=1, seconds=10.0):
    """CPU quadrature of a (H2O)n cluster from the same generator (cube side in angstrom, seed); points/s.  The compiled
    reference is single-threaded; the port (the oracle's C restatement) runs its blocks on `threads` OpenMP threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from pdynamo3_amd import basis, densities, molecules
    from pdynamo3_amd.definitions import functional_ids
    if kind == "reference" and not oracle.reference_available():
        kind = "port"
    lib = oracle.reference() if kind == "reference" else oracle.oracle()
    Z, xa, rots, _ = molecules.water_cluster(n_waters, cube, seed)
    xyz = molecules.to_bohr(xa)
    spec = basis.element_spec(basis.load_basis(), Z)
    P = densities.cluster_density(rots, spherical=spherical)
    h = lib.setup(2, Z, xyz, spec, spherical, functional_ids(functional))
    # The reference's first call evaluates and caches the basis values (inCore); the in-core calls after it are timed,
    # as the GPU's are, repeated to about `seconds` of CPU work.
    threads = 1 if kind == "reference" else threads
    h.integrate(P, threads)
    _, _, _, sec = h.integrate(P, threads)
    reps = max(1, int(np.ceil((seconds - sec) / max(sec, 1e-3))))
    total = sec
    for _ in range(reps):
        total += h.integrate(P, threads)[3]
    reps += 1
    return {"value": h.npoints * reps / total, "unit": "grid points/s", "cores": threads, "kind": kind,
            "sample": "(H2O){:d} cluster, {:.0f} A cube, seed {:d}, {:s}, 6-31G* {:s}, Medium grid ({:d} points), "
                      "{:d} in-core quadrature calls after the first, {:.1f} s".format(
                          n_waters, cube, seed, functional.upper(), "spherical" if spherical else "Cartesian",
                          h.npoints, reps, total)}


def uks_leg(grid, ids, P, steps, spec, spherical):
    """Unrestricted quadratures (xq_integrate_uks) with alpha = 0.55 P and beta = 0.45 P, both precisions: device ms from
    the first upload to the last result (the two density stages, the spin-polarised functional and the two Fock stages,
    with the host waits between them), and the mixed path's E_xc and V_xc against the FP64 path's."""
    Pa, Pb = 0.55 * P, 0.45 * P
    out, res = {}, {}
    for precision in ("fp64", "mixed"):
        grid.set_basis(spec, spherical, precision)
        Va, Vb = np.zeros_like(P), np.zeros_like(P)
        grid.integrate_uks(ids, Pa, Pb, Va, Vb)
        dev = []
        for _ in range(steps):
            Va[:] = 0.0
            Vb[:] = 0.0
            E, N, st = grid.integrate_uks(ids, Pa, Pb, Va, Vb)
            dev.append(st["device_ms"])
        res[precision] = (E, N, Va.copy(), Vb.copy())
        out[precision + "_ms_per_call"] = float(np.mean(dev))
    (e0, n0, a0, b0), (e1, n1, a1, b1) = res["fp64"], res["mixed"]
    out.update({"E_xc": e0, "N_e": n0, "dExc_mixed_vs_fp64": e1 - e0,
                "max_dVxc_mixed_vs_fp64": float(max(np.max(np.abs(a1 - a0)), np.max(np.abs(b1 - b0)))),
                "calls": steps, "densities": "alpha 0.55 P, beta 0.45 P of the workload's density"})
    return out


def gradient_leg(grid, ids, P, steps, n_waters, functional, spherical, cpu_kind):
    """XC nuclear gradients (xq_gradients, FP64) on the workload: wall time per call (P upload and the gradient download
    included), the translational sum of the gradient, the same for the unrestricted gradients (xq_gradients_uks), and the
    CPU's time for the restricted call on the bounded sample."""
    G = grid.gradients(ids, P)
    wall = []
    for _ in range(steps):
        t = time.perf_counter()
        G = grid.gradients(ids, P)
        wall.append((time.perf_counter() - t) * 1e3)
    out = {"ms_per_call": float(np.mean(wall)), "calls": steps, "atoms": int(G.shape[0]),
           "max_abs_gradient": float(np.max(np.abs(G))), "max_abs_translational_sum": float(np.max(np.abs(G.sum(0))))}
    # Unrestricted (xq_gradients_uks) at alpha = 0.55 P, beta = 0.45 P.
    Gu = grid.gradients_uks(ids, 0.55 * P, 0.45 * P)
    wall = []
    for _ in range(steps):
        t = time.perf_counter()
        Gu = grid.gradients_uks(ids, 0.55 * P, 0.45 * P)
        wall.append((time.perf_counter() - t) * 1e3)
    out["unrestricted"] = {"ms_per_call": float(np.mean(wall)), "calls": steps,
                           "densities": "alpha = 0.55 P, beta = 0.45 P",
                           "max_abs_translational_sum": float(np.max(np.abs(Gu.sum(0))))}
    if cpu_kind != "none":
        try:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle
            from pdynamo3_amd import basis, densities, molecules
            from pdynamo3_amd.definitions import functional_ids
            kind = "reference" if cpu_kind == "reference" and oracle.reference_available() else "port"
            lib = oracle.reference() if kind == "reference" else oracle.oracle()
            Z, xa, rots, _ = molecules.water_cluster(n_waters, 9.0 * (n_waters / 16.0) ** (1.0 / 3.0), 1)
            h = lib.setup(2, Z, molecules.to_bohr(xa), basis.element_spec(basis.load_basis(), Z), spherical,
                          functional_ids(functional))
            Ps = densities.cluster_density(rots, spherical=spherical)
            t = time.perf_counter()
            h.gradients(Ps, 1)
            out["cpu_baseline"] = {"seconds_per_call": time.perf_counter() - t, "cores": 1, "kind": kind,
                                   "sample": "(H2O){:d} cluster, same generator, {:s}, Medium grid ({:d} points), one "
                                             "gradient call".format(n_waters, functional.upper(), h.npoints)}
        except Exception as exc:
            out["cpu_baseline"] = {"value": None, "error": repr(exc)}
    return out


def scf_leg(name, compare, log_fn):
    """Restricted Kohn-Sham SCF of a workload cluster (scf.RKSDFT: one-electron and electron-fit integrals on the GPU, the
    XC quadrature on the GPU, ODA then DIIS on the host, converged to 1e-6 Eh and an RMS density change of 1e-6), from
    the diagonal guess.  compare=False: one full SCF per precision, wall time of the set-up and of the iterations with
    the time per phase.  compare=True: the FP64 SCF, and at every iteration the mixed-precision quadrature at the same
    density on a second handle, with its E_xc, N_e and V_xc differences from the FP64 quadrature."""
    from pdynamo3_amd import basis, integrator, molecules, scf
    n_w, cube, seed, functional = WORKLOADS[name]
    Z, xa, _, _ = molecules.water_cluster(n_w, cube, seed)
    xyz = molecules.to_bohr(xa)
    spec = basis.element_spec(basis.load_basis(), Z)
    out = {"workload": "(H2O){:d} cluster, {:.0f} A cube, seed {:d}, 6-31G* Cartesian, {:s}, Medium grid, J fitted in "
                       "DGauss A1 (dgauss-a1-dftjfit)".format(n_w, cube, seed, functional.upper())}
    for precision in (["fp64"] if compare else ["fp64", "mixed"]):
        t0 = time.perf_counter()
        s = scf.RKSDFT(Z, xyz, functional=functional, precision=precision, spherical=False)
        setup_s = time.perf_counter() - t0
        gm, rows, t_cmp = None, [], 0.0
        if compare:
            gm = integrator.QuadratureGrid(Z, xyz, 2)
            gm.set_basis(spec, False, "mixed")
        t1 = time.perf_counter()
        while True:
            P = s.density()
            e, V = s.xc(P)
            if gm is not None:
                tc = time.perf_counter()
                Vm = np.zeros_like(P)
                em, nm, _ = gm.integrate(s.ids, P, Vm)
                _, n64, _ = s.grid.integrate(s.ids, P, np.zeros_like(P))
                t_cmp += time.perf_counter() - tc
                rows.append({"iteration": s.iteration + 1, "dExc": em - e, "dNe": nm - n64,
                             "max_dVxc": float(np.max(np.abs(Vm - V)))})
            if s.advance(e, V):
                break
        iter_s = time.perf_counter() - t1 - t_cmp
        res = s.summary()
        leg = {"energy": res["energy"], "converged": res["converged"], "iterations": res["iterations"],
               "setup_s": setup_s, "iterations_s": iter_s, "total_s": setup_s + iter_s,
               "timings": res["timings"], "points": s.grid.number_of_points, "functions": s.n,
               "fit_functions": s.integrals.n_fit, **s.integrals.stats()}
        if rows:
            leg["mixed_vs_fp64_per_iteration"] = rows
            leg["max_abs_dExc"] = max(abs(r["dExc"]) for r in rows)
            leg["max_dVxc"] = max(r["max_dVxc"] for r in rows)
            leg["note"] = "iterations_s excludes the comparison quadratures"
            gm.close()
        s.close()
        out[precision] = leg
        log_fn("scf {:s} {:s}:".format(name, precision), json.dumps({k: v for k, v in leg.items()
                                                                      if k != "mixed_vs_fp64_per_iteration"}))
    if not compare:
        out["dE_mixed_vs_fp64"] = out["mixed"]["energy"] - out["fp64"]["energy"]
    return out


def progress(tag, log_fn, every=32):
    """A log function for the request loop that reports every `every`-th finished request (and the 8th, for short runs)."""
    count = [0]

    def log(message):
        if "latency" in message:
            count[0] += 1
            if count[0] % every == 0 or count[0] == 8:
                log_fn("serving {:s}: {:d} requests done".format(tag, count[0]))
    return log


def serving_leg(args, log_fn):
    """The serving stream in both precisions: SCF requests on (H2O)N, N log-uniform in [4, 64] (seed 3), BLYP/PBE by coin
    flip, Poisson arrivals, each a restricted Kohn-Sham SCF to convergence.  The mixed loop's capacity is measured
    first, on a calibration stream of 24 other requests (seed 7) that all arrive at once, after a warm-up pass of 8
    more (seed 8) that fills the device memory cache and loads the kernels; the arrival rate is serve-load x that
    capacity.  The FP64 loop then serves the same stream (same
    arrival times) after its own warm-up and calibration, so the two legs differ only in the precision."""
    import dataclasses
    from pdynamo3_amd import serving
    base = serving.request_stream(n_requests=args.serve_requests, seed=3, rate=1.0)
    out, rate = {}, None
    for precision in ("mixed", "fp64"):
        srv = serving.Server(precision=precision, max_jobs=args.serve_slots, scf="rks")
        srv.run(serving.request_stream(n_requests=8, seed=8, rate=1.0e6), log=progress(precision + " warm-up", log_fn))
        done, cl = srv.run(serving.request_stream(n_requests=24, seed=7, rate=1.0e6),
                           log=progress(precision + " calibration", log_fn))
        capacity = len(done) / cl["wall_s"]
        if rate is None:
            rate = args.serve_load * capacity
        reqs = [dataclasses.replace(r, arrival=r.arrival / rate) for r in base]
        finished, loop = srv.run(reqs, log=progress(precision + " stream", log_fn))
        srv.close()
        leg = serving.summarize(finished, loop)
        leg.update({"arrival_rate_per_s": rate, "capacity_requests_per_s": capacity, "load": rate / capacity,
                    "slots": args.serve_slots, "hbm_budget_bytes": srv.hbm_budget, "precision": precision,
                    "stream": "{:d} requests, (H2O)N with N log-uniform in [4, 64] (seed 3), 6-31G* Cartesian, BLYP/PBE "
                              "by coin flip, Poisson arrivals at the mixed loop's capacity; each request runs a "
                              "restricted Kohn-Sham SCF (J fitted, XC quadrature on the GPU) from the diagonal guess to "
                              "1e-6 Eh and an RMS density change of 1e-6".format(len(reqs))})
        out[precision] = leg
        log_fn("serving {:s}:".format(precision), json.dumps(leg))
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has a single MI355X, and the quadrature runs on one "
              "GPU; across GPUs it would shard by grid block. Run with --gpus 1.", file=sys.stderr)
        return 2
    from pdynamo3_amd import basis, densities, integrator, molecules
    from pdynamo3_amd.definitions import functional_ids
    n_w, cube, seed, functional = WORKLOADS[args.workload]
    spherical = bool(args.spherical)
    Z, xa, rots, _ = molecules.water_cluster(n_w, cube, seed)
    xyz = molecules.to_bohr(xa)
    b = basis.load_basis("6-31g_st")
    spec = basis.element_spec(b, Z)
    n = basis.number_of_functions(b, Z, spherical)
    P = densities.cluster_density(rots, spherical=spherical)
    ids = functional_ids(functional)
    log("workload", args.workload, "atoms", len(Z), "functions", n)
    t0 = time.perf_counter()
    grid = integrator.QuadratureGrid(Z, xyz, 2)
    grid_s = time.perf_counter() - t0
    G = grid.number_of_points
    log("grid", G, "points", grid.number_of_blocks, "blocks", "{:.3f} s".format(grid_s))
    fp64 = None
    if not args.skip_fp64:
        fp64 = run_leg(grid, ids, P, "fp64", args.steps, args.warmup, spec, spherical)
        log("fp64: E", fp64["E"], "N", fp64["N"], "device ms", np.mean(fp64["device_ms"]), "rho", np.mean(fp64["rho_ms"]),
            "fock", np.mean(fp64["fock_ms"]))
    mixed = run_leg(grid, ids, P, "mixed", args.steps, args.warmup, spec, spherical)
    log("mixed: E", mixed["E"], "N", mixed["N"], "device ms", np.mean(mixed["device_ms"]), "rho",
        np.mean(mixed["rho_ms"]), "fock", np.mean(mixed["fock_ms"]))
    st = mixed["stats"]
    ms = float(np.mean(mixed["device_ms"]))
    value = G / (ms * 1e-3)
    # Roofline of the dominant kernel: density_rows_kernel, X = P_red Phi per row chunk with the chunk's share of rho and
    # grad rho, timed alone by the library's events on its stream (stats[9]; the density stage adds the per-block
    # functional kernel, the Fock stage is two kernels).  Algorithmic bytes (stats[10]): every value of a kept function
    # at a point read once (17 bytes: 5 Phi digits and three 32-bit fixed-point gradients), P digits and row exponents,
    # the point exponents, and each row chunk's four shares per point written once.  Its int8 matrix work at peak is
    # far shorter than its bytes at peak, so the HBM roof is the one that binds.
    rho_ms = float(np.mean(mixed["product_ms"]))
    rho_bytes = st["product_bytes"]
    flops = 2.0 * st["bf_squares"]                    # X = P_red Phi: 2 nb^2 G per block, as algorithmic flops
    int8_ops = 2.0 * 15.0 * st["bf_squares_padded"]   # what the matrix cores issue: 15 digit products, rows padded to 32
    t_hbm, t_mfma = rho_bytes / (PEAK_HBM_GBS * 1e9), int8_ops / (PEAK_INT8_TOPS * 1e12)
    if t_hbm >= t_mfma:
        achieved, peak, unit, bound = rho_bytes / (rho_ms * 1e-3) / 1e9, PEAK_HBM_GBS, "GB/s", "hbm"
    else:
        achieved, peak, unit, bound = int8_ops / (rho_ms * 1e-3) / 1e12, PEAK_INT8_TOPS, "TOP/s", "mfma"
    traffic, traffic_commit = measured_traffic("density_rows_kernel", args.workload)
    roofline = {"kernel": "density_rows_kernel", "bound": bound, "achieved": achieved, "peak": peak, "unit": unit,
                "frac": achieved / peak, "traffic": traffic,
                "traffic_source": None if traffic is None else
                "profiles/density_rows_kernel_{:s}_pmc.txt (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, at commit "
                "{:s})".format(args.workload, traffic_commit or "unrecorded"),
                "algorithmic_bytes": rho_bytes, "algorithmic_flops": flops,
                "algorithmic_tflops_per_s": flops / (rho_ms * 1e-3) / 1e12, "hardware_int8_ops": int8_ops,
                "kernel_ms": rho_ms, "density_stage_ms": float(np.mean(mixed["rho_ms"])),
                "density_stage_bytes": st["rho_bytes"], "fock_stage_ms": float(np.mean(mixed["fock_ms"])),
                "fock_stage_bytes": st["fock_bytes"], "fock_algorithmic_flops": 4.0 * st["bf_squares"]}
    out = {"metric": METRIC, "value": value, "unit": "grid points/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "int8 digits x int8 -> int32 (exact) with f64 functional and sums", "data": "synthetic",
           "config": {"workload": "(H2O){:d} cluster, {:.0f} A cube, seed {:d}, 6-31G* {:s} ({:d} functions), {:s}, "
                                  "Medium grid ({:d} points)".format(n_w, cube, seed,
                                                                     "spherical" if spherical else "Cartesian", n,
                                                                     functional.upper(), G),
                      "parallelism": "single GPU"},
           "grid_points": G, "grid_build_s": grid_s, "basis_eval_s": mixed["basis_s"],
           "host_wall_ms_per_call_pcie_inclusive": float(np.mean(mixed["wall_ms"])),
           "call_ms_p50": float(np.percentile(mixed["wall_ms"], 50)),
           "call_ms_p99": float(np.percentile(mixed["wall_ms"], 99)),
           "E_xc": mixed["E"], "N_e": mixed["N"], "device_bytes": mixed["device_bytes"], "roofline": roofline}
    if fp64 is not None:
        ms64 = float(np.mean(fp64["device_ms"]))
        out["fp64_value"] = G / (ms64 * 1e-3)
        out["fp64_ms_per_step"] = ms64
        out["speedup_mixed_vs_fp64"] = ms64 / ms
        out["dExc_mixed_vs_fp64"] = mixed["E"] - fp64["E"]
        out["max_dVxc_mixed_vs_fp64"] = float(np.max(np.abs(mixed["V"] - fp64["V"])))
    if args.uks_steps > 0:
        out["uks"] = uks_leg(grid, ids, P, args.uks_steps, spec, spherical)
        log("uks:", json.dumps(out["uks"]))
        grid.set_basis(spec, spherical, "mixed")
    if args.grad_steps > 0:
        out["xc_gradients"] = gradient_leg(grid, ids, P, args.grad_steps, args.cpu_sample_waters, functional, spherical,
                                           args.cpu_baseline)
        log("gradients:", json.dumps(out["xc_gradients"]))
    grid.close()
    if args.scf != "none":
        out["scf"] = {}
        for name in [w for w in args.scf.split(",") if w]:
            out["scf"][name] = scf_leg(name, name != "h2o16", log)
    if args.serve_requests > 0:
        serve = serving_leg(args, log)
        out["latency_ms_p50"] = 1e3 * serve["mixed"]["latency_s_p50"]
        out["latency_ms_p99"] = 1e3 * serve["mixed"]["latency_s_p99"]
        out["serving"] = serve["mixed"]
        out["serving_fp64"] = serve["fp64"]
    if args.cpu_baseline != "none":
        cn, ccube, cseed, cfun = WORKLOADS[args.cpu_sample]
        try:
            out["cpu_baseline"] = cpu_baseline(args.cpu_baseline, cn, ccube, cseed, cfun, spherical)
        except Exception as exc:  # the bench line is still printed; the failure is named
            out["cpu_baseline"] = {"value": None, "error": repr(exc)}
        try:
            # The restatement on the host cores this job may use (16 per GPU on the box).
            out["cpu_baseline_all_cores"] = cpu_baseline("port", cn, ccube, cseed, cfun, spherical,
                                                         threads=min(16, os.cpu_count() or 1))
        except Exception as exc:
            out["cpu_baseline_all_cores"] = {"value": None, "error": repr(exc)}
        if args.cpu_sample_waters > 0:
            try:
                # A smaller sample too: per-point cost grows with the kept functions per block, so the rate differs.
                w = args.cpu_sample_waters
                out["cpu_baseline_small_sample"] = cpu_baseline(args.cpu_baseline, w, 9.0 * (w / 16.0) ** (1.0 / 3.0), 1,
                                                                functional, spherical, seconds=5.0)
            except Exception as exc:
                out["cpu_baseline_small_sample"] = {"value": None, "error": repr(exc)}
    print(json.dumps(out), flush=True)
    if args.copy_calibration:
        copy_calibration()
    return 0


if __name__ == "__main__":
    sys.exit(main())
