"""Benchmark: solves/s and p50/p99 latency of the linear-static voxel FEA solve on one MI355X,
mixed precision against plain FP64 MG-PCG, at <= 1e-6 relative displacement error.

Workload (one GPU): the 512x128x128 void block (25 610 499 DOF): seeded spherical voids (radii
U[2, 8] voxels, 20 % void fraction, void stiffness 1e-3), E = 70 GPa, nu = 0.33, x = 0 clamped,
-z pressure of 5 MPa on the top face over x > 200 mm, 6-level geometric multigrid, Chebyshev degree-3
smoothing (the workload as configured). Mixed precision: FP32 outer operator, FP32 V-cycle whose
level-0 and level-1 (child-scale) smoothing takes the element products on the matrix cores (BF16
pieces, FP32 accumulation),
FP64 coarsest solve, FP32 CG recurrences with FP64 dot products, the FP64 solution and the true
residual recomputed in FP64 at least every 20 iterations. The plain FP64 MG-PCG runs the same
multigrid in FP64. Both solves run to the reference solver's default tolerance, relative residual
1e-6.

A step is one complete solve from zero with the vectors resident in HBM (the hierarchy is built
once, as a served part's is). Every timed solve is checked (<= 1e-6) against a plain FP64 GPU solve
to 1e-11; the reported errors are against the FP64 oracle, the reference's own VoxelFEA::solve
(oracle/_ref) continued to a relative residual of 1e-10.
The CPU baseline times that reference solver, compiled from its sources, on the host share of the
GPU: build, smoother estimates, operator apply, V-cycle and one full solve of the same workload.
The default line also carries degree-2 smoothing with the FP32 VALU smoother (both precisions), the
batched load-case leg and a bounded serving leg as extra keys.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--dims 512,128,128] [--no-cpu] [--no-extras]
    python bench.py --batched [--nrhs 16]      (16 load cases of the 128x64x64 SIMP block, one solve)
    python bench.py --serve N [--rate 200]     (the serving stream through the request loop)
    python bench.py --build-only [--steps K]   (hierarchy build on the device against the host build)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))


def _package():
    import importlib.util
    if "parts_sim_amd" in sys.modules:
        return sys.modules["parts_sim_amd"]
    pkg = os.path.join(ROOT, "parts-sim_amd")
    spec = importlib.util.spec_from_file_location("parts_sim_amd", os.path.join(pkg, "__init__.py"), submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["parts_sim_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


METRIC = "solves/s + p50/p99 latency at 1e-6 rel. error; mixed vs plain FP64 MG-PCG"
HBM_PEAK = 8.0e12
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fine_operator_fp32_void_block_512x128x128.txt")
SMOOTH_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "smoothing_step_mfma_void_block_512x128x128.txt")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def measured_traffic(dims, path=TRAFFIC_FILE):
    """HBM bytes per launch of a kernel from its committed counter summary (rocprofv3 FETCH_SIZE and
    WRITE_SIZE runs scaled by a copy kernel's readings), or None."""
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch") and ("dims %d,%d,%d" % dims) in line:
                    return float(line.split(":")[1].split()[0])
    except OSError:
        pass
    return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, coarsest, tol, reps=3):
    """The reference's own CPU solver (oracle/_ref, compiled from its sources) on the host share of
    the GPU: hierarchy build, smoother estimates, best-of-`reps` operator apply and V-cycle, and one
    full solve of the workload from zero to `tol`, each timed (bench.cpp:29-37, 63-87). The solve is
    then continued from its solution to 1e-10, which is the FP64 oracle solution the GPU results are
    checked against. Returns (cpu_baseline dict, oracle u) or (None, None) when oracle/_ref is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    # the reference's thread pool takes PARTS_SIM_THREADS (parallel.cpp:7-16); 16 = the host cores
    # one GPU of the box is given
    os.environ.setdefault("PARTS_SIM_THREADS", "16")
    import cpu
    if not cpu.available("ref"):
        return None, None
    t = time.perf_counter()
    m = cpu.Model(w["dims"], w["density"], w["nu"], w["bc"], coarsest, kind="ref")
    build_s = time.perf_counter() - t
    t = time.perf_counter()
    m.prepare_smoothers()
    smoothers_s = time.perf_counter() - t
    x = np.random.default_rng(0).standard_normal(m.level(0)["nDof"])

    def best(fn):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return min(ts)

    apply_s = best(lambda: m.apply(x))
    vcycle_s = best(lambda: m.precondition(x))
    t = time.perf_counter()
    r = m.solve(w["f"], tol=tol, max_iter=400)
    solve_s = time.perf_counter() - t
    log("CPU reference: build %.2f s, smoothers %.2f s, apply %.3f s, V-cycle %.3f s, solve %.2f s (%d its)"
        % (build_s, smoothers_s, apply_s, vcycle_s, solve_s, r["iterations"]))
    t = time.perf_counter()
    ro = m.solve(w["f"], tol=1e-10, max_iter=400, x0=r["u"])
    log("CPU reference continued to 1e-10: %d more its, %.2f s, residual %.2e" % (ro["iterations"], time.perf_counter() - t, ro["residual"]))
    if not ro["converged"]:
        raise RuntimeError("the reference's FP64 oracle solve did not reach 1e-10")
    n_dof = 3 * W_nodes(w["dims"])
    out = {"value": round(1.0 / solve_s, 5), "unit": "solves/s", "cores": int(cpu.ref_threads()), "kind": "reference", "cpu_model": cpu_model(),
           "sample": "one full solve of the same workload from zero to %g (%d PCG iterations, %.2f s) with the reference's own VoxelFEA::solve; "
                     "hierarchy build %.2f s and smoother estimates %.2f s excluded as on the GPU; %d threads, the host cores one GPU of the "
                     "box is given" % (tol, r["iterations"], solve_s, build_s, smoothers_s, int(cpu.ref_threads())),
           "iterations": int(r["iterations"]), "solve_s": round(solve_s, 3), "build_s": round(build_s, 3), "smoothers_s": round(smoothers_s, 3),
           "apply_s": round(apply_s, 4), "vcycle_s": round(vcycle_s, 4), "dof_iterations_per_s": float(n_dof * r["iterations"] / solve_s),
           "oracle_solve": {"tol": 1e-10, "extra_iterations": int(ro["iterations"]), "residual": ro["residual"]}}
    return out, ro["u"]


def W_nodes(dims):
    return int(np.prod([int(d) + 1 for d in dims]))


def serving_leg(P, W, n_requests, rate, inflight, tol, pack_max=0, pack_max_dof=0, edge_max=256, degree=3, mfma=1):
    """Serving workload: a seeded Poisson stream of void-block parts (edges log-uniform 32-256 voxels,
    aspect 1-4, nu U[0.28, 0.36], 1-8 load cases, 30 % warm re-solves of an earlier part with 1 % of
    its solid voxels removed and floating voxels pruned, warm-started from the earlier solution)
    through the device-resident request loop; latency is arrival -> result. Parts of at most
    pack_max_dof DOFs (0: the server's default) run in its shared-launch pack of up to pack_max
    requests (1: no packing). Returns the summary."""
    import threading
    reqs = W.serving_stream(n_requests, seed=5, rate=rate, edge_max=edge_max)
    srv = P.Server(max_inflight=inflight, pack_max=pack_max, pack_max_dof=pack_max_dof)
    parts, results, lock = {}, {}, threading.Lock()
    done_evt = {r["id"]: threading.Event() for r in reqs}

    loads = W.serving_loads

    def waiter(r, ticket):
        u, res = srv.wait(ticket)
        t = time.perf_counter()
        with lock:
            results[r["id"]] = (r, res, t, u)
        done_evt[r["id"]].set()

    def client(r):
        # one thread per request: a warm re-solve waits for its parent's result without holding up
        # later arrivals
        if r["warm"]:
            done_evt[r["parent"]].wait()
            base = parts[r["parent"]]
            dens = W.remove_voxels(base["density"], r["remove_fraction"], r["remove_seed"])
            dens, _ = P.prune_floating(base["dims"], dens, W.held_nodes(base["bc"]))
            x0 = results[r["parent"]][3]
            F = loads(base, r["nrhs"], r["remove_seed"])
            tk = srv.submit(base["dims"], dens, base["nu"], base["bc"], F, x0=x0, coarsest_max_dof=1200, tol=tol, precision="mixed",
                            cheb_degree=degree, mfma_smoother=mfma)
        else:
            w = parts[r["id"]]
            tk = srv.submit(w["dims"], w["density"], w["nu"], w["bc"], loads(w, r["nrhs"], r["part_seed"]), coarsest_max_dof=1200, tol=tol,
                            precision="mixed", cheb_degree=degree, mfma_smoother=mfma)
        waiter(r, tk)
        with lock:
            n = len(results)
        if n % 10 == 0:
            log("serving: %d of %d requests done" % (n, len(reqs)))

    for r in reqs:
        if not r["warm"]:
            parts[r["id"]] = W.void_block(r["dims"], seed=r["part_seed"], nu=r["nu"], E=r["E"])
    threads = []
    t0 = time.perf_counter()
    for r in reqs:
        delay = t0 + r["t"] - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        r["arrival"] = t0 + r["t"]
        th = threading.Thread(target=client, args=(r,))
        th.start()
        threads.append(th)
    for th in threads:
        th.join()
    srv.close()
    lat = np.array([(t - r["arrival"]) * 1e3 for r, res, t, u in results.values()])
    span = max(t for r, res, t, u in results.values()) - t0
    cases = sum(r["nrhs"] for r in reqs)
    bad = [res for r, res, t, u in results.values() if not res["converged"]]
    out = {"metric": "serving: solves/s + p50/p99 latency (arrival to result), mixed precision", "requests": len(reqs), "load_cases": cases,
           "solves_per_s": round(cases / span, 2), "requests_per_s": round(len(reqs) / span, 2), "p50_ms": round(float(np.percentile(lat, 50)), 1),
           "p99_ms": round(float(np.percentile(lat, 99)), 1), "max_inflight": inflight, "arrival_rate": rate, "unconverged": len(bad),
           "mean_build_ms": round(float(np.mean([res["build_ms"] for r, res, t, u in results.values()])), 1),
           "mean_solve_ms": round(float(np.mean([res["solve_ms"] for r, res, t, u in results.values()])), 1),
           "mean_queue_ms": round(float(np.mean([res["queue_ms"] for r, res, t, u in results.values()])), 1),
           "packed_requests": int(sum(res["packed"] for r, res, t, u in results.values())), "pack_max": pack_max, "pack_max_dof": pack_max_dof,
           "dof_total": int(sum(3 * W.nodes(r["dims"]) * r["nrhs"] for r in reqs)), "tol": tol, "edge_max": edge_max}
    return out


def serve(args, P, W):
    print(json.dumps(serving_leg(P, W, args.serve, args.rate, args.inflight, args.tol, args.pack_max, args.pack_max_dof, args.edge_max, args.degree,
                                 1 if args.smoother == "mfma" else 0)), flush=True)


def packed_leg(P, W, n_parts, tol, reps=3, eager=False):
    """Packed solves in isolation: the first n_parts cold parts of the serving stream (seed 5) of at
    most 3 * 2^20 DOFs, each with its 1-8 load cases, solved one model after another
    (psx_solve_multi_device, each model's own graph) and all together (psx_solve_packed, shared
    launches). Hierarchies are built and the smoothers estimated before the timed region; the
    packed results must equal the lone ones bit for bit."""
    reqs = [r for r in W.serving_stream(400, seed=5) if not r["warm"] and 3 * W.nodes(r["dims"]) <= 3 << 20][:n_parts]
    parts = []
    for r in reqs:
        w = W.void_block(r["dims"], seed=r["part_seed"], nu=r["nu"], E=r["E"])
        parts.append((w, W.serving_loads(w, r["nrhs"], r["part_seed"])))
    models = [P.Model(w["dims"], w["density"], w["nu"], w["bc"], 1200, device_build=True) for w, _ in parts]
    for m in models:
        m.estimate_smoothers()
    d_f = [P.DeviceArray(F.size, F.reshape(-1)) for _, F in parts]
    d_u = [P.DeviceArray(F.size) for _, F in parts]
    nr = [F.shape[0] for _, F in parts]
    kw = dict(tol=tol, max_iter=2000, precision="mixed", eager=eager)
    lone_t, pack_t, same = [], [], True
    for rep in range(reps + 1):
        t = time.perf_counter()
        for m, f, u, k in zip(models, d_f, d_u, nr):
            m.solve_multi_device(k, f, u, **kw)
        lone_t.append(time.perf_counter() - t)
        U = [u.download() for u in d_u]
        t = time.perf_counter()
        infos = P.solve_packed(models, d_f, d_u, nr, **kw)
        pack_t.append(time.perf_counter() - t)
        same = same and all(np.array_equal(a, u.download()) for a, u in zip(U, d_u))
        if not all(i["converged"] for inf in infos for i in inf):
            raise RuntimeError("a packed solve did not converge")
    for m in models:
        m.close()
    lone, pack = float(np.mean(lone_t[1:])), float(np.mean(pack_t[1:]))
    cases = sum(nr)
    return {"metric": "packed solves: load cases/s of small serving parts in shared launches vs one model after another",
            "parts": len(parts), "load_cases": cases, "dof_max": int(max(3 * W.nodes(w["dims"]) for w, _ in parts)),
            "lone_sequential_ms": round(1e3 * lone, 2), "packed_ms": round(1e3 * pack, 2), "speedup": round(lone / pack, 3),
            "solves_per_s_lone": round(cases / lone, 1), "solves_per_s_packed": round(cases / pack, 1),
            "bitwise_equal": bool(same), "eager": eager, "tol": tol}


def batched(args, P, W):
    dims = tuple(int(v) for v in args.dims.split(",")) if args.dims != "512,128,128" else (128, 64, 64)
    if args.profile_solves:
        # profiling runs: --steps multi-instance solves of one precision only
        w = W.simp_block(dims, seed=4, nrhs=args.nrhs)
        m = P.Model(w["dims"], w["density"], w["nu"], w["bc"], 1200, device_build=True)
        m.estimate_smoothers()
        d_f, d_u = P.DeviceArray(w["f"].size, w["f"].reshape(-1)), P.DeviceArray(w["f"].size)
        for k in range(args.steps):
            t = time.perf_counter()
            infos = m.solve_multi_device(args.nrhs, d_f, d_u, tol=args.tol, max_iter=3000, precision=args.profile_solves, check_every=8)
            log("%s batched solve %d: %d iterations, %.2f ms" % (args.profile_solves, k, max(i["iterations"] for i in infos), 1e3 * (time.perf_counter() - t)))
        return
    print(json.dumps(batched_leg(P, W, dims, args.nrhs, args.steps, args.warmup, args.tol, args.degree, 1 if args.smoother == "mfma" else 0)),
          flush=True)


def batched_leg(P, W, dims, nrhs, steps, warmup, tol, degree=3, mf=1):
    """Batched load cases: the 128x64x64 SIMP block (x ~ U[0.001, 1] per voxel, stiffness x^3,
    E0 = 200 GPa, nu = 0.3, x = 0 clamped) with 16 load cases (1 kN point loads at random nodes of
    the x = 128 face, directions uniform on the sphere), all 16 in one multi-instance solve (one
    launch per V-cycle level and CG step), mixed precision, tol 1e-6; against the same 16 solved one
    after another in mixed precision and in plain FP64. solves/s counts load cases. Returns the summary."""
    w = W.simp_block(dims, seed=4, nrhs=nrhs)
    F = w["f"]
    k, n = F.shape
    m = P.Model(w["dims"], w["density"], w["nu"], w["bc"], 1200, device_build=True)
    m.estimate_smoothers()
    log("workload", w["name"], "DOF per case", n, "levels", m.num_levels)
    d_f = P.DeviceArray(F.size, F.reshape(-1))
    d_u = P.DeviceArray(F.size)
    m.solve_multi_device(k, d_f, d_u, tol=1e-11, max_iter=3000, precision="fp64")
    U_ref = d_u.download().reshape(k, n)

    def errs(U):
        return [float(np.linalg.norm(U[c] - U_ref[c]) / np.linalg.norm(U_ref[c])) for c in range(k)]

    def timed(fn, steps, warmup):
        lat = []
        for s in range(warmup + steps):
            t = time.perf_counter()
            out = fn()
            if s >= warmup:
                lat.append(time.perf_counter() - t)
        return np.array(lat), out

    def multi(precision, mfma=0, deg=degree):
        return m.solve_multi_device(k, d_f, d_u, tol=tol, max_iter=3000, precision=precision, check_every=8, mfma_smoother=mfma,
                                    cheb_degree=deg)

    # the configuration of the line: Chebyshev degree 3, the smoothing of level 0 (per-voxel densities)
    # and level 1 (all own elements, child-scale form) as element products on the matrix cores
    lat_b, infos = timed(lambda: multi("mixed", mf), steps, warmup)
    e_b = errs(d_u.download().reshape(k, n))
    if not all(i["converged"] for i in infos) or max(e_b) > 1e-6:
        raise RuntimeError("batched mixed solve missed the accuracy bar")
    # also the coarse levels' own element products on the matrix cores (BF16 MFMA over the 16 cases)
    lat_mf, infos_mf = timed(lambda: multi("mixed", mf | 2), steps, warmup)
    e_mf = errs(d_u.download().reshape(k, n))
    if not all(i["converged"] for i in infos_mf) or max(e_mf) > 1e-6:
        raise RuntimeError("batched mixed solve with the matrix-core own elements missed the accuracy bar")
    # degree 2 with the FP32 VALU smoother (the earlier configuration)
    lat_el, infos_el = timed(lambda: multi("mixed", 0, 2), steps, warmup)
    e_el = errs(d_u.download().reshape(k, n))
    if not all(i["converged"] for i in infos_el) or max(e_el) > 1e-6:
        raise RuntimeError("batched mixed solve with the VALU smoother missed the accuracy bar")
    lat_b64, infos64 = timed(lambda: multi("fp64"), max(2, steps // 2), 1)
    e_b64 = errs(d_u.download().reshape(k, n))
    # one load case after another (the reference's way), same device vectors
    d_f1, d_u1 = P.DeviceArray(n), P.DeviceArray(n)

    def sequential(precision):
        its = []
        for c in range(k):
            d_f1.upload(F[c])
            its.append(m.solve_device(d_f1, d_u1, tol=tol, max_iter=3000, precision=precision, check_every=8, cheb_degree=degree,
                                      mfma_smoother=mf if precision == "mixed" else 0)["iterations"])
        return its

    lat_s, its_s = timed(lambda: sequential("mixed"), max(2, steps // 2), 1)
    lat_s64, its_s64 = timed(lambda: sequential("fp64"), max(2, steps // 2), 1)
    log("batched mixed %.1f ms (%d cases), batched FP64 %.1f ms, sequential mixed %.1f ms, sequential FP64 %.1f ms"
        % (1e3 * lat_b.mean(), k, 1e3 * lat_b64.mean(), 1e3 * lat_s.mean(), 1e3 * lat_s64.mean()))
    its = [i["iterations"] for i in infos]
    out = {"metric": "batched load cases: solves/s at 1e-6 rel. error, one multi-instance launch per MG level", "value": round(k / lat_b.mean(), 3),
           "unit": "solves/s", "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": round(1e3 * lat_b.mean(), 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
           "config": {"workload": w["name"] + ", SIMP x^3 densities U[0.001,1], E0 200 GPa, nu 0.3, x=0 clamped, 1 kN point loads on x=%d, "
                                   "Chebyshev degree %d, level-0 and level-1 smoothing %s" % (dims[0], degree, "on the matrix cores" if mf else "FP32 VALU"),
                      "load_cases": k, "dof_per_case": n, "levels": m.num_levels, "tol": tol, "parallelism": "single GPU"},
           "p50_ms": round(1e3 * float(np.percentile(lat_b, 50)), 3), "p99_ms": round(1e3 * float(np.percentile(lat_b, 99)), 3),
           "iterations_max": int(max(its)), "iterations_mean": float(np.mean(its)), "rel_error_vs_fp64_max": max(e_b),
           "batched_fp64": {"solves_per_s": round(k / lat_b64.mean(), 3), "ms": round(1e3 * lat_b64.mean(), 3), "rel_error_max": max(e_b64),
                            "iterations_max": int(max(i["iterations"] for i in infos64))},
           "sequential_mixed": {"solves_per_s": round(k / lat_s.mean(), 3), "ms": round(1e3 * lat_s.mean(), 3), "iterations_max": int(max(its_s))},
           "sequential_fp64": {"solves_per_s": round(k / lat_s64.mean(), 3), "ms": round(1e3 * lat_s64.mean(), 3), "iterations_max": int(max(its_s64))},
           "speedup_vs_sequential_fp64": round(lat_s64.mean() / lat_b.mean(), 3),
           "speedup_vs_sequential_mixed": round(lat_s.mean() / lat_b.mean(), 3),
           "batched_mfma_own": {"solves_per_s": round(k / lat_mf.mean(), 3), "ms": round(1e3 * lat_mf.mean(), 3), "rel_error_max": max(e_mf),
                                "iterations_max": int(max(i["iterations"] for i in infos_mf)),
                                "form": "also the coarse levels' own element products as BF16 MFMA (v_mfma_f32_16x16x32_bf16) over the 16 cases"},
           "batched_valu_degree2": {"solves_per_s": round(k / lat_el.mean(), 3), "ms": round(1e3 * lat_el.mean(), 3), "rel_error_max": max(e_el),
                                    "iterations_max": int(max(i["iterations"] for i in infos_el)),
                                    "form": "Chebyshev degree 2, FP32 VALU smoothing"}}
    m.close()
    return out


def build_only(args, P, w, coarsest):
    """Hierarchy build (levels, Galerkin coarse elements, coarsest inverse) plus the Lanczos smoother
    estimates of one part, from host arrays to a model ready to solve: device kernels against the host
    CPU build, --steps times each after one warm-up."""
    out = {}
    for name, dev in (("device", True), ("host", False)):
        ts = []
        for s in range(args.steps + 1):
            t = time.perf_counter()
            m = P.Model(w["dims"], w["density"], w["nu"], w["bc"], coarsest, device_build=dev)
            m.estimate_smoothers()
            dt = time.perf_counter() - t
            levels = m.num_levels
            m.close()
            if s:
                ts.append(dt)
        out[name] = {"ms_mean": round(1e3 * float(np.mean(ts)), 2), "ms_min": round(1e3 * float(np.min(ts)), 2)}
        log("%s build + smoother estimates: %.1f ms mean, %.1f ms best (%d levels)" % (name, 1e3 * np.mean(ts), 1e3 * np.min(ts), levels))
    print(json.dumps({"metric": "hierarchy build + smoother estimates (ms)", "workload": w["name"], "dof": 3 * int(np.prod([d + 1 for d in w["dims"]])),
                      "steps": args.steps, "device": out["device"], "host": out["host"],
                      "speedup": round(out["host"]["ms_mean"] / out["device"]["ms_mean"], 2)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--dims", default="512,128,128")
    ap.add_argument("--tol", type=float, default=1e-6, help="relative residual tolerance (the reference solver's default, voxel_fea.hpp:113)")
    ap.add_argument("--degree", type=int, default=3, help="Chebyshev smoother degree (3: the workload as configured)")
    ap.add_argument("--smoother", default="mfma", choices=["mfma", "valu"],
                    help="mixed precision: level-0 and level-1 smoothing as the element products on the matrix cores, or the FP32 VALU stencils")
    ap.add_argument("--pre-degree", type=int, default=0, help="Chebyshev degree of the pre-smoothing (0 = --degree)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extras", action="store_true", help="skip the extra legs of the default line (C3 as specified, batched, serving)")
    ap.add_argument("--serve-requests", type=int, default=40, help="requests of the default line's serving leg")
    ap.add_argument("--serve", type=int, default=0, help="serving workload: this many requests of the Poisson stream")
    ap.add_argument("--rate", type=float, default=200.0, help="serving arrival rate (requests/s)")
    ap.add_argument("--inflight", type=int, default=4, help="serving: admission workers (requests built or solved alone at once)")
    ap.add_argument("--pack-max", type=int, default=0, help="serving: requests in the shared-launch pack (0 = the server's 16, 1 = no packing)")
    ap.add_argument("--edge-max", type=int, default=256, help="serving: longest part edge of the stream (voxels)")
    ap.add_argument("--pack-max-dof", type=int, default=0, help="serving: largest part (DOFs) that joins the pack (0 = the server's 100 000)")
    ap.add_argument("--packed", type=int, default=0, help="packed solves: this many small parts of the serving stream, lone vs one pack")
    ap.add_argument("--batched", action="store_true", help="batched load cases: the 128x64x64 SIMP block, --nrhs cases in one multi-instance solve")
    ap.add_argument("--nrhs", type=int, default=16)
    ap.add_argument("--profile-solves", default="", choices=["", "mixed", "fp64", "ref"],
                    help="profiling runs: build on the device, then --steps solves of this precision only (no reference solve, no CPU leg)")
    ap.add_argument("--build-only", action="store_true",
                    help="hierarchy build + smoother estimates of the workload, on the device and on the host, --steps times each")
    ap.add_argument("--reps", type=int, default=50, help="back-to-back launches per kernel of the per-level timings (the roofline figure)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="counter runs: three 1 GiB device copies (the calibration kernel), then the per-level operator and V-cycle timings")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (a part or request runs wholly on one GPU)", file=sys.stderr)
        sys.exit(2)
    P = _package()
    import parts_sim_amd.workloads as W
    if args.serve:
        serve(args, P, W)
        return
    if args.batched:
        batched(args, P, W)
        return
    if args.packed:
        print(json.dumps(packed_leg(P, W, args.packed, args.tol, max(1, args.steps // 2), eager=bool(os.environ.get("PSX_GRAPHS") == "0"))), flush=True)
        return

    dims = tuple(int(v) for v in args.dims.split(","))
    coarsest = 1200
    w = W.void_block(dims, seed=1)
    n_dof = 3 * W.nodes(dims)
    log("workload", w["name"], "DOF", n_dof, "void fraction %.3f" % w["void_fraction"])
    if args.build_only:
        build_only(args, P, w, coarsest)
        return
    if args.profile_solves:
        m = P.Model(w["dims"], w["density"], w["nu"], w["bc"], coarsest, device_build=True)
        m.estimate_smoothers()
        d_f, d_u = P.DeviceArray(n_dof, w["f"]), P.DeviceArray(n_dof)
        for k in range(args.steps):
            t = time.perf_counter()
            info = m.solve_device(d_f, d_u, tol=args.tol, max_iter=2000, precision=args.profile_solves, check_every=8, cheb_degree=args.degree,
                                  mfma_smoother=1 if args.profile_solves == "mixed" and args.smoother == "mfma" else 0)
            log("%s solve %d: %d iterations, %.2f ms" % (args.profile_solves, k, info["iterations"], 1e3 * (time.perf_counter() - t)))
        return
    # hierarchy build + smoother estimates: on the host CPU (first, so that it also carries the
    # process's GPU start-up) and on the device (the model the solves use)
    t0 = time.perf_counter()
    mh = P.Model(w["dims"], w["density"], w["nu"], w["bc"], coarsest)
    mh.estimate_smoothers()
    setup_host_s = time.perf_counter() - t0
    mh.close()
    t0 = time.perf_counter()
    m = P.Model(w["dims"], w["density"], w["nu"], w["bc"], coarsest, device_build=True)
    m.estimate_smoothers()
    setup_s = time.perf_counter() - t0
    log("setup %.3f s on the device (host build %.3f s), levels %d, device MB %.0f" % (setup_s, setup_host_s, m.num_levels, m.device_bytes / 2 ** 20))
    if args.kernels_only:
        a, b = P.DeviceArray(1 << 27), P.DeviceArray(1 << 27)
        for _ in range(3):
            b.copy_from(a)
        # the same launches and repetitions as the bench line's roofline figure
        print(m.profile(reps=args.reps), flush=True)
        return
    d_f = P.DeviceArray(n_dof, w["f"])
    d_u = P.DeviceArray(n_dof)

    # FP64 reference solution (tight tolerance) for the error of every timed solve
    ref = m.solve_device(d_f, d_u, tol=1e-11, max_iter=2000, precision="fp64")
    u_ref = d_u.download()
    log("FP64 reference solve: %d its, residual %.2e" % (ref["iterations"], ref["residual"]))

    def run(precision, steps, warmup, **kw):
        lat, errs, its, us = [], [], [], []
        for k in range(warmup + steps):
            t = time.perf_counter()
            info = m.solve_device(d_f, d_u, tol=args.tol, max_iter=2000, precision=precision, check_every=8, **kw)
            dt = time.perf_counter() - t
            if k >= warmup:
                u = d_u.download()
                lat.append(dt)
                errs.append(float(np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref)))
                its.append(info["iterations"])
                if not info["converged"] or errs[-1] > 1e-6:
                    raise RuntimeError("%s solve missed the accuracy bar: %s err %.2e" % (precision, info, errs[-1]))
                if not us or not np.array_equal(u, us[-1]):
                    us.append(u)
        return np.array(lat), np.array(errs), its, info, us

    # timed region: barrier-equivalent (the device is idle between calls: every solve ends on a
    # host read of its final residual) and one stream
    mf = 1 if args.smoother == "mfma" else 0
    lat64, err64, its64, info64, us64 = run("fp64", max(3, args.steps // 2), 1, cheb_degree=args.degree, pre_degree=args.pre_degree)
    log("FP64 MG-PCG: %.2f ms/solve, %d its, err %.2e" % (1e3 * lat64.mean(), its64[0], err64.max()))
    lat, err, its, info, us = run("mixed", args.steps, args.warmup, cheb_degree=args.degree, pre_degree=args.pre_degree, mfma_smoother=mf)
    log("mixed: %.2f ms/solve (p50 %.2f p99 %.2f), %d its, %d folds, err %.2e" % (1e3 * lat.mean(), 1e3 * np.percentile(lat, 50),
                                                                                   1e3 * np.percentile(lat, 99), its[0], info["folds"], err.max()))
    solves_per_s = len(lat) / lat.sum()
    fp64_per_s = len(lat64) / lat64.sum()
    ref_mode = m.solve_device(d_f, d_u, tol=args.tol, max_iter=2000, precision="ref", check_every=8)

    # roofline of the dominant kernel: the level-0 Chebyshev step y = x + c1 (x - d) + c2 D^-1 (r - K x)
    # (three of the V-cycle's six level-0 launches at degree 3), in the smoother's form, and of the outer
    # FP32 operator y = K x; timed with HIP events on the model's stream over back-to-back launches
    prof = m.profile(reps=args.reps)
    line0 = [l for l in prof.splitlines() if l.startswith("L0 ")][0]
    field = lambda name: float(line0.split(name + " ")[1].split("ms")[0]) * 1e-3
    t_apply32 = field("matvec32")
    t_smooth = field("smooth-mfma2" if mf else "smooth32")
    nodes = W.nodes(dims)
    nvox = dims[0] * dims[1] * dims[2]
    alg_bytes = 24.0 * nodes + 4.0 * nvox + 2.0 * nodes  # x read + y written (FP32), voxel word, node info word
    smooth_bytes = 48.0 * nodes + 4.0 * nvox + 2.0 * nodes  # x, r, d read + y written (FP32), voxel word, node info word
    achieved = alg_bytes / t_apply32
    achieved_s = smooth_bytes / t_smooth
    log(prof.strip())
    traffic = measured_traffic(dims)
    traffic_s = measured_traffic(dims, SMOOTH_TRAFFIC_FILE) if mf else None

    levels = m.num_levels
    extras = {}
    if not args.no_extras:
        # degree-2 smoothing with the FP32 VALU smoother, both precisions (the configuration of the
        # earlier figures; faster per iteration, more iterations)
        lat2, err2, its2, info2, us2 = run("mixed", 3, 1, cheb_degree=2)
        lat2d, err2d, its2d, info2d, us2d = run("fp64", 3, 1, cheb_degree=2)
        extras["degree2_valu"] = {"solves_per_s": round(len(lat2) / lat2.sum(), 3), "ms_per_solve": round(1e3 * lat2.mean(), 3),
                                  "p99_ms": round(1e3 * float(np.percentile(lat2, 99)), 3), "iterations": int(its2[0]), "folds": int(info2["folds"]),
                                  "fp64_solves_per_s": round(len(lat2d) / lat2d.sum(), 3), "fp64_iterations": int(its2d[0]),
                                  "speedup_mixed_vs_fp64": round(lat2d.mean() / lat2.mean(), 3),
                                  "smoother": "Chebyshev degree 2, FP32 VALU stencil", "rel_error_vs_gpu_fp64_1e-11": float(err2.max())}
        log("degree 2, VALU smoother: mixed %.2f ms/solve, %d its; FP64 %.2f ms" % (1e3 * lat2.mean(), its2[0], 1e3 * lat2d.mean()))
        m.close()
        for name, fn in (("batched", lambda: batched_leg(P, W, (128, 64, 64), 16, max(3, args.steps // 2), 1, args.tol, args.degree, mf)),
                         ("serving", lambda: serving_leg(P, W, args.serve_requests, args.rate, args.inflight, args.tol, args.pack_max,
                                                         args.pack_max_dof, 256, args.degree, mf)),
                         ("packed", lambda: packed_leg(P, W, 16, args.tol))):
            t = time.perf_counter()
            try:
                extras[name] = fn()
            except Exception as e:  # reported in the line, not hidden
                extras[name] = {"error": "%s: %s" % (type(e).__name__, e)}
            log("%s leg: %.1f s" % (name, time.perf_counter() - t))

    cpu_base, u_oracle = None, None
    if not args.no_cpu:
        cpu_base, u_oracle = cpu_baseline(w, coarsest, args.tol)
        log("CPU baseline:", cpu_base)

    def oracle_err(ulist):
        return max(float(np.linalg.norm(u - u_oracle) / np.linalg.norm(u_oracle)) for u in ulist)

    if u_oracle is not None:
        err_mixed, err_fp64 = oracle_err(us), oracle_err(us64)
        if "degree2_valu" in extras:
            extras["degree2_valu"]["rel_error_vs_fp64"] = oracle_err(us2)
        err_source = "the reference's own FP64 VoxelFEA::solve (oracle/_ref) to a relative residual of 1e-10"
    else:
        err_mixed, err_fp64 = float(err.max()), float(err64.max())
        err_source = "a plain FP64 GPU solve to a relative residual of 1e-11 (oracle/_ref absent)"

    out = {
        "metric": METRIC, "value": round(solves_per_s, 3), "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * lat.mean(), 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "void block %dx%dx%d, seeded spherical voids 20%% at 1e-3 stiffness, E 70 GPa, nu 0.33, x=0 clamped, "
                               "5 MPa top pressure x>200 mm, %d-level MG, Chebyshev degree %d, level-0 and level-1 smoothing %s" %
                               (dims + (levels, args.degree, "as element products on the matrix cores (BF16 pieces, FP32 accumulation)" if mf
                                        else "as the FP32 VALU stencil")),
                   "dof": n_dof, "levels": levels, "tol": args.tol, "parallelism": "single GPU"},
        "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3), "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
        "rel_error_vs_fp64": err_mixed, "rel_error_reference": err_source, "rel_error_vs_gpu_fp64_1e-11": float(err.max()),
        "iterations": int(its[0]), "folds": int(info["folds"]),
        "dof_iterations_per_s": float(n_dof * its[0] / lat.mean()),
        "fp64": {"solves_per_s": round(fp64_per_s, 3), "ms_per_solve": round(1e3 * lat64.mean(), 3), "iterations": int(its64[0]),
                 "p50_ms": round(1e3 * float(np.percentile(lat64, 50)), 3), "p99_ms": round(1e3 * float(np.percentile(lat64, 99)), 3),
                 "rel_error_vs_fp64": err_fp64},
        "speedup_mixed_vs_fp64": round(solves_per_s / fp64_per_s, 3),
        "reference_algorithm_iterations": int(ref_mode["iterations"]),
        "setup_s": round(setup_s, 3), "setup_host_build_s": round(setup_host_s, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved_s / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(achieved_s / HBM_PEAK, 4), "traffic": traffic_s,
                     "kernel": ("level-0 Chebyshev step as element products on the matrix cores (kFineEl<CHEB2>)" if mf
                                else "level-0 Chebyshev step, FP32 VALU stencil (kFine4<float, CHEB2>)"),
                     "algorithmic_bytes": smooth_bytes, "avg_ms": round(t_smooth * 1e3, 4)},
        "roofline_outer_operator": {"bound": "hbm", "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                                    "frac": round(achieved / HBM_PEAK, 4), "traffic": traffic,
                                    "kernel": "fine-level FP32 operator y = K x (kFine4<float, APPLY>)", "algorithmic_bytes": alg_bytes,
                                    "avg_ms": round(t_apply32 * 1e3, 4)},
        "cpu_baseline": cpu_base,
    }
    out.update(extras)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
