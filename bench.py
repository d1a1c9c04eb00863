"""Headline benchmark: solves/s and p50/p99 latency at relative residual 1e-8 for the
single large solve, mixed precision vs plain FP64 on one MI355X, plus the dominant kernel's
HBM roofline fraction and the CPU oracle timed on the host's cores.

Workload: 3D variable-coefficient diffusion -div(k grad u) = f on [0,1]^3, 257^3 nodes, 7-point
flux stencil, k = exp(0.5 g) with g a seeded Gaussian random field smoothed by a Gaussian filter
of sigma = 4 cells and normalised to unit variance, f seeded N(0,1) with zero mean, Dirichlet 0
on the x/y faces, Robin (alpha=1, beta=1, gamma=0, outward normal) on the z faces.  MG-preconditioned
CG, V(2,2) red-black Gauss-Seidel, 6 levels (257..9, dense FP64 coarsest solve).  The plain
path is FP64 throughout; the mixed path runs FP32 MG-PCG inner solves with FP16 coarse-level
storage inside FP64 iterative refinement.  A step is one complete solve from a zero initial
guess with the operator, hierarchy and right-hand side resident in HBM.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--n 257] [--no-cpu-baseline]
    python bench.py --only mixed|fp64 [--steps K] [--warmup W]   (one path only, for profiling)
    python bench.py --workload batch [--steps K] [--count 512]  (512 uneven 2D problems)
    python bench.py --workload nonsym [--count 24] [--rd-steps 1000]
    python bench.py --workload serve [--count 400] [--rates 50,200]
    python bench.py --kernels 4,3 [--precision mixed|fp64] [--steps K]
        (K launches of plan kernels only: 0 apply, 1 in-place sweep, 2 residual, 3 copy,
         4 fused sweep; for hardware-counter runs)
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solves/s + p99 latency @ rel-resid 1e-8, mixed vs FP64 GPU; % HBM peak"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def make_problem(n: int):
    from scipy.ndimage import gaussian_filter

    from pde_solver_amd import capi
    from pde_solver_amd.capi import BC

    g = np.random.default_rng(2).standard_normal((n, n, n))
    g = gaussian_filter(g, sigma=4.0, mode="reflect", truncate=4.0)
    g = (g - g.mean()) / g.std()
    f = np.random.default_rng(20).standard_normal((n, n, n))
    f -= f.mean()
    s = capi.Spec(n, n, n, a=-1.0, op_kind=capi.OP_FLUX)
    s.kappa = np.exp(0.5 * g)
    s.f_field = -f                        # reference sign: L u + f = 0
    s.bcs = [BC.dirichlet(0.0)] * 4 + [BC.robin(1.0, 1.0, 0.0)] * 2
    s.bc_convention = capi.BCCONV_OUTWARD
    return s


def make_batch_problems(count: int = 512, seed: int = 3):
    """512 independent 2D problems: n = 2^(5+K)+1 with K ~ U{0..5} (log-uniform over
    33..1025); the first half Poisson -lap u = f, the rest Helmholtz -lap u + c u = f with
    c ~ U(0, 50); f a sum of 8 Fourier modes sin(pi p x + phi) sin(pi q y + psi) with p, q ~
    U{1..8}, amplitudes N(0, 1), phases U(0, 2 pi); each face Dirichlet 0 / Neumann 0 /
    Robin(1, 1, 0) (outward normal) drawn uniformly.  A Poisson problem that drew Neumann on
    all four faces is singular; it gets the Helmholtz shift c = 1."""
    from pde_solver_amd import capi
    from pde_solver_amd.capi import BC

    rng = np.random.default_rng(seed)
    kinds = [BC.dirichlet(0.0), BC.neumann(0.0), BC.robin(1.0, 1.0, 0.0)]
    specs = []
    for t in range(count):
        n = 2 ** (5 + int(rng.integers(0, 6))) + 1
        c = 0.0 if t < count // 2 else float(rng.uniform(0.0, 50.0))
        x = np.linspace(0.0, 1.0, n)
        f = np.zeros((n, n))
        for _ in range(8):
            p, q = rng.integers(1, 9, size=2)
            amp = rng.standard_normal()
            phi, psi = rng.uniform(0.0, 2 * np.pi, size=2)
            f += amp * np.outer(np.sin(np.pi * q * x + psi), np.sin(np.pi * p * x + phi))
        faces = [kinds[k] for k in rng.integers(0, 3, size=4)]
        if c == 0.0 and all(fc.kind == capi.NEUMANN for fc in faces):
            c = 1.0
        s = capi.Spec(n, n, a=-1.0, b=-1.0, e=c)
        s.f_field = -f                     # reference sign: L u + f = 0
        s.bcs = faces + [BC(), BC()]
        s.bc_convention = capi.BCCONV_OUTWARD
        specs.append(s)
    return specs


def fourier_source(n, rng):
    """Smooth seeded source: sum of 8 modes sin(pi p x + phi) sin(pi q y + psi), p, q ~ U{1..8},
    amplitudes N(0, 1), phases U(0, 2 pi)."""
    x = np.linspace(0.0, 1.0, n)
    f = np.zeros((n, n))
    for _ in range(8):
        p, q = rng.integers(1, 9, size=2)
        amp = rng.standard_normal()
        phi, psi = rng.uniform(0.0, 2 * np.pi, size=2)
        f += amp * np.outer(np.sin(np.pi * q * x + psi), np.sin(np.pi * p * x + phi))
    return f


def make_nonsym_problems(count: int = 128, seed: int = 4):
    """Convection-diffusion -eps lap u + b.grad u = f, first-order upwind, Dirichlet 0:
    n = 2^(7+K)+1 with K ~ U{0..3} (129..1025), eps = 10^U(-3,-1), b = (cos t, sin t) with
    t ~ U(0, 2 pi), 8-mode Fourier source."""
    from pde_solver_amd import capi

    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(count):
        n = 2 ** (7 + int(rng.integers(0, 4))) + 1
        eps = float(10.0 ** rng.uniform(-3.0, -1.0))
        th = float(rng.uniform(0.0, 2 * np.pi))
        s = capi.Spec(n, n, a=-eps, b=-eps, c=float(np.cos(th)), d=float(np.sin(th)),
                      convection=capi.CONV_UPWIND)
        s.f_field = -fourier_source(n, rng)
        specs.append(s)
    return specs


def bench_nonsym(args) -> int:
    """Nonsymmetric workload: Jacobi-preconditioned BiCGStab and GMRES(30) to rel-resid 1e-8 on
    the convection-diffusion problems (a bounded sample, one resident plan after the other),
    mixed precision vs FP64; and reaction-diffusion u_t = lap u + u(1 - u) on 2049^2, 1000 RK4
    steps at dt = 0.2 h^2, FP32 state vs FP64."""
    from scipy.ndimage import gaussian_filter
    from pde_solver_amd import capi, hip
    from pde_solver_amd.capi import BC

    ok, note = hip.is_available()
    log(f"device: {note}")
    if not ok:
        raise SystemExit("no usable gfx9 device")
    specs = make_nonsym_problems(128)[: args.count]
    ctx = hip.Context()

    def kopts(method, precision, tol=1e-8):
        return capi.options(method, tol=tol, tol_mode=capi.TOL_REL, precond=capi.PC_JACOBI,
                            bc_rows=capi.ROWS_FOLDED, gmres_restart=30, max_iter=10000,
                            precision=precision)

    results = {}
    for method, mname in ((capi.BICGSTAB, "bicgstab"), (capi.GMRES, "gmres")):
        sols, oks = {}, {}
        for precision, pname in ((capi.FP64, "fp64"), (capi.MIXED, "mixed")):
            plans = [ctx.plan(s, kopts(method, precision)) for s in specs]
            lat, its, conv, us, ok = [], [], 0, [], []
            t0 = time.perf_counter()
            for pl in plans:
                t1 = time.perf_counter()
                st, rep = pl.solve(check=False)
                lat.append((time.perf_counter() - t1) * 1e3)
                its.append(rep.inner_iterations if precision == capi.MIXED else rep.iterations)
                conv += int(st == 0)
                ok.append(st == 0)
            wall = time.perf_counter() - t0
            for pl in plans:
                us.append(pl.download())
                pl.close()
            sols[pname] = us
            oks[pname] = ok
            results[f"{mname}_{pname}"] = {
                "solves_per_s": round(len(specs) / wall, 2), "converged": conv,
                "problems": len(specs), "latency_ms_p50": round(pct(lat, 50), 2),
                "latency_ms_p99": round(pct(lat, 99), 2), "iterations_total": int(sum(its))}
            log(f"{mname} {pname}: {results[f'{mname}_{pname}']['solves_per_s']} solves/s, "
                f"{conv}/{len(specs)} converged, {sum(its)} iterations")
        errs = [float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
                for a, b in zip(sols["mixed"], sols["fp64"])]
        results[f"{mname}_mixed"]["max_diff_vs_fp64"] = max(errs)
        if method == capi.BICGSTAB:
            seq64, seq64_ok = sols["fp64"], oks["fp64"]
        else:
            gm64, gm64_ok = sols["fp64"], oks["fp64"]

    # batched BiCGStab: the same problems solved together in lock step by a resident batch
    # (one launch per operation over every unfinished problem, 8 iterations per graph replay),
    # on the sample above and on all 128 problems; FP64 and mixed (FP32 Krylov vectors, FP64
    # solution and residual).  The first solve of a batch captures its graph and is not timed.
    all_specs = make_nonsym_problems(128)
    for label, sp in (("sample", specs), ("all", all_specs)):
        for precision, pname in ((capi.FP64, "fp64"), (capi.MIXED, "mixed")):
            bt = ctx.batch(sp, kopts(capi.BICGSTAB, precision), capi.BatchConfig(use_graph=1))
            bt.solve(check=False)
            t0 = time.perf_counter()
            st, reps = bt.solve(check=False)
            wall = time.perf_counter() - t0
            key = f"bicgstab_batched_{label}_{pname}"
            results[key] = {"solves_per_s": round(len(sp) / wall, 2), "batch_ms": round(wall * 1e3, 1),
                            "converged": int(sum(r.status == 0 for r in reps)), "problems": len(sp),
                            "iterations_max": int(max(r.iterations for r in reps)),
                            "iterations_total": int(sum(r.iterations for r in reps))}
            if label == "sample":   # over the problems both solves converged on
                diffs = [float(np.linalg.norm(bt.download(i) - seq64[i]) / max(np.linalg.norm(seq64[i]), 1e-300))
                         for i in range(len(sp)) if reps[i].status == 0 and seq64_ok[i]]
                results[key]["max_diff_vs_sequential_fp64"] = max(diffs) if diffs else None
            bt.close()
            log(f"bicgstab batched {label} {pname}: {results[key]['solves_per_s']} solves/s "
                f"({results[key]['batch_ms']} ms per batch), {results[key]['converged']}/{len(sp)} converged")

    # batched GMRES(30), FP64: per-problem Arnoldi bases and Hessenberg systems in lock step, one
    # graph per restart cycle (mixed-precision GMRES is not batched)
    for label, sp in (("sample", specs), ("all", all_specs)):
        bt = ctx.batch(sp, kopts(capi.GMRES, capi.FP64), capi.BatchConfig(use_graph=1))
        bt.solve(check=False)
        t0 = time.perf_counter()
        st, reps = bt.solve(check=False)
        wall = time.perf_counter() - t0
        key = f"gmres_batched_{label}_fp64"
        results[key] = {"solves_per_s": round(len(sp) / wall, 2), "batch_ms": round(wall * 1e3, 1),
                        "converged": int(sum(r.status == 0 for r in reps)), "problems": len(sp),
                        "iterations_max": int(max(r.iterations for r in reps)),
                        "iterations_total": int(sum(r.iterations for r in reps))}
        if label == "sample":
            diffs = [float(np.linalg.norm(bt.download(i) - gm64[i]) / max(np.linalg.norm(gm64[i]), 1e-300))
                     for i in range(len(sp)) if reps[i].status == 0 and gm64_ok[i]]
            results[key]["max_diff_vs_sequential_fp64"] = max(diffs) if diffs else None
        bt.close()
        log(f"gmres batched {label} fp64: {results[key]['solves_per_s']} solves/s "
            f"({results[key]['batch_ms']} ms per batch), {results[key]['converged']}/{len(sp)} converged")

    # reaction-diffusion, 2049^2, 1000 RK4 steps
    n = 2049
    h = 1.0 / (n - 1)
    rd = capi.Spec(n, n, a=-1.0, b=-1.0, ut=1.0)
    rd.nonlinear = [(capi.NL_POWER, -1.0, 1), (capi.NL_POWER, 1.0, 2)]   # N(u) = -u + u^2
    rd.initial = gaussian_filter(np.random.default_rng(44).uniform(0.0, 1.0, (n, n)), 2.0,
                                 mode="reflect")
    rd.bcs = [BC.neumann(0.0)] * 4 + [BC(), BC()]
    steps = args.rd_steps
    rdres = {}
    for precision, pname in ((capi.MIXED, "fp32"), (capi.FP64, "fp64")):
        tc = capi.time_config(capi.TI_RK4, dt=0.2 * h * h, frames=steps + 1, precision=precision)
        es = 4 if precision == capi.MIXED else 8
        for fused in (1, 0):
            hip.tune_set("time.fused", fused)
            ctx.time_series(rd, capi.time_config(capi.TI_RK4, dt=0.2 * h * h, frames=3, precision=precision),
                            keep_frames=False)
            _, u, rep = ctx.time_series(rd, tc, keep_frames=False)
            sec = rep.solve_ms / 1e3
            if fused:
                # one launch per step: u_n read once, u_{n+1} written once
                rdres[pname] = {"steps": steps, "ms": round(rep.solve_ms, 2),
                                "us_per_step": round(rep.solve_ms * 1e3 / steps, 2),
                                "steps_per_s": round(steps / sec, 1),
                                "GB_per_s": round(2 * es * rd.n * steps / sec / 1e9, 1),
                                "hbm_frac": round(2 * es * rd.n * steps / sec / 1e9 / HBM_PEAK_GBS, 3),
                                "stage_equivalent_GB_per_s": round(17 * es * rd.n * steps / sec / 1e9, 1),
                                "u_mean": float(u.mean())}
                rdres[pname + "_u"] = u
            else:
                # one launch per stage (17 state accesses per node and step)
                rdres[pname]["per_stage_launches_ms"] = round(rep.solve_ms, 2)
                rdres[pname]["per_stage_GB_per_s"] = round(17 * es * rd.n * steps / sec / 1e9, 1)
                rdres[pname]["fused_equals_per_stage"] = bool(np.array_equal(u, rdres[pname + "_u"]))
        hip.tune_set("time.fused", 1)
        log(f"RD {pname}: fused {rdres[pname]['ms']} ms for {steps} steps ({rdres[pname]['GB_per_s']} GB/s), "
            f"per-stage launches {rdres[pname]['per_stage_launches_ms']} ms, "
            f"bitwise equal: {rdres[pname]['fused_equals_per_stage']}")
    rdres["fp32"]["max_rel_diff_vs_fp64"] = float(np.abs(rdres.pop("fp32_u") - rdres["fp64_u"]).max()
                                                   / np.abs(rdres["fp64_u"]).max())
    rdres.pop("fp64_u")

    mixed_total = 2 * len(specs) / (len(specs) / results["bicgstab_mixed"]["solves_per_s"]
                                     + len(specs) / results["gmres_mixed"]["solves_per_s"])
    fp64_total = 2 * len(specs) / (len(specs) / results["bicgstab_fp64"]["solves_per_s"]
                                    + len(specs) / results["gmres_fp64"]["solves_per_s"])
    out = {
        "metric": METRIC,
        "value": round(mixed_total, 2),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": round(1e3 / mixed_total, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 Krylov vectors + f64 solution/residual",
        "data": "synthetic",
        "config": {"workload": f"nonsymmetric: the first {len(specs)} of 128 upwind "
                               "convection-diffusion problems (129^2..1025^2, eps 10^U(-3,-1)), "
                               "Jacobi-PC BiCGStab and GMRES(30) to rel-resid 1e-8, one resident "
                               "plan after the other; RD 2049^2 RK4 1000 steps",
                   "problems": len(specs)},
        "krylov": results,
        "fp64_solves_per_s": round(fp64_total, 2),
        "bicgstab_batched_all_mixed_solves_per_s": results["bicgstab_batched_all_mixed"]["solves_per_s"],
        "bicgstab_batched_all_fp64_solves_per_s": results["bicgstab_batched_all_fp64"]["solves_per_s"],
        "gmres_batched_all_fp64_solves_per_s": results["gmres_batched_all_fp64"]["solves_per_s"],
        "reaction_diffusion_2049": rdres,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


def make_serve_requests(count: int, seed: int = 5):
    """Request stream: 2D (3 in 4) with n = 2^(5+K)+1, K ~ U{0..6}, or 3D with K ~ U{0..3};
    operator Poisson / lognormal variable-coefficient diffusion (flux form) / upwind
    convection-diffusion; MG-PCG for the symmetric ones, Jacobi-PC BiCGStab or GMRES(30) for
    convection-diffusion (kept to 2D n <= 513 and 3D n <= 33, where GMRES(30) converges);
    rel-resid 1e-8.  Coefficient fields and sources are drawn once per size and shared."""
    from scipy.ndimage import gaussian_filter
    from pde_solver_amd import capi
    from pde_solver_amd.capi import BC

    rng = np.random.default_rng(seed)
    kappas, sources = {}, {}

    def field(key, shape, maker):
        if key not in maker[0]:
            maker[0][key] = maker[1](shape)
        return maker[0][key]

    def kappa_of(shape):
        g = gaussian_filter(np.random.default_rng(len(shape) * 1000 + shape[-1]).standard_normal(shape),
                            4.0, mode="reflect")
        g = (g - g.mean()) / max(g.std(), 1e-12)
        return np.exp(0.5 * g)

    def source_of(shape):
        f = np.random.default_rng(shape[-1] + 7 * len(shape)).standard_normal(shape)
        return f - f.mean()

    reqs = []
    for _ in range(count):
        d3 = rng.uniform() < 0.25
        n = 2 ** (5 + int(rng.integers(0, 4 if d3 else 7))) + 1
        kind = int(rng.integers(0, 3))
        if kind == 2 and n > (33 if d3 else 513):
            kind = int(rng.integers(0, 2))
        shape = (n, n, n) if d3 else (n, n)
        if kind == 0:
            s = capi.Spec(*shape[::-1], a=-1.0, b=-1.0, az=-1.0 if d3 else 0.0)
        elif kind == 1:
            s = capi.Spec(*shape[::-1], a=-1.0, op_kind=capi.OP_FLUX)
            s.kappa = field(shape, shape, (kappas, kappa_of))
        else:
            eps = float(10.0 ** rng.uniform(-3.0, -1.0))
            th = float(rng.uniform(0.0, 2 * np.pi))
            s = capi.Spec(*shape[::-1], a=-eps, b=-eps, az=-eps if d3 else 0.0,
                          c=float(np.cos(th)), d=float(np.sin(th)), convection=capi.CONV_UPWIND)
        s.f_field = field(shape, shape, (sources, source_of))
        if kind < 2 and d3:
            s.bcs = [BC()] * 4 + [BC.robin(1.0, 1.0, 0.0)] * 2
            s.bc_convention = capi.BCCONV_OUTWARD
        elif kind < 2:
            kinds = [BC.dirichlet(0.0), BC.neumann(0.0), BC.robin(1.0, 1.0, 0.0)]
            faces = [kinds[k] for k in rng.integers(0, 3, size=4)]
            if all(f.kind == capi.NEUMANN for f in faces):
                faces[0] = BC.dirichlet(0.0)
            s.bcs = faces + [BC(), BC()]
            s.bc_convention = capi.BCCONV_OUTWARD
        method = capi.MGPCG if kind < 2 else (capi.BICGSTAB if rng.uniform() < 0.5 else capi.GMRES)
        reqs.append((s, method))
    return reqs


def bench_serve(args) -> int:
    """Serving: Poisson arrivals at each rate, requests submitted to the continuous-batching
    server as they arrive, latency from arrival to completion; mixed precision vs FP64."""
    from pde_solver_amd import capi, hip

    ok, note = hip.is_available()
    log(f"device: {note}")
    if not ok:
        raise SystemExit("no usable gfx9 device")
    reqs = make_serve_requests(args.count)
    rates = [float(r) for r in args.rates.split(",")]
    ctx = hip.Context()

    def opts(method, precision):
        if method == capi.MGPCG:
            return capi.options(capi.MGPCG, tol=1e-8, tol_mode=capi.TOL_REL, precond=capi.PC_MG,
                                bc_rows=capi.ROWS_FOLDED, precision=precision, max_iter=200)
        return capi.options(method, tol=1e-8, tol_mode=capi.TOL_REL, precond=capi.PC_JACOBI,
                            bc_rows=capi.ROWS_FOLDED, gmres_restart=30, max_iter=20000,
                            precision=precision)

    results = {}
    hi = max(rates)
    runs = [(capi.MIXED, "mixed", lam, 2) for lam in rates] + \
           [(capi.FP64, "fp64", lam, 2) for lam in rates] + \
           [(capi.MIXED, "mixed_no_batch_lanes", hi, -1), (capi.FP64, "fp64_no_batch_lanes", hi, -1)]
    for precision, pname, lam, lanes in runs:
        srv = ctx.server(capi.ServerConfig(max_inflight=256, streams=4, cache_operators=1,
                                           batch_lanes=lanes, batch_max=64))
        arr = np.cumsum(np.random.default_rng(55).exponential(1.0 / lam, size=len(reqs)))
        pending = {}
        lat, fails = [], 0
        t0 = time.perf_counter()
        k = 0
        last_done = t0
        while k < len(reqs) or pending:
            now = time.perf_counter() - t0
            while k < len(reqs) and arr[k] <= now:
                spec, method = reqs[k]
                pending[srv.submit(spec, opts(method, precision))] = arr[k]
                k += 1
            for tk in list(pending):
                st, _, rep = srv.poll(tk)
                if st == capi.PENDING:
                    continue
                tdone = time.perf_counter() - t0
                lat.append((tdone - pending.pop(tk)) * 1e3)
                fails += int(st != 0)
                last_done = tdone
            if k < len(reqs) and not pending:
                time.sleep(max(0.0, min(arr[k] - (time.perf_counter() - t0), 0.01)))
            else:
                time.sleep(0.0002)
        stats = srv.stats()
        srv.close()
        key = f"{pname}_lambda{int(lam)}"
        results[key] = {"requests": len(reqs), "failed": fails,
                        "solves_per_s": round(len(reqs) / last_done, 1),
                        "p50_ms": round(pct(lat, 50), 2), "p99_ms": round(pct(lat, 99), 2),
                        "cache_hits": int(stats["cache_hits"]),
                        "batched": int(stats["batched"]), "lane_batches": int(stats["batches"]),
                        "peak_GiB": round(stats["peak_bytes"] / 2**30, 2)}
        log(f"{key}: {results[key]}")
    out = {
        "metric": METRIC,
        "value": results[f"mixed_lambda{int(hi)}"]["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": len(reqs),
        "warmup": 0,
        "ms_per_step": results[f"mixed_lambda{int(hi)}"]["p50_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner solves + f64 refinement",
        "data": "synthetic",
        "config": {"workload": f"serving: {len(reqs)} requests per rate, Poisson arrivals at "
                               f"{args.rates} req/s, 2D 33^2..2049^2 and 3D 33^3..257^3, Poisson / "
                               "var-coef / conv-diff, MG-CG / BiCGStab / GMRES, rel-resid 1e-8",
                   "requests": len(reqs), "rates": rates},
        "latency_ms": {"p50": results[f"mixed_lambda{int(hi)}"]["p50_ms"],
                       "p99": results[f"mixed_lambda{int(hi)}"]["p99_ms"]},
        "serving": results,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


def batch_options(precision: int, tol: float):
    from pde_solver_amd import capi
    return capi.options(capi.MGPCG, tol=tol, tol_mode=capi.TOL_REL, precond=capi.PC_MG,
                        bc_rows=capi.ROWS_FOLDED, precision=precision, mg_pre=2, mg_post=2,
                        max_iter=200, inner_tol=1e-4)


def max_rel_err(us, refs):
    return max(float(np.linalg.norm(u - r) / max(np.linalg.norm(r), 1e-300)) for u, r in zip(us, refs))


def batch_run(ctx, specs, precision, tol, fp16, steps, warmup, u_ref, roofline=False):
    """One resident lock-step batch of every problem (autotuned shapes, graph replay): solves/s,
    batch p50/p99, error against u_ref and, with roofline, the level-0 launches' HIP-event
    timing."""
    from pde_solver_amd import capi, hip
    tuned = capi.BatchConfig(autotune=1, use_graph=1)
    hip.tune_set("mixed.fp16", fp16)
    try:
        t0 = time.perf_counter()
        b = ctx.batch(specs, batch_options(precision, tol), tuned)
        setup = time.perf_counter() - t0
    finally:
        hip.tune_set("mixed.fp16", 1)
    for _ in range(warmup):
        b.solve()
    lat, dev = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        t1 = time.perf_counter()
        _, reps = b.solve()
        lat.append((time.perf_counter() - t1) * 1e3)
        dev.append(reps[0].solve_ms)
    wall = time.perf_counter() - t0
    us = [b.download(i) for i in range(len(specs))]
    out = {"solves_per_s": round(len(specs) * steps / wall, 1),
           "batch_ms_p50": round(pct(lat, 50), 3), "batch_ms_p99": round(pct(lat, 99), 3),
           "device_ms_p50": round(pct(dev, 50), 3), "setup_s": round(setup, 2),
           "max_err_vs_fp64_oracle": max_rel_err(us, u_ref),
           "max_rel_residual": max(r.rel_residual for r in reps),
           "all_converged": all(r.status == 0 for r in reps),
           "resident_GiB": round(b.bytes() / 2**30, 2),
           "iterations_max": max(r.iterations for r in reps),
           "inner_iterations_max": max(r.inner_iterations for r in reps)}
    if roofline:
        kern = {}
        for k, name in ((0, "sweep"), (1, "residual"), (2, "apply_dot"), (3, "residual_fp64")):
            ms, nb = b.time_kernel(k, 10)
            kern[name] = {"avg_ms": round(ms, 4), "GB/s": round(nb / (ms * 1e-3) / 1e9, 1),
                          "algorithmic_bytes": nb}
        out["level0_kernels"] = kern
    b.close()
    return out


def batch_plain(ctx, specs, u_ref):
    """The plain FP64 path: one resident plan per problem, solved one after the other."""
    from pde_solver_amd import capi
    plans = [ctx.plan(s, batch_options(capi.FP64, 1e-8)) for s in specs]
    for pl in plans[:8]:
        pl.solve()
    lat = []
    t0 = time.perf_counter()
    for pl in plans:
        t1 = time.perf_counter()
        pl.solve()
        lat.append((time.perf_counter() - t1) * 1e3)
    wall = time.perf_counter() - t0
    err = max_rel_err([pl.download() for pl in plans], u_ref)
    for pl in plans:
        pl.close()
    return {"solves_per_s": round(len(specs) / wall, 1),
            "latency_ms_p50": round(pct(lat, 50), 3), "latency_ms_p99": round(pct(lat, 99), 3),
            "max_err_vs_fp64_oracle": err}


def batch_cpu(specs, seconds):
    """The CPU oracle's FP64 MG-PCG over the batch's problems in batch order (a random size
    mix) until about `seconds` of work."""
    from oracle import oracle as O
    from pde_solver_amd import capi
    threads = cpu_threads()
    O.set_threads(threads)
    t0 = time.perf_counter()
    done = 0
    while done < len(specs) and (time.perf_counter() - t0 < seconds or done < 8):
        O.solve(specs[done], batch_options(capi.FP64, 1e-8))
        done += 1
    total = time.perf_counter() - t0
    cpu = {"value": round(done / total, 2), "unit": "solves/s", "cores": threads,
           "kind": "port", "build": CPU_BUILD["flags"],
           "sample": f"the first {done} problems of the batch "
                     f"({sum(s.n for s in specs[:done]) / 1e6:.1f}M nodes), FP64 MG-PCG "
                     f"to rel 1e-8, one after the other, {total:.1f}s wall"}
    log(f"cpu baseline (batch): {cpu['value']} solves/s on {threads} threads (sample)")
    return cpu


_THREADS = []


def cpu_threads() -> int:
    """Host cores this process may use, read once: after the OpenMP runtime binds its threads
    (OMP_PROC_BIND=close) the main thread's own affinity is a single core."""
    if not _THREADS:
        t = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        _THREADS.append(max(1, min(t, 64)))
    return _THREADS[0]


CPU_BUILD = {"flags": "-O3 -march=x86-64-v3 (prebuilt)"}


def cpu_setup() -> None:
    """The CPU baseline's conditions: OpenMP threads bound close, one per host core this
    process may use, and the oracle rebuilt with -march=native on this host (the prebuilt
    x86-64-v3 library when no compiler is there).  Called before anything loads OpenMP."""
    import subprocess
    import tempfile
    os.environ["OMP_PROC_BIND"] = "close"
    os.environ["OMP_PLACES"] = "cores"
    os.environ["OMP_NUM_THREADS"] = str(cpu_threads())
    out = tempfile.mkdtemp(prefix="oracle_native_")
    try:
        r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native",
                            f"NATIVE_OUT={out}"], capture_output=True, text=True, timeout=180)
        lib = os.path.join(out, "liboracle.so")
        if r.returncode == 0 and os.path.exists(lib):
            from oracle import oracle as O
            O.use_library(lib)
            CPU_BUILD["flags"] = "-O3 -march=native (built on this host)"
        else:
            log(f"native oracle build failed, using the prebuilt library: {r.stderr[-300:]}")
    except Exception as e:   # no make / compiler on this host
        log(f"native oracle build unavailable ({e}), using the prebuilt library")


def bench_batch(args) -> int:
    """Batched uneven sizes: 512 2D problems per step.  Mixed batched (FP32 MG-PCG with FP64
    residual replacement) vs FP64 batched vs the plain FP64 path (one resident plan solved after
    the other), errors against an FP64 oracle solve (FP64 batch to rel-resid 1e-12)."""
    from pde_solver_amd import capi, hip

    ok, note = hip.is_available()
    log(f"device: {note}")
    if not ok:
        raise SystemExit("no usable gfx9 device")
    t = time.perf_counter()
    specs = make_batch_problems(args.count)
    nodes = sum(s.n for s in specs)
    sizes = {}
    for s in specs:
        sizes[s.nx] = sizes.get(s.nx, 0) + 1
    log(f"{len(specs)} problems, {nodes / 1e6:.1f}M nodes, sizes {dict(sorted(sizes.items()))}, "
        f"generated in {time.perf_counter() - t:.1f}s")
    ctx = hip.Context()
    tuned = capi.BatchConfig(autotune=1, use_graph=1)

    t = time.perf_counter()
    ref = ctx.batch(specs, batch_options(capi.FP64, 1e-12), tuned)
    ref.solve()
    u_ref = [ref.download(i) for i in range(len(specs))]
    ref.close()
    log(f"FP64 oracle batch (rel-resid 1e-12) in {time.perf_counter() - t:.1f}s incl. set-up")

    def run(precision, tol, fp16, steps, roofline=False):
        return batch_run(ctx, specs, precision, tol, fp16, steps, args.warmup, u_ref, roofline)

    mixed = run(capi.MIXED, 1e-8, 0, args.steps, roofline=True)
    log(f"mixed batched (FP32 coarse): {mixed['solves_per_s']} solves/s, batch p50 "
        f"{mixed['batch_ms_p50']} ms, max err {mixed['max_err_vs_fp64_oracle']:.2e}")
    mixed16 = run(capi.MIXED, 1e-8, 1, args.steps)
    log(f"mixed batched (FP16 coarse): {mixed16['solves_per_s']} solves/s")
    fp64 = run(capi.FP64, 1e-8, 1, args.steps, roofline=True)
    log(f"fp64 batched: {fp64['solves_per_s']} solves/s, batch p50 {fp64['batch_ms_p50']} ms")
    tight_mixed = run(capi.MIXED, 1e-10, 0, max(1, args.steps // 2))
    tight_fp64 = run(capi.FP64, 1e-10, 1, max(1, args.steps // 2))
    log(f"at rel-resid 1e-10: mixed {tight_mixed['solves_per_s']} (err "
        f"{tight_mixed['max_err_vs_fp64_oracle']:.1e}), fp64 {tight_fp64['solves_per_s']} solves/s")

    plain = batch_plain(ctx, specs, u_ref)
    log(f"plain fp64 (sequential resident plans): {plain['solves_per_s']} solves/s")
    cpu = None if args.no_cpu_baseline else batch_cpu(specs, 15.0)

    sw = mixed["level0_kernels"]["sweep"]
    out = {
        "metric": METRIC,
        "value": mixed["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["batch_ms_p50"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 MG-PCG + f64 residual replacement",
        "data": "synthetic",
        "config": {"workload": f"batched uneven sizes: {len(specs)} 2D problems per step, n "
                               "log-uniform in {33..1025}, half Poisson / half Helmholtz "
                               "c~U(0,50), 8-mode Fourier RHS, D/N/R per face, MG-PCG V(2,2) "
                               "to rel-resid 1e-8",
                   "problems": len(specs), "nodes": nodes},
        "latency_ms": {"p50": mixed["batch_ms_p50"], "p99": mixed["batch_ms_p99"]},
        "roofline": {"bound": "hbm", "kernel": "kb_sweep<P32> (red+black, one launch) over every "
                                               "problem's level 0",
                     "achieved": sw["GB/s"], "peak": 8000.0, "unit": "GB/s",
                     "frac": round(sw["GB/s"] / 8000.0, 3), "traffic": None,
                     "algorithmic_bytes": sw["algorithmic_bytes"], "avg_ms": sw["avg_ms"]},
        "mixed_batched": mixed,
        "mixed_batched_fp16_coarse": mixed16,
        "fp64_batched": fp64,
        "plain_fp64": plain,
        "at_rel_resid_1e-10": {"mixed_batched": tight_mixed, "fp64_batched": tight_fp64},
        "speedup_mixed_batched_vs_plain_fp64": round(mixed["solves_per_s"] / plain["solves_per_s"], 2),
        "speedup_mixed_batched_vs_fp64_batched": round(mixed["solves_per_s"] / fp64["solves_per_s"], 2),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


def options(precision: int, tol: float):
    from pde_solver_amd import capi
    return capi.options(capi.MGPCG, tol=tol, tol_mode=capi.TOL_REL, precond=capi.PC_MG,
                        bc_rows=capi.ROWS_FOLDED, precision=precision, mg_max_levels=6,
                        mg_pre=2, mg_post=2, max_iter=200, inner_tol=1e-4)


def timed_solves(plan, steps: int, warmup: int):
    for _ in range(warmup):
        plan.solve()
    lat, dev, reps = [], [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        t = time.perf_counter()
        _, rep = plan.solve()
        lat.append((time.perf_counter() - t) * 1e3)
        dev.append(rep.solve_ms)
        reps.append(rep)
    wall = time.perf_counter() - t0
    return wall, lat, dev, reps


def pct(v, q):
    return float(np.percentile(np.asarray(v), q))


def cpu_baseline(spec, n_threads: int, runs: int = 5):
    """The oracle's FP64 MG-PCG on the same problem: median solve-phase time of `runs` solves
    (the oracle's own steady_clock around the solve; set-up reported separately)."""
    from oracle import oracle as O
    from pde_solver_amd import capi
    O.set_threads(n_threads)
    solve_s, total_s = [], []
    for _ in range(runs):
        t = time.perf_counter()
        st, u, rep = O.solve(spec, options(capi.FP64, 1e-8))
        total_s.append(time.perf_counter() - t)
        solve_s.append(rep.solve_ms / 1e3)
    return {"status": st, "solve_s": statistics.median(solve_s), "total_s": statistics.median(total_s),
            "iterations": rep.iterations, "runs": runs}, u


def measure_traffic(n: int):
    """HBM bytes per launch of the fine-level FP32 sweep from two rocprofv3 counter passes over
    this script's kernel-timing mode (FETCH_SIZE, then WRITE_SIZE), each scaled by what the
    same counter reads for the FP32 copy of one level-0 vector (N * 4 bytes read and written).
    Returns (bytes, detail) or (None, reason)."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile
    if not shutil.which("rocprofv3"):
        return None, "rocprofv3 not found"
    med = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        out = tempfile.mkdtemp(prefix="pmc_")
        cmd = ["timeout", "-s", "KILL", "150", "rocprofv3", "--pmc", counter, "--output-format",
               "csv", "-d", out, "-o", "pmc", "--", sys.executable, os.path.join(ROOT, "bench.py"),
               "--kernels", "4,3", "--precision", "mixed", "--steps", "4", "--n", str(n)]
        env = dict(os.environ, TMPDIR="/tmp")
        log(f"traffic: {counter} pass ...")
        try:
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=180, env=env, cwd="/tmp")
        except subprocess.TimeoutExpired:
            return None, f"{counter} pass timed out"
        files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
        if r.returncode not in (0, 139) or not files:
            return None, f"{counter} pass failed (rc {r.returncode})"
        vals = {}
        with open(files[0]) as fh:
            for row in csv.DictReader(fh):
                name = row.get("Kernel_Name", "")
                if row.get("Counter_Name") != counter:
                    continue
                sweep = "k_rbgs_reg3" in name or "k_rbgs_ring3" in name
                key = "sweep" if sweep else ("copy" if "k_convert" in name else None)
                if key:
                    vals.setdefault(key, []).append(float(row["Counter_Value"]))
        if "sweep" not in vals or "copy" not in vals:
            return None, f"{counter}: kernels not found in the counter output"
        med[counter] = {k: statistics.median(v) for k, v in vals.items()}
        shutil.rmtree(out, ignore_errors=True)
    nbytes = float(n) ** 3 * 4.0
    scale_f = nbytes / (med["FETCH_SIZE"]["copy"] * 1024.0)
    scale_w = nbytes / (med["WRITE_SIZE"]["copy"] * 1024.0)
    traffic = med["FETCH_SIZE"]["sweep"] * 1024.0 * scale_f + med["WRITE_SIZE"]["sweep"] * 1024.0 * scale_w
    detail = {"fetch_KiB": med["FETCH_SIZE"]["sweep"], "write_KiB": med["WRITE_SIZE"]["sweep"],
              "copy_fetch_KiB": med["FETCH_SIZE"]["copy"], "copy_write_KiB": med["WRITE_SIZE"]["copy"],
              "fetch_scale": round(scale_f, 4), "write_scale": round(scale_w, 4)}
    return traffic, detail


def bench_cfg1(ctx, with_cpu: bool):
    """The plain CPU workload on both sides: 129^2 Poisson, Dirichlet 0, f = 2 pi^2 sin sin;
    CG to relative residual 1e-10 and Jacobi to max|du| < 1e-10.  GPU: median of 5 (CG) or 3
    (Jacobi) one-shot solves, set-up included as a caller sees it.  CPU: the oracle around the
    whole solve call (the right-hand side build is negligible at this size), median of 5 CG
    solves, and a 5000-sweep Jacobi sample whose per-sweep time is scaled to the full count."""
    from pde_solver_amd import capi
    n = 129
    s = capi.Spec(n, n, a=-1.0, b=-1.0)
    X, Y, _ = s.coords()
    exact = np.sin(np.pi * X) * np.sin(np.pi * Y)
    s.f_field = -2 * np.pi ** 2 * exact
    h = 1.0 / (n - 1)
    closed = 2 * np.pi ** 2 / ((8 / h ** 2) * np.sin(np.pi * h / 2) ** 2) - 1.0
    cg = capi.options(capi.CG, tol=1e-10, tol_mode=capi.TOL_REL, max_iter=2000)
    jac = capi.options(capi.JACOBI, tol=1e-10, max_iter=60000)
    out = {"workload": "129^2 Poisson, Dirichlet 0, f = 2 pi^2 sin(pi x) sin(pi y); CG to "
                       "rel-resid 1e-10, Jacobi to max|du| < 1e-10",
           "mms_closed_form": closed}
    for name, opt in (("cg", cg), ("jacobi", jac)):
        ctx.solve(s, opt)
        ms = []
        for _ in range(5 if name == "cg" else 3):
            t = time.perf_counter()
            st, u, rep = ctx.solve(s, opt)
            ms.append((time.perf_counter() - t) * 1e3)
        out[f"gpu_{name}"] = {"ms_median": round(statistics.median(ms), 3), "iterations": rep.iterations,
                              "device_ms": round(rep.solve_ms, 3), "status": st,
                              "mms_linf": float(np.abs(u - exact).max())}
    log(f"cfg1: GPU CG {out['gpu_cg']['ms_median']} ms ({out['gpu_cg']['iterations']} its), Jacobi "
        f"{out['gpu_jacobi']['ms_median']} ms ({out['gpu_jacobi']['iterations']} sweeps)")
    out["_cpu"] = (s, cg, exact) if with_cpu else None
    return out


def bench_cfg1_cpu(out):
    """The CPU half of bench_cfg1 (run after every GPU measurement)."""
    from pde_solver_amd import capi
    s, cg, exact = out.pop("_cpu")
    from oracle import oracle as O
    O.set_threads(cpu_threads())
    ms = []
    for _ in range(5):
        t = time.perf_counter()
        st, u, rep = O.solve(s, cg)
        ms.append((time.perf_counter() - t) * 1e3)
    out["cpu_cg"] = {"ms_median": round(statistics.median(ms), 3), "iterations": rep.iterations,
                     "mms_linf": float(np.abs(u - exact).max()), "cores": cpu_threads(),
                     "build": CPU_BUILD["flags"]}
    sample = capi.options(capi.JACOBI, tol=0.0, max_iter=5000)
    ms = []
    for _ in range(3):
        t = time.perf_counter()
        _, _, rep = O.solve(s, sample)
        ms.append((time.perf_counter() - t) * 1e3)
    per = statistics.median(ms) / rep.iterations
    full = out["gpu_jacobi"]["iterations"]
    out["cpu_jacobi"] = {"ms_per_sweep": round(per, 5), "sweeps": full,
                         "ms_full_estimate": round(per * full, 1), "cores": cpu_threads(),
                         "sample": f"median of 3 runs of {rep.iterations} sweeps",
                         "build": CPU_BUILD["flags"]}
    log(f"cfg1: CPU CG {out['cpu_cg']['ms_median']} ms, Jacobi {per * full:.0f} ms (from a sample)")


def batched_summary(ctx, steps: int, warmup: int, with_cpu: bool):
    """The batched workload (512 uneven 2D problems) in the default line: mixed batched and
    FP64 batched lock-step solves, the plain FP64 path, the batched sweep's HIP-event roofline
    and the CPU oracle on a bounded sample.  Errors are against the FP64 batch at 1e-12."""
    from pde_solver_amd import capi
    specs = make_batch_problems(512)
    ref = ctx.batch(specs, batch_options(capi.FP64, 1e-12), capi.BatchConfig(autotune=1, use_graph=1))
    ref.solve()
    u_ref = [ref.download(i) for i in range(len(specs))]
    ref.close()
    mixed = batch_run(ctx, specs, capi.MIXED, 1e-8, 0, steps, warmup, u_ref, roofline=True)
    fp64 = batch_run(ctx, specs, capi.FP64, 1e-8, 1, steps, warmup, u_ref)
    plain = batch_plain(ctx, specs, u_ref)
    # rel-resid 1e-8 leaves errors up to ~2e-8 on the near-singular problems; at 2e-9 both
    # batched paths stay within 1e-8 of the FP64 solution
    acc = {"tol": 2e-9,
           "mixed_batched": batch_run(ctx, specs, capi.MIXED, 2e-9, 0, steps, warmup, u_ref),
           "fp64_batched": batch_run(ctx, specs, capi.FP64, 2e-9, 1, steps, warmup, u_ref)}
    acc["speedup_mixed_vs_fp64_batched"] = round(acc["mixed_batched"]["solves_per_s"] /
                                                 acc["fp64_batched"]["solves_per_s"], 2)
    acc["speedup_mixed_batched_vs_plain_fp64"] = round(acc["mixed_batched"]["solves_per_s"] /
                                                       plain["solves_per_s"], 2)
    sw = mixed.pop("level0_kernels")["sweep"]
    out = {"workload": "512 2D problems, n log-uniform in 33..1025, half Poisson / half Helmholtz "
                       "c~U(0,50), 8-mode Fourier RHS, D/N/R per face, MG-PCG V(2,2) to rel-resid 1e-8",
           "nodes": sum(s.n for s in specs),
           "mixed_batched": mixed, "fp64_batched": fp64, "plain_fp64": plain,
           "speedup_mixed_batched_vs_plain_fp64": round(mixed["solves_per_s"] / plain["solves_per_s"], 2),
           "speedup_mixed_batched_vs_fp64_batched": round(mixed["solves_per_s"] / fp64["solves_per_s"], 2),
           "roofline": {"bound": "hbm", "kernel": "kb_sweep<P32> over every problem's level 0",
                        "achieved": sw["GB/s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(sw["GB/s"] / HBM_PEAK_GBS, 4),
                        "algorithmic_bytes": sw["algorithmic_bytes"], "avg_ms": sw["avg_ms"]},
           "at_error_1e-8": acc,
           "cpu_baseline": None,
           "_cpu": specs if with_cpu else None}
    log(f"batched: mixed {mixed['solves_per_s']}, fp64 {fp64['solves_per_s']}, plain fp64 "
        f"{plain['solves_per_s']} solves/s")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=257)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only", choices=["mixed", "fp64"], default=None,
                    help="run W + K solves of one path and print their times (profiling runs)")
    ap.add_argument("--kernels", default=None,
                    help="comma list of plan kernels to launch K times each (counter runs)")
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--workload", choices=["single", "batch", "nonsym", "serve"], default="single",
                    help="single: the 257^3 solve (default); batch: 512 uneven 2D problems; "
                         "nonsym: convection-diffusion Krylov + reaction-diffusion RK4; "
                         "serve: Poisson-arrival request stream")
    ap.add_argument("--rates", default="50,200,500,1000,2000", help="serve: arrival rates, req/s")
    ap.add_argument("--count", type=int, default=None,
                    help="batch: problems per step (512); nonsym: problems timed (24 of 128)")
    ap.add_argument("--rd-steps", type=int, default=1000, help="nonsym: RK4 steps at 2049^2")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter passes that measure roofline.traffic")
    ap.add_argument("--no-extras", action="store_true",
                    help="default workload: only the 257^3 line (no 129^2 or batched rows)")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="launch-shape knob for A/B timing (pdehip_tune_set), repeatable")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        return 2

    if not args.no_cpu_baseline and not args.kernels and not args.only:
        cpu_setup()
    if args.workload == "batch":
        if args.count is None:
            args.count = 512
        return bench_batch(args)
    if args.workload == "nonsym":
        if args.count is None:
            args.count = 24
        return bench_nonsym(args)
    if args.workload == "serve":
        if args.count is None:
            args.count = 400
        return bench_serve(args)

    # counter passes first, in child processes, before this process touches the GPU
    traffic, traffic_detail = None, "not measured (--no-traffic)"
    if not args.kernels and not args.only and not args.no_traffic:
        traffic, traffic_detail = measure_traffic(args.n)
        log(f"traffic: {traffic} bytes per sweep ({traffic_detail})")

    from pde_solver_amd import capi, hip

    ok, note = hip.is_available()
    log(f"device: {note}")
    if not ok:
        raise SystemExit("no usable gfx9 device")
    for kv in args.tune:
        key, val = kv.split("=")
        hip.tune_set(key, int(val))
        log(f"tuning {key} = {val}")
    n = args.n
    t = time.perf_counter()
    spec = make_problem(n)
    log(f"inputs {n}^3 generated in {time.perf_counter() - t:.1f}s")
    ctx = hip.Context()

    if args.kernels:
        prec = capi.MIXED if args.precision == "mixed" else capi.FP64
        plan = ctx.plan(spec, options(prec, 1e-8))
        res = {}
        for k in [int(v) for v in args.kernels.split(",")]:
            ms, nbytes = plan.time_kernel(k, prec, reps=args.steps)
            res[str(k)] = {"avg_ms": round(ms, 5), "algorithmic_bytes": nbytes,
                           "GB/s": round(nbytes / (ms * 1e-3) / 1e9, 1)}
        print(json.dumps({"kernels": res, "precision": args.precision, "n": n,
                          "launches_each": args.steps + 1}), flush=True)
        plan.close()
        ctx.close()
        return 0

    if args.only:
        plan = ctx.plan(spec, options(capi.MIXED if args.only == "mixed" else capi.FP64, 1e-8))
        wall, lat, dev, reps = timed_solves(plan, args.steps, args.warmup)
        print(json.dumps({"path": args.only, "n": n, "steps": args.steps, "warmup": args.warmup,
                          "solves_per_s": round(args.steps / wall, 3),
                          "device_ms": [round(v, 3) for v in dev],
                          "iterations": reps[-1].iterations,
                          "inner_iterations": reps[-1].inner_iterations}), flush=True)
        plan.close()
        ctx.close()
        return 0

    # FP64 oracle solution on the GPU (plain path driven to rel-resid 1e-12)
    t = time.perf_counter()
    ref_plan = ctx.plan(spec, options(capi.FP64, 1e-12))
    _, rep_ref = ref_plan.solve()
    u_ref = ref_plan.download()
    ref_plan.close()
    log(f"FP64 reference solve (rel 1e-12): {rep_ref.iterations} iterations, "
        f"{time.perf_counter() - t:.1f}s incl. set-up")

    plan64 = ctx.plan(spec, options(capi.FP64, 1e-8))
    planmx = ctx.plan(spec, options(capi.MIXED, 1e-8))
    log(f"plans resident: fp64 {plan64.bytes() / 2**30:.2f} GiB, mixed {planmx.bytes() / 2**30:.2f} GiB")

    # ---- mixed (the measured path)
    wall_mx, lat_mx, dev_mx, reps_mx = timed_solves(planmx, args.steps, args.warmup)
    u_mx = planmx.download()
    log(f"mixed: {args.steps} solves in {wall_mx:.3f}s, refinements {reps_mx[-1].refinements}, "
        f"inner iterations {reps_mx[-1].inner_iterations}, rel_resid {reps_mx[-1].rel_residual:.2e}")

    # ---- plain FP64
    steps64 = max(3, args.steps // 2)
    wall_64, lat_64, dev_64, reps_64 = timed_solves(plan64, steps64, 1)
    u_64 = plan64.download()
    log(f"fp64: {steps64} solves in {wall_64:.3f}s, iterations {reps_64[-1].iterations}, "
        f"rel_resid {reps_64[-1].rel_residual:.2e}")

    nref = float(np.linalg.norm(u_ref))
    err_mx = float(np.linalg.norm(u_mx - u_ref) / nref)
    err_64 = float(np.linalg.norm(u_64 - u_ref) / nref)

    # ---- roofline of the dominant kernel (fine-level fused red-black GS sweep) with HIP events
    # on the plan's stream (the stream the solver launches it on)
    ms_mx, bytes_mx = planmx.time_kernel(4, capi.MIXED, reps=20)
    ms_64, bytes_64 = plan64.time_kernel(4, capi.FP64, reps=20)
    ms_ip, bytes_ip = planmx.time_kernel(1, capi.MIXED, reps=20)
    ms_ap, bytes_ap = plan64.time_kernel(0, capi.FP64, reps=20)
    ms_rs, bytes_rs = plan64.time_kernel(2, capi.FP64, reps=20)
    ms_cp, bytes_cp = plan64.time_kernel(3, capi.FP64, reps=20)
    gbs = lambda b, ms: b / (ms * 1e-3) / 1e9  # noqa: E731
    sweep_kernel = "k_rbgs_reg3" if hip.tune_get("smooth.reg") else "k_rbgs_ring3"
    roof = {"bound": "hbm", "kernel": f"{sweep_kernel} (red+black sweep, one launch), FP32, fine level {n}^3",
            "achieved": round(gbs(bytes_mx, ms_mx), 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs(bytes_mx, ms_mx) / HBM_PEAK_GBS, 4),
            "traffic": round(traffic) if traffic else None,
            "traffic_over_algorithmic": round(traffic / bytes_mx, 3) if traffic else None,
            "traffic_counters": traffic_detail,
            "algorithmic_bytes": bytes_mx, "avg_ms": round(ms_mx, 4)}
    kernels = {
        "rbgs_sweep_fp64": {"avg_ms": round(ms_64, 4), "GB/s": round(gbs(bytes_64, ms_64), 1)},
        "rbgs_inplace_2launch_fp32": {"avg_ms": round(ms_ip, 4), "GB/s": round(gbs(bytes_ip, ms_ip), 1)},
        "apply_fp64": {"avg_ms": round(ms_ap, 4), "GB/s": round(gbs(bytes_ap, ms_ap), 1)},
        "residual_fp64": {"avg_ms": round(ms_rs, 4), "GB/s": round(gbs(bytes_rs, ms_rs), 1)},
        "stream_copy_fp64": {"avg_ms": round(ms_cp, 4), "GB/s": round(gbs(bytes_cp, ms_cp), 1)},
    }
    log(f"roofline: rbgs fp32 {roof['achieved']} GB/s, rbgs fp64 {kernels['rbgs_sweep_fp64']['GB/s']} GB/s, "
        f"apply fp64 {kernels['apply_fp64']['GB/s']} GB/s, copy {kernels['stream_copy_fp64']['GB/s']} GB/s")

    # ---- a tolerance at which both paths' error against the FP64 solution is <= 1e-8 (rel-resid
    # 1e-8 leaves ~1.5e-8 on this problem)
    tol_acc = 2e-9
    acc = {"tol": tol_acc}
    for name, prec in (("mixed", capi.MIXED), ("fp64", capi.FP64)):
        pl = ctx.plan(spec, options(prec, tol_acc))
        w, la, _, rp = timed_solves(pl, 5, 1)
        acc[name] = {"solves_per_s": round(5 / w, 3), "p99_ms": round(pct(la, 99), 3),
                     "err_vs_fp64_oracle": float(np.linalg.norm(pl.download() - u_ref) / nref),
                     "rel_residual": rp[-1].rel_residual,
                     "iterations": rp[-1].inner_iterations if prec == capi.MIXED else rp[-1].iterations}
        pl.close()
    acc["speedup_mixed_vs_fp64"] = round(acc["mixed"]["solves_per_s"] / acc["fp64"]["solves_per_s"], 3)
    log(f"at tol {tol_acc}: mixed {acc['mixed']['solves_per_s']} (err {acc['mixed']['err_vs_fp64_oracle']:.1e}), "
        f"fp64 {acc['fp64']['solves_per_s']} (err {acc['fp64']['err_vs_fp64_oracle']:.1e})")

    plan64.close()
    planmx.close()
    extras = {}
    if not args.no_extras:
        extras["plain_cpu_workload_129"] = bench_cfg1(ctx, not args.no_cpu_baseline)
        extras["batched_512"] = batched_summary(ctx, 5, 2, not args.no_cpu_baseline)

    # ---- CPU baselines last (the OpenMP runtime binds this process's main thread): the oracle's
    # FP64 MG-PCG on the same problem, median of 5; then the 129^2 and batched rows' CPU legs
    cpu = None
    if not args.no_cpu_baseline:
        threads = cpu_threads()
        log(f"cpu baseline: oracle FP64 MG-PCG on {threads} threads, 5 solves ...")
        res, u_cpu = cpu_baseline(spec, threads)
        cpu = {"value": round(1.0 / res["solve_s"], 4), "unit": "solves/s", "cores": threads,
               "kind": "port", "build": CPU_BUILD["flags"],
               "omp": f"OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')} OMP_PLACES={os.environ.get('OMP_PLACES')}",
               "sample": f"median of {res['runs']} full {n}^3 FP64 MG-PCG solves to rel-resid 1e-8 "
                         f"({res['iterations']} iterations, solve phase {res['solve_s']:.2f}s, "
                         f"set-up excluded; {res['total_s']:.2f}s with set-up)",
               "rel_l2_vs_gpu_fp64": float(np.linalg.norm(u_cpu - u_64) / np.linalg.norm(u_64))}
        log(f"cpu baseline: {cpu['value']} solves/s")
        if extras:
            bench_cfg1_cpu(extras["plain_cpu_workload_129"])
            extras["batched_512"]["cpu_baseline"] = batch_cpu(extras["batched_512"].pop("_cpu"), 15.0)
    for e in extras.values():
        e.pop("_cpu", None)

    value = args.steps / wall_mx
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_mx / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 MG-PCG + f16 coarse storage inside f64 refinement",
        "data": "synthetic",
        "config": {"workload": f"3D var-coef diffusion -div(k grad u)=f, {n}^3, 7-pt flux stencil, "
                               "k=exp(0.5 g) (Gaussian field sigma=4 cells, seed 2), f~N(0,1) seed 20, "
                               "Dirichlet 0 x/y faces, Robin(1,1,0) z faces; MG-PCG V(2,2) RB-GS "
                               "6 levels, rel-resid 1e-8",
                   "n": n, "levels": 6, "tol": 1e-8},
        "latency_ms": {"p50": round(pct(lat_mx, 50), 3), "p99": round(pct(lat_mx, 99), 3),
                       "device_p50": round(pct(dev_mx, 50), 3)},
        "fp64": {"value": round(steps64 / wall_64, 3), "unit": "solves/s",
                 "p50_ms": round(pct(lat_64, 50), 3), "p99_ms": round(pct(lat_64, 99), 3),
                 "iterations": reps_64[-1].iterations, "rel_residual": reps_64[-1].rel_residual,
                 "err_vs_fp64_oracle": err_64},
        "mixed": {"refinements": reps_mx[-1].refinements, "inner_iterations": reps_mx[-1].inner_iterations,
                  "rel_residual": reps_mx[-1].rel_residual, "err_vs_fp64_oracle": err_mx},
        "speedup_mixed_vs_fp64": round(value / (steps64 / wall_64), 3),
        "at_error_1e-8": acc,
        "roofline": roof,
        "kernels": kernels,
        "cpu_baseline": cpu,
        **extras,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
