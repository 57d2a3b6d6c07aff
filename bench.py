"""Benchmark: pressure-Poisson multigrid solves per second on one MI355X.

Workload (256x192x192 stretched grid, the headline configuration):
  domain [0,4]x[0,3]x[0,3]; x uniform, y and z tanh-clustered at the centre
  (beta = 2); one prolate ellipsoid, semi-axes (0.5,0.25,0.25), centre
  (1.5,1.5,1.5), rotated 20 deg about z; RHS = discrete divergence
  (rhs_poisson, dt = 1) of seeded face velocities N(0,1) times
  sin(pi x/4) sin(pi y/3) sin(pi z/3).  One step = one solve from p = 0 to
  ||b - A p||_2 / ||b||_2 <= 1e-8 (after the compatibility projection).

value = FP32 V-cycles inside FP64 iterative refinement, solves/s, inputs
resident in HBM.  The same line also reports the FP64 multigrid's solve
rate, the refinement solve's error against the FP64 solve, the roofline
of the dominant kernel (HIP events around its launches) and the
reference multigrid restated in C (the oracle) timed on the host cores, whose
converged FP64 solution the GPU solutions are compared with
(error_vs_oracle).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from flow_structure_interaction_fem_ibm_amd import pp  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "MG Poisson solve/s @1e-8 resid (256x192x192): FP32+refine vs FP64, HBM GB/s"


def workload(nx=256, ny=192, nz=192, seed=2, extras=False):
    xn = np.linspace(0.0, 4.0, nx + 1)
    yn = pp.tanh_nodes(ny, 3.0, 2.0)
    zn = pp.tanh_nodes(nz, 3.0, 2.0)
    body = pp.Ellipsoid((1.5, 1.5, 1.5), (0.5, 0.25, 0.25), 20.0, 96)
    ib, fr, gs, ex = pp.build_geometry(xn, yn, zn, [body], mode=pp.PP_GEOM_ANALYTIC)
    rhs = divergence_rhs(xn, yn, zn, ib, seed)
    if extras:
        return xn, yn, zn, ib, gs, rhs, ex
    return xn, yn, zn, ib, gs, rhs


def time_frac_step(solver, ex, dev, reps=10):
    """correct_vel and update_pressure on device fields of the same grid
    (the steps either side of the solve; not part of the headline value):
    wall time per call around synchronised calls, and algorithmic GB/s
    (correct_vel 129 B/cell: p, u, v, w, three faces, three bc arrays and
    the flags read, u, v, w and the faces written; update_pressure 24 B/cell)."""
    import torch
    solver.set_vel_ghosts(ex["coeffD"], 2.0, 1.0)
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    f = [torch.randn(solver.shape, dtype=torch.float64, device=dev, generator=g) for _ in range(10)]
    cells = float(np.prod([n - 3 for n in solver.shape]))
    out = {}
    for name, call, bpc in (("correct_vel", lambda: solver.correct_vel_device(0.01, *f), 129.0),
                            ("update_pressure", lambda: solver.update_pressure_device(1, f[1], f[0]), 24.0)):
        call()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            call()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t) / reps * 1e3
        out[name] = {"ms": ms, "GBs": bpc * cells / (ms * 1e-3) / 1e9, "bytes_per_cell": bpc}
    return out


def time_solve_ad(solver, dev, reps=3):
    """The implicit momentum solve (pp_set_solve_ad_device +
    pp_solve_ad_device, Crank-Nicolson, GCM ghost velocities) on device
    fields of the same grid and geometry (not part of the headline value):
    nu = 1e-3, dt = 1e-3, face velocities and right-hand sides N(0,1), from
    u = v = w = 0 to max residual 1e-10.  Wall time per call around
    synchronised calls, iterations, and algorithmic GB/s at 564 B/cell per
    iteration: 6 line passes (x, y, z, two colours) of u, v and w together,
    76.5 B/cell each (half of the 7 FP64 row coefficients 28 and of the
    flags 0.5; per field half of r, var read and var written 12, and the
    other colour's var 4), and the three residuals, 105 B/cell."""
    import torch
    shape = solver.shape
    g = torch.Generator(device=dev)
    g.manual_seed(11)
    dt = 1e-3
    visc = torch.full(shape, 1e-3, dtype=torch.float64, device=dev)
    solver.set_solve_ad_device(dt, visc, visc, visc, visc)
    faces = [torch.randn(shape, dtype=torch.float64, device=dev, generator=g) for _ in range(3)]
    nl = [torch.randn(shape, dtype=torch.float64, device=dev, generator=g) for _ in range(3)]
    u = [torch.zeros(shape, dtype=torch.float64, device=dev) for _ in range(3)]
    st = solver.solve_ad_device(pp.PP_CRANK_NICOLSON1, *u, *nl, *faces, itermax_ad=100, restol_ad=1e-10)
    ms = []
    for _ in range(reps):
        for t in u:
            t.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = solver.solve_ad_device(pp.PP_CRANK_NICOLSON1, *u, *nl, *faces, itermax_ad=100, restol_ad=1e-10)
        torch.cuda.synchronize()
        ms.append((time.perf_counter() - t0) * 1e3)
    cells = float(np.prod([n - 3 for n in shape]))
    m = min(ms)
    its = max(st["iterations"], 1)
    return {"ms": m, "device_ms": st["ms"], "iterations": st["iterations"], "max_residual": st["max_residual"],
            "status": st["status"], "ms_per_iteration": m / its, "bytes_per_cell_per_iteration": 564.0,
            "GBs": 564.0 * cells * its / (m * 1e-3) / 1e9}


def divergence_rhs(xn, yn, zn, ib, seed):
    """rhs_poisson (CORE_POISSON.F90:281-288) with dt = 1 of face velocities
    N(0,1) * sin(pi x/4) sin(pi y/3) sin(pi z/3) sampled at face centres."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nx, ny, nz = len(xn) - 1, len(yn) - 1, len(zn) - 1
    xc, yc, zc = (0.5 * (a[1:] + a[:-1]) for a in (xn, yn, zn))
    env = lambda x, y, z: np.sin(np.pi * x / 4.0) * np.sin(np.pi * y / 3.0) * np.sin(np.pi * z / 3.0)
    # face_u at x nodes (nx+1), y/z centres, etc.
    u = rng.standard_normal((nz, ny, nx + 1)) * env(xn[None, None, :], yc[None, :, None], zc[:, None, None])
    v = rng.standard_normal((nz, ny + 1, nx)) * env(xc[None, None, :], yn[None, :, None], zc[:, None, None])
    w = rng.standard_normal((nz + 1, ny, nx)) * env(xc[None, None, :], yc[None, :, None], zn[:, None, None])
    div = ((u[:, :, 1:] - u[:, :, :-1]) / np.diff(xn)[None, None, :]
           + (v[:, 1:, :] - v[:, :-1, :]) / np.diff(yn)[None, :, None]
           + (w[1:, :, :] - w[:-1, :, :]) / np.diff(zn)[:, None, None])
    rhs = np.zeros((nz + 3, ny + 3, nx + 3))
    rhs[1:-2, 1:-2, 1:-2] = div
    return rhs * (1 - ib)


def progress(msg):
    """A progress line on stderr (the run's phases; the JSON line alone goes to stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


# Outer iterations the oracle's restated MG_Solver needs from p = 0 to
# relative residual 1e-8 on this workload, measured by `bench.py --cpu-full`
# (three full solves, 159 iterations each, 131-135 s on 16 threads; the log
# is profiles/cpu_baseline_256x192x192.md).
ORACLE_ITERATIONS_TO_1E8 = 159


def cpu_baseline(xn, yn, zn, ib, gs, rhs_projected, full=False, reps=3, sample_iterations=20, ref_tol=1e-13):
    """The oracle's restatement of the reference MG (CORE_MG.F90 MG_Solver +
    GCM_correct_pressure, CORE_POISSON.F90:638-676) on the same grid,
    geometry and projected RHS, on the host cores: the all-core variant
    (the same line rows in two-colour line order, OpenMP over the lines of a
    colour and over the cell loops).

    Default (a bounded sample, about 20 s): one warm-up outer iteration, then
    `sample_iterations` outer iterations from p = 0 timed together; value =
    1 / (seconds per iteration x ORACLE_ITERATIONS_TO_1E8).
    full = True: `reps` full solves to 1e-8 (best of), then one continued to
    ref_tol, whose zero-mean solution is returned for a direct comparison,
    and one outer iteration in the reference's own line order on one core."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib  # noqa: E402
    from oracle_lib import Oracle  # noqa: E402
    o = Oracle(xn, yn, zn)
    o.set_iblank(ib)
    o.set_stencils(gs)
    threads = oracle_lib.threads()
    p = np.zeros(ib.shape)
    t0 = time.perf_counter()
    o.mg_cycle(p, rhs_projected, zebra=1)           # warm-up (page faults, thread pool)
    progress(f"oracle warm-up iteration on {threads} threads: {time.perf_counter() - t0:.1f} s")
    common = {"cores": threads, "kind": "port", "unit": "solves/s", "cpu_model": cpu_model(), "compiler": "gcc",
              "flags": oracle_lib.build_flags()}
    if not full:
        p = np.zeros(ib.shape)
        t0 = time.perf_counter()
        rc, stats = o.mg_solve(p, rhs_projected, rel_tol=0.0, iter_max=sample_iterations, zebra=1)
        t = time.perf_counter() - t0
        per_it = t / stats["iterations"]
        progress(f"oracle sample: {stats['iterations']} iterations in {t:.1f} s ({per_it:.2f} s each)")
        out = dict(common, value=1.0 / (per_it * ORACLE_ITERATIONS_TO_1E8),
                   sample=(f"{stats['iterations']} outer iterations of the restated MG_Solver (x, y, z "
                           f"semi-coarsened V-cycles, line SOR, ghost rows) from p = 0 on the 256x192x192 problem, "
                           f"two-colour line order on {threads} OpenMP threads, after one warm-up iteration, "
                           f"scaled by the {ORACLE_ITERATIONS_TO_1E8} iterations a full solve to 1e-8 takes "
                           f"(measured by bench.py --cpu-full)"),
                   s_per_iteration=per_it, sample_s=t, iterations_to_1e8=ORACLE_ITERATIONS_TO_1E8,
                   rel_l2_after_sample=stats["l2"] / stats["b_l2"])
        return out, None
    times, stats = [], None
    for r in range(reps):
        p = np.zeros(ib.shape)
        t0 = time.perf_counter()
        rc, stats = o.mg_solve(p, rhs_projected, rel_tol=1e-8, iter_max=400, zebra=1)
        times.append(time.perf_counter() - t0)
        progress(f"oracle solve {r + 1}/{reps} to 1e-8: {stats['iterations']} iterations, {times[-1]:.1f} s")
        if rc != 0:
            raise RuntimeError(f"oracle MG did not reach 1e-8: {stats}")
    it_1e8 = stats["iterations"]
    t0 = time.perf_counter()
    rc, st_ref = o.mg_solve(p, rhs_projected, rel_tol=ref_tol, iter_max=600, zebra=1, mean_remove=0)
    t_ref = time.perf_counter() - t0
    progress(f"oracle continued to {ref_tol:g}: {st_ref['iterations']} more iterations, {t_ref:.1f} s")
    o.mean_remove(p)
    q = np.zeros(ib.shape)
    t0 = time.perf_counter()
    o.mg_cycle(q, rhs_projected, zebra=0)
    t_1core = time.perf_counter() - t0
    progress(f"one reference-order iteration on one core: {t_1core:.1f} s")
    best = min(times)
    out = dict(common, value=1.0 / best,
               sample=(f"full solves from p = 0 to relative residual 1e-8 ({it_1e8} outer iterations) of the "
                       f"256x192x192 problem, two-colour line order on {threads} OpenMP threads; best of {reps} "
                       f"after one warm-up iteration"),
               solve_s=times, iterations_to_1e8=it_1e8, reference_order_1core_s_per_iteration=t_1core,
               oracle_solution={"rel_l2": st_ref["l2"] / st_ref["b_l2"], "extra_iterations": st_ref["iterations"],
                                "s": t_ref, "stop": f"relative residual {ref_tol:g}"})
    return out, p


def certified_reference(o_factory, solver, d_p, d_rhs, common, bp, fl, tol=1e-13):
    """The FP64 device solve to `tol`, with its residual recomputed by the
    oracle's calc_residual (CORE_POISSON.F90:1111-1222) on the host: the
    solution of the reference's discrete system the GPU results are measured
    against (its own error is bounded by the condition number times tol)."""
    d_p.zero_()
    st = solver.solve_device(d_p, d_rhs, precision=pp.PP_PREC_FP64, **dict(common, rel_tol=tol, max_outer=400))
    p = d_p.cpu().numpy()
    o = o_factory()
    _, linf, _, l2 = o.calc_residual(p, bp)
    return p, {"solver": "FP64 multigrid + GMRES on the device", "tol": tol, "status": st["status"],
               "outer_iterations": st["outer_iterations"],
               "oracle_rel_l2": float(l2 / np.linalg.norm(bp[fl])), "oracle_linf": float(linf)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-full", action="store_true",
                    help="CPU baseline from full oracle solves (minutes) instead of a bounded sample, and the "
                         "GPU solutions against the oracle multigrid's own converged solution")
    ap.add_argument("--cpu-reps", type=int, default=3, help="timed oracle solves of --cpu-full")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--nu", type=int, default=2)
    ap.add_argument("--cycles", type=int, default=1)
    ap.add_argument("--graph", action="store_true", help="replay the V-cycles as a captured hipGraph")
    ap.add_argument("--nu-coarse", type=int, default=1, help="smoothing steps on the coarse levels (0: --nu)")
    ap.add_argument("--krylov", type=int, default=8,
                    help="flexible GMRES restart length around the V-cycles (0: plain refinement)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported (the solve is a single-GPU "
              f"kernel path; multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        sys.exit(2)
    t_setup = time.perf_counter()
    xn, yn, zn, ib, gs, rhs, gex = workload(extras=True)
    solver = pp.PressureSolver(xn, yn, zn)
    solver.set_geometry(ib, gs)
    t_setup = time.perf_counter() - t_setup
    dev = torch.device("cuda:0")
    d_rhs = torch.from_numpy(rhs).to(dev)
    d_p = torch.zeros(solver.shape, dtype=torch.float64, device=dev)
    common = dict(rel_tol=args.tol, max_outer=200, nu_pre=args.nu, nu_post=args.nu,
                  cycles_per_outer=args.cycles, krylov=args.krylov, use_graph=1 if args.graph else 0, nu_coarse=args.nu_coarse)

    def run(prec, steps, warm, profile=False):
        for _ in range(warm):
            d_p.zero_()
            solver.solve_device(d_p, d_rhs, precision=prec, **common)
        torch.cuda.synchronize()
        stats = []
        t0 = time.perf_counter()
        for _ in range(steps):
            d_p.zero_()
            stats.append(solver.solve_device(d_p, d_rhs, precision=prec, profile=int(profile), **common))
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps, stats

    progress(f"set-up {t_setup:.1f} s; timing {args.steps} mixed solves")
    ms_mixed, st_mixed = run(pp.PP_PREC_MIXED, args.steps, args.warmup)
    p_mixed = d_p.cpu().numpy()
    ms_fp64, st_fp64 = run(pp.PP_PREC_FP64, max(2, args.steps // 2), 1)
    p_fp64 = d_p.cpu().numpy()
    # one profiled solve of each precision: per-kernel-class HIP events
    _, prof_mixed = run(pp.PP_PREC_MIXED, 1, 0, profile=True)
    _, prof_fp64 = run(pp.PP_PREC_FP64, 1, 0, profile=True)
    for s in st_mixed + st_fp64:
        if s["status"] != "PP_OK":
            print(json.dumps({"error": "solve did not converge", "stats": s}), file=sys.stderr)
            sys.exit(1)
    fl = np.zeros(ib.shape, bool)
    fl[1:-2, 1:-2, 1:-2] = (ib[1:-2, 1:-2, 1:-2] == 0)

    def rel(a, b):
        return float(np.linalg.norm((a - b)[fl]) / np.linalg.norm(b[fl]))

    err = rel(p_mixed, p_fp64)
    # converged references: FP64 and mixed solves to 1e-12 (the precision
    # effect alone is their difference)
    tight = dict(common, rel_tol=1e-12)
    d_p.zero_()
    st_ref = solver.solve_device(d_p, d_rhs, precision=pp.PP_PREC_FP64, **tight)
    p_ref = d_p.cpu().numpy()
    d_p.zero_()
    st_mt = solver.solve_device(d_p, d_rhs, precision=pp.PP_PREC_MIXED, **tight)
    p_mt = d_p.cpu().numpy()
    progress(f"mixed {1e3 * ms_mixed:.2f} ms, FP64 {1e3 * ms_fp64:.2f} ms per solve; kernel timing, steps either side")
    km = prof_mixed[0]["kernels"]
    # dominant kernel: the kernel class with the largest device time on the
    # finest level (one kernel instantiation per class there); per-launch
    # algorithmic bytes / average launch time from HIP events on the
    # library's stream
    dom = max(km, key=lambda k: km[k]["ms_l0"])
    d = km[dom]
    n0 = max(d["launches_l0"], 1)
    avg_s = d["ms_l0"] * 1e-3 / n0
    bpl = d["bytes_l0"] / n0
    timing = "HIP events around each launch of one profiled solve"
    if dom in ("xline", "yline", "zline", "residual"):
        # 50 back-to-back launches of that kernel between two HIP events on
        # the library's stream (per-launch events add dispatch latency)
        reps = 50
        us, bpl = solver.time_kernel(dom, 0, pp.PP_PREC_MIXED, reps)
        avg_s, n0 = us * 1e-6, reps
        timing = f"HIP events around {reps} back-to-back launches"
    achieved = bpl / avg_s / 1e9 if avg_s > 0 else 0.0
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        t = json.load(open(tpath))
        if t.get("workload") == "256x192x192" and dom in t.get("kernels", {}):
            traffic = t["kernels"][dom]["hbm_bytes_per_launch"]
    cells = int(np.prod(ib[1:-2, 1:-2, 1:-2].shape))
    s0 = st_mixed[-1]
    out = {
        "metric": METRIC,
        "value": 1.0 / ms_mixed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_mixed * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 V-cycles + f64 refinement",
        "data": "synthetic (seeded divergence RHS, analytic ellipsoid geometry)",
        "config": {"workload": "256x192x192 tanh-stretched (beta=2 in y,z), rotated prolate ellipsoid, "
                               "divergence RHS, solve to 1e-8 relative residual",
                   "cells": cells, "ghosts": gs.n, "levels": solver.level_info(),
                   "nu_pre": args.nu, "nu_post": args.nu, "cycles_per_outer": args.cycles,
                   "krylov_restart": args.krylov, "hip_graph": bool(args.graph), "nu_coarse": args.nu_coarse},
        "fp64": {"solves_per_s": 1.0 / ms_fp64, "ms_per_solve": ms_fp64 * 1e3,
                 "outer_iterations": st_fp64[-1]["outer_iterations"], "rel_l2": st_fp64[-1]["rel_l2"]},
        "speedup_vs_fp64": ms_fp64 / ms_mixed,
        "mixed": {"outer_iterations": s0["outer_iterations"], "vcycles": s0["vcycles"], "rel_l2": s0["rel_l2"],
                  "abs_linf": s0["abs_linf"], "device_ms": s0["ms"], "algorithmic_GB": s0["bytes"] / 1e9,
                  "solve_GBs": s0["bytes"] / 1e9 / (s0["ms"] * 1e-3),
                  "history": s0["history"]},
        "error_vs_fp64_solve": err,
        "error_vs_fp64_converged": rel(p_mixed, p_ref),
        "error_fp64_solve_vs_fp64_converged": rel(p_fp64, p_ref),
        "mixed_vs_fp64_both_at_1e-12": rel(p_mt, p_ref),
        "converged_refs": {"fp64_rel_l2": st_ref["rel_l2"], "fp64_outer": st_ref["outer_iterations"],
                           "mixed_rel_l2": st_mt["rel_l2"], "mixed_outer": st_mt["outer_iterations"]},
        "roofline": {"bound": "hbm", "kernel": f"{dom} line pass, finest level" if dom.endswith("line") else dom,
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": ("profiles/traffic.json: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of this "
                                        "kernel on this workload in a separate run, scaled by a copy kernel's "
                                        "counters (not measured in this run)") if traffic else None,
                     "launches": n0, "avg_us": avg_s * 1e6, "bytes_per_launch": bpl, "timing": timing},
        "kernels_mixed": {k: {"ms": v["ms"], "GBs": (v["bytes"] / (v["ms"] * 1e-3) / 1e9) if v["ms"] > 0 else None,
                              "launches": v["launches"], "ms_l0": v["ms_l0"],
                              "GBs_l0": (v["bytes_l0"] / (v["ms_l0"] * 1e-3) / 1e9) if v["ms_l0"] > 0 else None}
                          for k, v in km.items()},
        "kernels_fp64": {k: {"ms": v["ms"], "launches": v["launches"]} for k, v in prof_fp64[0]["kernels"].items()},
        "setup_s": t_setup,
        "frac_step": time_frac_step(solver, gex, dev),
        "momentum": time_solve_ad(solver, dev),
        "cpu_baseline": None,
    }
    # the same projected RHS the GPU solved
    w = solver.null_vector()
    bp = rhs.copy()
    bp[fl] -= (w[fl] * rhs[fl]).sum() / w[fl].sum()

    def oracle():
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_lib import Oracle  # noqa: E402
        o = Oracle(xn, yn, zn)
        o.set_iblank(ib)
        o.set_stencils(gs)
        return o

    progress("oracle-certified FP64 reference and the tolerance sweep")
    p_cert, cert = certified_reference(oracle, solver, d_p, d_rhs, common, bp, fl)
    sweep = {}
    for tol in (1e-8, 1e-9, 1e-10, 1e-11, 1e-12):
        d_p.zero_()
        st = solver.solve_device(d_p, d_rhs, precision=pp.PP_PREC_MIXED, **dict(common, rel_tol=tol))
        sweep[f"{tol:g}"] = {"mixed_error": rel(d_p.cpu().numpy(), p_cert), "outer": st["outer_iterations"]}
    meets = [t for t, v in sweep.items() if v["mixed_error"] <= 1e-9]
    out["error_vs_oracle"] = {
        "reference": cert,
        "mixed_at_tol": rel(p_mixed, p_cert), "fp64_at_tol": rel(p_fp64, p_cert), "tol": args.tol,
        "mixed_by_tol": sweep,
        "loosest_tol_meeting_1e-9": max(meets, key=float) if meets else None}
    if not args.no_cpu:
        progress("CPU baseline (oracle)")
        out["cpu_baseline"], p_or = cpu_baseline(xn, yn, zn, ib, gs, bp, full=args.cpu_full, reps=args.cpu_reps)
        if p_or is not None:
            # the oracle multigrid's own converged solution (both zero fluid-volume mean)
            out["error_vs_oracle"]["vs_oracle_mg"] = {
                "mixed_at_tol": rel(p_mixed, p_or), "fp64_at_tol": rel(p_fp64, p_or),
                "certified_reference": rel(p_cert, p_or), "oracle": out["cpu_baseline"]["oracle_solution"]}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
