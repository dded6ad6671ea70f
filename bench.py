"""Benchmark: projections/s of the pressure projection at 2048^2, density ratio 1000:1.

One step = one full projection (PROJECTION_2D::Solve) of the Rayleigh-Taylor
workload (phi = y - 0.5 - 0.01 cos(8 pi x), rho 1000 above / 1 below, sigma 0,
u,v ~ U[-1,1] seed 2, dt = 1) from device-resident phi, u, v to device-resident
projected u, v and p: coefficient assembly, right-hand side, multigrid setup,
the solve to max|D u_proj| <= 1e-10 (and ||r||_2/||b||_2 <= 1e-10), the face
update and the FP64 divergence check.  Each step starts from the same
unprojected velocity (a device copy outside the timed interval) with a zero
initial pressure.

value = projections/s of the mixed-precision path (FP32 MG-PCG inside FP64
refinement); the FP64 path is timed the same way and reported beside it, as
are the adjoint's cost, per-kernel HBM rates at the finest level and, from the
CPU leg, the error of the mixed result against the FP64 CPU oracle.
"""
import argparse
import importlib
import json
import os
import re
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "projections/s at 2048² ρ 1000:1, div≤1e-10: mixed vs FP64; HBM GB/s"
HBM_PEAK_GBS = 8000.0

# Finest-level kernels of one FP32 inner iteration (and the FP64 refinement
# step), with their algorithmic HBM bytes per cell: every array the kernel
# must read or write once (halo re-reads and cache effects excluded).
# `which` is the p2d_time_kernel index: 50 back-to-back launches of the kernel
# on the solver's buffers (the workload's coefficients, patterned vectors),
# bracketed by one pair of HIP events on the context's stream, after the timed
# region.  (An event pair around every launch inside
# the solves perturbs them: the events' cache maintenance slows the timed kernel
# by ~40% and the projection by ~7%.)
KERNELS = [
    (2, "k_mg_pre<float,true>", 41,
     "Krylov update r -= a q, x += a p, 3 red-black sweeps from 0, residual, restriction: "
     "reads r q beta_x beta_y x p (24 B) and the packed binary16 weights wx|wy (4 B), writes r x x_pre "
     "(12 B) and coarse r (1 B)"),
    (3, "k_mg_post<float,true>", 25,
     "prolongation, 3 black-red sweeps, r.z: reads r x_pre beta_x beta_y (16 B), the packed weights "
     "(4 B) and coarse x (1 B), writes z (4 B)"),
    (0, "k_dir<float>", 24,
     "p = z - mean z + beta p_old, q = A p, p.q: reads z p_old beta_x beta_y (16 B), writes p q (8 B)"),
    (4, "k_correct_resid<float,float>", 56,
     "FP64 refinement: reads b p_hi (16 B) the FP32 p_lo (4 B) x (4 B) beta_x beta_y (16 B), writes p_hi (8 B) "
     "p_lo (4 B) r (4 B)"),
]
DOMINANT = "k_mg_pre<float,true>"

# Algorithmic bytes per cell of one whole mixed projection, by the same rule as KERNELS:
# a V-cycle is the finest pre + post + dir (41 + 25 + 24) plus, on each coarser level l
# (N / 4^l cells), pre (reads r beta_x beta_y and the packed weights, writes x and coarse r:
# 21) and post (reads r x_pre beta_x beta_y, the packed weights and coarse x, writes x: 25),
# summing to 46 N / 3; a refinement pass is 56; the once-per-projection kernels are the
# coefficients (reads phi, writes beta_x beta_y jc in FP64: 32), the right-hand side (reads
# u v beta_x beta_y, writes b and the FP32 residual: 44), the multigrid setup (FP64 beta in,
# FP32 beta and packed weights out, coarse beta, over the hierarchy: 44), the face update
# (68, the pressure's low part FP32) and the divergence check (16).
# An iteration is one full V-cycle with its direction; each pass ends with one more finest
# pre-smoothing (its update and convergence test, 41), after which the rest of that
# V-cycle is not launched.  An FP32 pass that converges at its direction finish skips
# that head (its near-end direction kernels read the residual, 4 B/cell more, and the
# refinement the pending direction, 4 more); the count keeps the head: an upper bound.
VCYCLE_B_PER_CELL = 41 + 25 + 24 + 46.0 / 3.0
HEAD_B_PER_CELL = 41
REFINE_B_PER_CELL = 56
SETUP_B_PER_CELL = 32 + 44 + 44 + 68 + 16


# The FP64 path moves every Krylov vector, coefficient and weight (wx and wy apart) in FP64:
# finest pre 90 (r q beta_x beta_y wx wy x p read, r x x_pre written, coarse r), post 58,
# dir 48, coarser levels 2 x 54 N / 3; its refinement reads b p_hi x beta_x beta_y and writes
# p_hi r in FP64, and reads and writes the FP32 p_lo (64).
FP64_VCYCLE_B_PER_CELL = 90 + 58 + 48 + 108.0 / 3.0
FP64_HEAD_B_PER_CELL = 90
FP64_REFINE_B_PER_CELL = 64


def projection_bytes(N, stats, fp64=False):
    """Algorithmic bytes of one projection: a full V-cycle per iteration, a pass-ending
    finest pre-smoothing and a refinement per pass, and the setup."""
    if fp64:
        return N * (stats["iterations"] * FP64_VCYCLE_B_PER_CELL
                    + stats["outer"] * (FP64_HEAD_B_PER_CELL + FP64_REFINE_B_PER_CELL) + SETUP_B_PER_CELL)
    return N * (stats["iterations"] * VCYCLE_B_PER_CELL + stats["outer"] * (HEAD_B_PER_CELL + REFINE_B_PER_CELL)
                + SETUP_B_PER_CELL)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_mg_pre_rt{n}_mixed_pmc.md")  # per workload size


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_path(P2D, torch, w, precision, steps, warmup, dev):
    n = w.n
    ctx = P2D.Context(n, n, w.h, precision=precision, rtol=1e-10, div_tol=1e-10, max_iterations=2000)
    stream = torch.cuda.current_stream(dev)
    ctx.set_stream(stream.cuda_stream)
    phi = torch.from_numpy(w.phi).to(dev)
    u0 = torch.from_numpy(w.u).to(dev)
    v0 = torch.from_numpy(w.v).to(dev)
    u, v = u0.clone(), v0.clone()
    ctx.set_fluid(phi, w.rho_minus, w.rho_plus, w.sigma)
    times, stats = [], None
    for k in range(warmup + steps):
        u.copy_(u0)
        v.copy_(v0)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        _, _, st = ctx.project(u, v, w.dt)
        e1.record(stream)
        e1.synchronize()
        if k >= warmup:
            times.append(e0.elapsed_time(e1))
        stats = st
        log(f"  {'mixed' if precision else 'fp64 '} step {k}: {e0.elapsed_time(e1):.3f} ms iters={st['iterations']} "
            f"outer={st['outer']} max_div={st['max_div']:.2e} rel_res={st['rel_residual']:.2e}")
    return ctx, times, stats, (u, v)


def time_adjoint(ctx, torch, w, uv, steps, dev):
    """The adjoint of the last projection (loss 0.5 ||u_proj - u_t||^2, u_t seed 3), timed like the forward."""
    W = importlib.import_module("2d_engine_amd.workloads")
    ut, vt = W.target_velocity(w.n)
    gu = (uv[0] - torch.from_numpy(ut).to(dev)).contiguous()
    gv = (uv[1] - torch.from_numpy(vt).to(dev)).contiguous()
    stream = torch.cuda.current_stream(dev)
    times, st = [], None
    for k in range(steps + 1):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        *_, st = ctx.adjoint(gu, gv)
        e1.record(stream)
        e1.synchronize()
        if k:
            times.append(e0.elapsed_time(e1))
    return statistics.median(times), st


def time_reinit(P2D, torch, ctx, w, dev, reps=5):
    """The kernels around the projection on the same workload: 25 Sussman WENO5
    re-initialisation steps (dt = h/5) of the level set scaled by 2 (not a distance), timed
    like a projection, and the step kernel alone (p2d_time_kernel); one WENO5 level-set
    advection stage, one ENO3 velocity advection stage and one explicit viscosity step."""
    src = torch.from_numpy(2.0 * w.phi).to(dev)
    phi = src.clone()
    stream = torch.cuda.current_stream(dev)
    times = []
    for k in range(reps + 1):
        phi.copy_(src)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        ctx.reinitialize(phi, iterations=25, scaling=5.0, epsilon=1e-6)
        e1.record(stream)
        e1.synchronize()
        if k:
            times.append(e0.elapsed_time(e1))
    step_ms = ctx.time_kernel(P2D.TIME_REINIT_STEP, 50)
    N = w.n * w.n
    # one WENO5 advection stage by the workload's velocity (dt = 0.4 h), device-resident
    u = torch.from_numpy(w.u).to(dev)
    v = torch.from_numpy(w.v).to(dev)
    adv = []
    for k in range(reps + 1):
        phi.copy_(src)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        ctx.advect_levelset(phi, u, v, 0.4 * w.h)
        e1.record(stream)
        e1.synchronize()
        if k:
            adv.append(e0.elapsed_time(e1))
    uu, vv = u.clone(), v.clone()
    vel = []
    for k in range(reps + 1):
        uu.copy_(u)
        vv.copy_(v)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        ctx.advect_velocity(uu, vv, 0.4 * w.h)
        e1.record(stream)
        e1.synchronize()
        if k:
            vel.append(e0.elapsed_time(e1))
    visc = []
    for k in range(reps + 1):
        uu.copy_(u)
        vv.copy_(v)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        ctx.apply_viscosity(uu, vv, 0.25 * w.h, 1.78e-5, 1.137e-3)
        e1.record(stream)
        e1.synchronize()
        if k:
            visc.append(e0.elapsed_time(e1))
    return {"reinit_ms_25_steps": round(statistics.median(times), 4),
            "reinit_step_avg_launch_ms": round(step_ms, 5), "reinit_step_bytes_per_cell": 24,
            "reinit_step_achieved_GBs": round(24 * N / (step_ms * 1e-3) / 1e9, 1),
            "advect_stage_ms": round(statistics.median(adv), 4),
            "velocity_advect_stage_ms": round(statistics.median(vel), 4),
            "viscosity_step_ms": round(statistics.median(visc), 4),
            "accounting": "re-initialisation step: reads phi and the sign function, writes phi (24 B/cell, FP64), "
                          "~600 FP64 VALU instructions per cell in the compiled kernel (four WENO5 derivatives with 16 correctly rounded quotients, Godunov), VALU-bound; advection stage: "
                          "reads phi and the two face velocities, writes phi (32 B/cell), two WENO5 derivatives "
                          "per cell, timed with its two device copies; velocity stage (ENO3, both components): reads u, v, writes u, v, timed with its four device copies; explicit viscosity step (air/water 1.78e-5/1.137e-3): reads phi, u, v, writes u, v, timed with its four device copies"}


def time_step(P2D, torch, w, dev, reps=3):
    """One whole TVD-RK3 air-water time step (p2d_advance, mixed-precision projections) on
    the same workload: three WENO5 level-set stages, three re-initialisations of 25 steps,
    three velocity stages each with ENO3 advection, gravity, air/water viscosity, wall
    velocities and a projection; dt = 0.25 h.  Device-resident, timed on the context's
    stream with events; each repetition restarts from the same state."""
    n = w.n
    ctx = P2D.Context(n, n, w.h, precision=P2D.MIXED, rtol=1e-10, div_tol=1e-10, max_iterations=2000)
    stream = torch.cuda.current_stream(dev)
    ctx.set_stream(stream.cuda_stream)
    src = [torch.from_numpy(a).to(dev) for a in (w.phi, w.u, w.v)]
    cur = [a.clone() for a in src]
    dt = 0.25 * w.h
    kw = dict(rho_minus=1.0, rho_plus=1000.0, sigma=0.0728, mu_minus=1.78e-5, mu_plus=1.137e-3)
    times, st = [], None
    for k in range(reps + 1):
        for a, b in zip(cur, src):
            a.copy_(b)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        _, _, _, st = ctx.advance(cur[0], cur[1], cur[2], dt, allow_not_converged=True, **kw)
        e1.record(stream)
        e1.synchronize()
        if k:
            times.append(e0.elapsed_time(e1))
    ms = statistics.median(times)
    return {"advance_ms": round(ms, 3), "steps_per_s": round(1000.0 / ms, 2), "projections_per_step": 3,
            "last_projection": {"iterations": st["iterations"], "outer": st["outer"], "max_div": st["max_div"]},
            "setup": f"rt {n}x{n}, u,v~U[-1,1] seed 2, dt 0.25 h, rho 1/1000, sigma 0.0728, mu 1.78e-5/1.137e-3, "
                     "25 re-initialisation steps per call, mixed-precision projections to max|D u| <= 1e-10"}


def time_semi_implicit(P2D, torch, n, dev, reps=3):
    """The oil-water semi-implicit viscosity step (p2d_visc_oil_water: coefficients, then a
    BiCGStab solve per MAC component) on an n x n vertical-pipe grid at the reference script's
    properties (r in [0.02, 1.28], z in [0, 3.14], Re 0.94983, oil/water viscosity 601:1,
    dt 1e-4, tolerance 1e-6 absolute), device-resident, timed with events."""
    import numpy as np
    dx, dy = 1.26 / n, 3.14 / n
    x0 = 0.02 + 0.5 * dx
    xs = x0 + np.arange(n) * dx
    ys = (np.arange(n) + 0.5) * dy
    phi = np.ascontiguousarray(xs[None, :] - (0.7 + 0.05 * np.sin(2 * np.pi * ys[:, None] / 3.14)))
    rng = np.random.default_rng(5)
    u, v = rng.uniform(-1, 1, (n, n + 1)), rng.uniform(-1, 1, (n + 1, n))
    vs = P2D.ViscositySolver(n, n, dx, dy)
    stream = torch.cuda.current_stream(dev)
    vs.set_stream(stream.cuda_stream)
    tp = torch.from_numpy(phi).to(dev)
    src = [torch.from_numpy(a).to(dev) for a in (u, v)]
    cur = [a.clone() for a in src]
    kw = dict(x_min=x0, Re=0.94983, rho_oil=0.905, rho_water=0.995, mu_oil=6.01, mu_water=0.01)
    times, sx, sy = [], None, None
    for k in range(reps + 1):
        for a, b in zip(cur, src):
            a.copy_(b)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        _, _, sx, sy = vs.oil_water(tp, cur[0], cur[1], 1e-4, allow_not_converged=True, **kw)
        e1.record(stream)
        e1.synchronize()
        if k:
            times.append(e0.elapsed_time(e1))
    ms = statistics.median(times)
    its = sx["iterations"] + sy["iterations"]
    return {"step_ms": round(ms, 3), "iterations_x": sx["iterations"], "iterations_y": sy["iterations"],
            "residual_x": sx["max_abs_residual"], "residual_y": sy["max_abs_residual"],
            "ms_per_iteration": round(ms / max(its, 1), 4),
            "setup": f"vertical pipe {n}x{n}, oil-water semi-implicit, dimensionless, reference boundary coupling"}


def _timed(torch, stream, fn, reps, warmup=1):
    """Median ms of fn() over reps runs after warmup runs, HIP events on the stream."""
    times, out = [], None
    for k in range(warmup + reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        out = fn()
        e1.record(stream)
        e1.synchronize()
        if k >= warmup:
            times.append(e0.elapsed_time(e1))
    return statistics.median(times), out


def time_configs(P2D, torch, W, dev, steps=5):
    """The other configurations, each on its own workload, device-resident, timed like the
    headline: C1 drops64 1024^2 FP64 MG-PCG; C3 its forward + adjoint; C4 drops512 4096^2
    mixed forward at its CFL step then the adjoint (only the forward pressure kept), with the
    device memory the run holds (hipMemGetInfo before the context and after each phase)."""
    stream = torch.cuda.current_stream(dev)
    out = {}

    def run(w, precision, adjoint, **cfg):
        T = lambda a: torch.from_numpy(a).to(dev)
        torch.cuda.synchronize()
        free = [torch.cuda.mem_get_info(dev)[0]]
        ctx = P2D.Context(w.n, w.n, w.h, precision=precision, rtol=1e-10, div_tol=1e-10, max_iterations=2000,
                          **cfg)
        ctx.set_stream(stream.cuda_stream)
        free.append(torch.cuda.mem_get_info(dev)[0])
        phi, u0, v0 = T(w.phi), T(w.u), T(w.v)
        u, v = u0.clone(), v0.clone()
        ctx.set_fluid(phi, w.rho_minus, w.rho_plus, w.sigma)

        def fwd():
            u.copy_(u0)
            v.copy_(v0)
            return ctx.project(u, v, w.dt)[2]
        ms, st = _timed(torch, stream, fwd, steps, warmup=2)
        free.append(torch.cuda.mem_get_info(dev)[0])
        N = w.n * w.n
        pb = projection_bytes(N, st, fp64=precision == P2D.FP64)
        r = {"ms_median": round(ms, 4), "projections_per_s": round(1000.0 / ms, 2),
             "iterations": st["iterations"], "outer": st["outer"], "max_div": st["max_div"], "dt": w.dt,
             "algorithmic_GBs": round(pb / (ms * 1e-3) / 1e9, 1)}
        if adjoint:
            ut, vt = W.target_velocity(w.n)
            fwd()
            gu = (u - T(ut)).contiguous()
            gv = (v - T(vt)).contiguous()
            ams, sta = _timed(torch, stream, lambda: ctx.adjoint(gu, gv)[4], max(3, steps // 2))
            free.append(torch.cuda.mem_get_info(dev)[0])
            r["adjoint"] = {"ms_median": round(ams, 4), "iterations": sta["iterations"], "status": sta["status"],
                            "forward_plus_adjoint_over_forward": round((ms + ams) / ms, 3)}
        r["device_memory"] = {"context_bytes": free[0] - free[1], "peak_bytes": free[0] - min(free),
                              "peak_bytes_per_cell": round((free[0] - min(free)) / N, 1),
                              "how": "hipMemGetInfo (torch.cuda.mem_get_info) before the context, after it, "
                                     "after the forward runs and after the adjoint runs; peak = the largest drop "
                                     "in free memory, inputs and gradients included"}
        del ctx
        return r

    w1 = W.drops64(1024)
    out["c1_drops64_1024_fp64"] = run(w1, P2D.FP64, True)
    out["c1_drops64_1024_fp64"]["solver"] = "FP64 MG-PCG (8-level V-cycle, 3+3 red-black sweeps)"
    out["c1_drops64_1024_fp64_2_2_sweeps"] = run(w1, P2D.FP64, False, mg_pre_sweeps=2, mg_post_sweeps=2)
    out["c1_drops64_1024_fp64_2_2_sweeps"]["solver"] = "FP64 MG-PCG (8-level V-cycle, 2+2 red-black sweeps)"
    out["c1_drops64_1024_mixed"] = run(w1, P2D.MIXED, False)
    w4 = W.drops512(4096)
    out["c4_drops512_4096_mixed"] = run(w4, P2D.MIXED, True)
    out["c4_drops512_4096_fp64"] = run(w4, P2D.FP64, False)
    out["c4_drops512_4096_mixed"]["speedup_vs_fp64"] = round(
        out["c4_drops512_4096_fp64"]["ms_median"] / out["c4_drops512_4096_mixed"]["ms_median"], 3)
    out["note"] = ("C3 (adjoint at 1024^2) is c1_drops64_1024_fp64.adjoint; C4's dt is the CFL step of its "
                   "state (workloads.DROPS512_CFL_DT)")
    return out


def read_traffic(n):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 PMC summary of
    this workload size, or None when there is none."""
    try:
        with open(TRAFFIC_FILE.format(n=n)) as f:
            text = f.read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", text)
    return float(m.group(1)) if m else None


def cpu_baseline(w):
    """The oracle's serial FP64 MG-PCG projection of the same workload (one full projection)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    t = time.perf_counter()
    p, u, v, st = O.project(w.phi, w.rho_minus, w.rho_plus, w.sigma, w.u, w.v, w.h, solver=O.SOLVER_MG, tol=1e-12)
    dt = time.perf_counter() - t
    base = {"value": 1.0 / dt, "unit": "projections/s", "cores": 1, "kind": "port",
            "sample": f"one full projection of the same {w.n}^2 RT workload by the oracle "
                      f"(serial FP64 MG-PCG to relative residual 1e-12, {st['iterations']} iterations), {dt:.1f} s"}
    # the reference's live solver path (CSR assembly + serial IC(0)-PCG, PCG_METHOD.cpp:176-264),
    # bounded to 5 and 20 iterations: per-iteration cost at this size (its full solve runs
    # for thousands of iterations, far beyond a bench's budget)
    t0 = time.perf_counter()
    O.project(w.phi, w.rho_minus, w.rho_plus, w.sigma, w.u, w.v, w.h, solver=O.SOLVER_IC0, max_iteration=5)
    t1 = time.perf_counter()
    _, _, _, st20 = O.project(w.phi, w.rho_minus, w.rho_plus, w.sigma, w.u, w.v, w.h, solver=O.SOLVER_IC0,
                              max_iteration=20)
    t2 = time.perf_counter()
    base["reference_path_sample"] = {
        "solver": "oracle restatement of the reference's CSR assembly + serial IC(0)-PCG",
        "seconds_5_iterations": round(t1 - t0, 3), "seconds_20_iterations": round(t2 - t1, 3),
        "seconds_per_iteration": round((t2 - t1 - (t1 - t0)) / 15.0, 4),
        "residual_after_20": float(st20["residual"])}

    # the reference's other solvers on the same system, bounded the same way: serial CG
    # (CG_METHOD.cpp:13-73) and Jacobi-PCG on one and on all host threads
    # (PCG_METHOD.cpp:266-383, the threaded path)
    def per_iteration(solver, threads=0):
        a = time.perf_counter()
        O.project(w.phi, w.rho_minus, w.rho_plus, w.sigma, w.u, w.v, w.h, solver=solver, max_iteration=5,
                  threads=threads)
        b = time.perf_counter()
        O.project(w.phi, w.rho_minus, w.rho_plus, w.sigma, w.u, w.v, w.h, solver=solver, max_iteration=25,
                  threads=threads)
        c = time.perf_counter()
        return round((c - b - (b - a)) / 20.0, 4)

    all_threads = host_threads()
    base["other_reference_solvers"] = {
        "cg_serial_seconds_per_iteration": per_iteration(O.SOLVER_CG, 1),
        "jacobi_pcg_1_thread_seconds_per_iteration": per_iteration(O.SOLVER_JACOBI, 1),
        f"jacobi_pcg_{all_threads}_threads_seconds_per_iteration": per_iteration(O.SOLVER_JACOBI, all_threads),
        "how": "oracle restatements, 25 minus 5 bounded iterations (CSR assembly cancels), same workload"}
    # iterations to the stopping rule (relative 1e-10) at 256^2 and 512^2 of the same
    # workload family, extrapolated to this size (CG iterations grow like n for Poisson)
    W = importlib.import_module("2d_engine_amd.workloads")
    conv = {}
    for label, solver, threads, per_it in (("ic0_serial", O.SOLVER_IC0, 1,
                                            base["reference_path_sample"]["seconds_per_iteration"]),
                                           (f"jacobi_{all_threads}_threads", O.SOLVER_JACOBI, all_threads,
                                            base["other_reference_solvers"][
                                                f"jacobi_pcg_{all_threads}_threads_seconds_per_iteration"])):
        its = {}
        for m in (256, 512):
            wm = W.rt(m)
            a = time.perf_counter()
            _, _, _, stm = O.project(wm.phi, wm.rho_minus, wm.rho_plus, wm.sigma, wm.u, wm.v, wm.h, solver=solver,
                                     tol=1e-10, threads=threads)
            its[m] = (stm["iterations"], round(time.perf_counter() - a, 3))
        est_its = its[512][0] * (w.n / 512.0)
        conv[label] = {"iterations_and_seconds_256": its[256], "iterations_and_seconds_512": its[512],
                       f"estimated_iterations_{w.n}": int(round(est_its)),
                       f"estimated_seconds_{w.n}": round(est_its * per_it, 1)}
    base["reference_solvers_to_convergence"] = conv
    base["host"] = {"nproc": os.cpu_count(), "threads_used_for_all_core": all_threads, "cpu": cpu_model()}
    return base, (p, u, v)


def host_threads():
    """The host threads this process may use: the affinity mask, capped by OMP_NUM_THREADS."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison leg")
    ap.add_argument("--no-configs", action="store_true", help="skip the other configurations (1024^2, 4096^2)")
    ap.add_argument("--pmc-calibration", action="store_true",
                    help="end with three 256 MiB device copies (reference for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is a design (y-strips), not built", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    P2D = importlib.import_module("2d_engine_amd")
    W = P2D.workloads
    dev = torch.device("cuda:0")
    log(f"bench: building RT workload n={args.n}")
    w = W.rt(args.n)
    N = w.n * w.n

    ctx_m, t_m, st_m, uv_m = time_path(P2D, torch, w, P2D.MIXED, args.steps, args.warmup, dev)
    p_m = ctx_m.pressure(mean_zero=True)
    fp64 = None
    if not args.no_fp64:
        ctx_d, t_d, st_d, _ = time_path(P2D, torch, w, P2D.FP64, args.steps, args.warmup, dev)
        ms_d = statistics.median(t_d)
        pb_d = projection_bytes(N, st_d, fp64=True)
        fp64 = {"projections_per_s": 1000.0 / ms_d, "ms_median": ms_d,
                "iterations": st_d["iterations"], "outer": st_d["outer"], "max_div": st_d["max_div"],
                "algorithmic_bytes": pb_d, "achieved_GBs": round(pb_d / (ms_d * 1e-3) / 1e9, 1)}
    ms = statistics.median(t_m)
    value = 1000.0 / ms
    adj_ms, st_a = time_adjoint(ctx_m, torch, w, uv_m, max(3, args.steps // 4), dev)
    adjoint = {"ms_median": adj_ms, "iterations": st_a["iterations"],
               "forward_plus_adjoint_over_forward": (ms + adj_ms) / ms}

    # per-kernel HBM rates at the finest level: algorithmic bytes per launch over the
    # average launch duration (after the adjoint, which overwrites no forward state read here)
    kernels = {}
    for which, name, bpc, what in KERNELS:
        kms = ctx_m.time_kernel(which, 50)
        kbytes = bpc * N
        kernels[name] = {"avg_launch_ms": round(kms, 5), "bytes_per_launch": kbytes,
                         "achieved_GBs": round(kbytes / (kms * 1e-3) / 1e9, 1), "bytes_per_cell": bpc,
                         "accounting": what}
    levelset = time_reinit(P2D, torch, ctx_m, w, dev)
    step = time_step(P2D, torch, w, dev)
    semi = time_semi_implicit(P2D, torch, w.n, dev)
    configs = time_configs(P2D, torch, W, dev) if not args.no_configs else None
    dom = kernels[DOMINANT]
    traffic = read_traffic(w.n)
    pbytes = projection_bytes(N, st_m)
    per_projection = {"algorithmic_bytes": pbytes, "achieved_GBs": round(pbytes / (ms * 1e-3) / 1e9, 1),
                      "frac": round(pbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                      "vcycles": st_m["iterations"], "passes": st_m["outer"],
                      "accounting": f"N x (iterations x {VCYCLE_B_PER_CELL:.1f} + passes x "
                                    f"({HEAD_B_PER_CELL} + {REFINE_B_PER_CELL}) + {SETUP_B_PER_CELL}) bytes, "
                                    "see bench.py VCYCLE_B_PER_CELL"}
    roofline = {"bound": "hbm", "achieved": dom["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(dom["achieved_GBs"] / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": DOMINANT + " (the largest share of the solve's kernel time)",
                "bytes_per_launch": dom["bytes_per_launch"], "avg_launch_ms": dom["avg_launch_ms"]}
    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "projections/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 inner / f64 refinement", "data": "synthetic",
        "config": {"workload": f"rt {w.n}x{w.n}, rho 1000:1, sigma 0, u,v~U[-1,1] seed 2, dt 1",
                   "stopping": "max|D u_proj| <= 1e-10 and ||r||/||b|| <= 1e-10",
                   "solver": "FP32 MG-PCG in FP64 iterative refinement", "parallelism": "1 GPU"},
        "mixed": {"iterations": st_m["iterations"], "outer": st_m["outer"], "max_div": st_m["max_div"],
                  "rel_residual": st_m["rel_residual"]},
        "fp64": fp64,
        "speedup_vs_fp64": round(value / fp64["projections_per_s"], 3) if fp64 else None,
        "adjoint": adjoint,
        "roofline": roofline,
        "kernels": kernels,
        "per_projection": per_projection,
        "levelset": levelset,
        "time_step": step,
        "viscosity_semi_implicit": semi,
        "configs": configs,
    }
    if not args.no_cpu:
        log("bench: CPU baseline (oracle, serial)")
        base, (p_o, u_o, v_o) = cpu_baseline(w)
        line["cpu_baseline"] = base
        um, vm = uv_m[0].cpu().numpy(), uv_m[1].cpu().numpy()
        p_o = p_o - p_o.mean()
        du = max(np.abs(um - u_o).max(), np.abs(vm - v_o).max())
        line["mixed"]["error_vs_fp64_oracle"] = {
            "pressure_rel_l2": float(np.linalg.norm(p_m - p_o) / np.linalg.norm(p_o)),
            "velocity_max_abs": float(du),
            "velocity_rel_max": float(du / max(np.abs(u_o).max(), np.abs(v_o).max()))}
    print(json.dumps(line), flush=True)
    if args.pmc_calibration:
        x = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()


if __name__ == "__main__":
    main()
