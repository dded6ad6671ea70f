"""Benchmark: projections/s of the pressure projection at 2048^2, density ratio 1000:1.

One step = one full projection (PROJECTION_2D::Solve) of the Rayleigh-Taylor
workload (phi = y - 0.5 - 0.01 cos(8 pi x), rho 1000 above / 1 below, sigma 0,
u,v ~ U[-1,1] seed 2, dt = 1) from device-resident phi, u, v to device-resident
projected u, v and p: coefficient assembly, right-hand side, the solve to
max|D u_proj| <= 1e-10 (and ||r||_2/||b||_2 <= 1e-10), the face update and the
FP64 divergence check.  Each step starts from the same unprojected velocity
(a device copy outside the timed interval) with a zero initial pressure.

value = projections/s of the mixed-precision path (FP32 MG-PCG inside FP64
refinement); the FP64 path is timed the same way and reported beside it.
"""
import argparse
import importlib
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "projections/s at 2048² ρ 1000:1, div≤1e-10: mixed vs FP64; HBM GB/s"
HBM_PEAK_GBS = 8000.0


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
    return ctx, times, stats


def cpu_baseline(w):
    """The oracle's serial FP64 MG-PCG projection of the same workload (one full projection)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    t = time.perf_counter()
    _, u, v, st = O.project(w.phi, w.rho_minus, w.rho_plus, w.sigma, w.u, w.v, w.h, solver=O.SOLVER_MG, tol=1e-10)
    dt = time.perf_counter() - t
    return {"value": 1.0 / dt, "unit": "projections/s", "cores": 1, "kind": "port",
            "sample": f"one full projection of the same {w.n}^2 RT workload by the oracle "
                      f"(serial FP64 CSR-free MG-PCG, {st['iterations']} iterations, rel 1e-10), {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is a design (y-strips), not built", file=sys.stderr)
        sys.exit(2)
    import torch
    P2D = importlib.import_module("2d_engine_amd")
    W = P2D.workloads
    dev = torch.device("cuda:0")
    log(f"bench: building RT workload n={args.n}")
    w = W.rt(args.n)
    N = w.n * w.n

    ctx_m, t_m, st_m = time_path(P2D, torch, w, P2D.MIXED, args.steps, args.warmup, dev)
    fp64 = None
    if not args.no_fp64:
        ctx_d, t_d, st_d = time_path(P2D, torch, w, P2D.FP64, args.steps, args.warmup, dev)
        fp64 = {"projections_per_s": 1000.0 / statistics.median(t_d), "ms_median": statistics.median(t_d),
                "iterations": st_d["iterations"], "outer": st_d["outer"], "max_div": st_d["max_div"]}
    ms = statistics.median(t_m)
    value = 1000.0 / ms

    # FP32 direction update + stencil (k_dir<float>) at the finest level.
    # Algorithmic bytes per launch: read z, p_old, beta_x, beta_y; write p_new, q = 6 * 4 * N.
    reps = 200
    kms = ctx_m.time_kernel(0, reps)
    kbytes = 6 * 4 * N
    achieved = kbytes / (kms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_dir<float> (FP32 direction update + 5-point GFM stencil + p.q)",
                "bytes_per_launch": kbytes, "avg_launch_ms": kms}
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
        "roofline": roofline,
    }
    if not args.no_cpu:
        log("bench: CPU baseline (oracle, serial)")
        line["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
