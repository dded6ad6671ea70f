"""The projection at the workloads' full sizes (1024^2, 2048^2, 4096^2), where the CPU
oracle is too slow to run in a test: size-independent properties, all recomputed on the
host from the GPU's output.

- max|D u_proj| from the returned velocities (numpy, FP64) meets the stopping rule;
- mixed and FP64 paths give the same projected velocity and mean-zero pressure;
- linearity (sigma = 0): P(a u1 + b u2) = a P(u1) + b P(u2);
- idempotence: projecting a projected field changes it by no more than the solver tolerance;
- the wall faces stay zero;
- at 4096^2 (512 drops, at the workload's CFL step) the adjoint against central differences.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _div(u, v, h):
    return ((u[:, 1:] - u[:, :-1]) + (v[1:, :] - v[:-1, :])) / h


def _walls_zero(u, v):
    return np.all(u[:, 0] == 0) and np.all(u[:, -1] == 0) and np.all(v[0] == 0) and np.all(v[-1] == 0)


def test_rt2048_mixed_and_fp64_agree(P2D, W):
    w = W.rt(2048)
    print("case full size rt 2048 mixed vs fp64", flush=True)
    out = {}
    for prec in (P2D.MIXED, P2D.FP64):
        ctx = P2D.Context(w.n, w.n, w.h, precision=prec, rtol=1e-10, div_tol=1e-10)
        ctx.set_fluid(w.phi, w.rho_minus, w.rho_plus, w.sigma)
        u, v, st = ctx.project(w.u, w.v, w.dt)
        div = np.abs(_div(u, v, w.h)).max()
        print(f"  prec={prec} iters={st['iterations']} outer={st['outer']} host max|Du|={div:.2e}")
        assert div <= 1e-10
        assert _walls_zero(u, v)
        out[prec] = (u, v, ctx.pressure(mean_zero=True))
    (um, vm, pm), (ud, vd, pd) = out[P2D.MIXED], out[P2D.FP64]
    scale = max(np.abs(ud).max(), np.abs(vd).max())
    assert max(np.abs(um - ud).max(), np.abs(vm - vd).max()) <= 1e-9 * scale
    assert np.linalg.norm(pm - pd) <= 1e-9 * np.linalg.norm(pd)


def test_rt2048_linearity_and_idempotence(P2D, W):
    w = W.rt(2048)
    print("case full size rt 2048 linearity", flush=True)
    ctx = P2D.Context(w.n, w.n, w.h, precision=P2D.MIXED, rtol=1e-12, div_tol=1e-11)
    ctx.set_fluid(w.phi, w.rho_minus, w.rho_plus, 0.0)
    u1, v1 = w.u, w.v
    u2, v2 = W.uniform_velocity(w.n, 17)
    a, b = 0.75, -1.5
    p1u, p1v, _ = ctx.project(u1, v1)
    p2u, p2v, _ = ctx.project(u2, v2)
    pcu, pcv, _ = ctx.project(a * u1 + b * u2, a * v1 + b * v2)
    scale = max(np.abs(pcu).max(), np.abs(pcv).max())
    err = max(np.abs(pcu - (a * p1u + b * p2u)).max(), np.abs(pcv - (a * p1v + b * p2v)).max())
    print(f"  linearity error {err:.2e} (scale {scale:.2e})")
    assert err <= 1e-9 * scale
    qu, qv, st = ctx.project(p1u, p1v)
    d = max(np.abs(qu - p1u).max(), np.abs(qv - p1v).max())
    print(f"  idempotence change {d:.2e}, iterations {st['iterations']}")
    assert d <= 1e-9 * max(np.abs(p1u).max(), np.abs(p1v).max())


@pytest.mark.parametrize("nu", [3, 2])
def test_drops64_1024_fp64(P2D, W, nu):
    """The 1024^2 FP64 case with the default 3+3 sweeps and with the 2+2 red-black pre- and
    post-sweeps the configuration names (the hierarchy keeps its full depth: a four-level
    cycle would leave a 128^2 bottom that no exact solve handles at this cost)."""
    w = W.drops64(1024)
    print(f"case full size drops64 1024 fp64 nu={nu}", flush=True)
    ctx = P2D.Context(w.n, w.n, w.h, precision=P2D.FP64, rtol=1e-10, div_tol=1e-10, mg_pre_sweeps=nu,
                      mg_post_sweeps=nu)
    ctx.set_fluid(w.phi, w.rho_minus, w.rho_plus, w.sigma)
    u, v, st = ctx.project(w.u, w.v, w.dt)
    div = np.abs(_div(u, v, w.h)).max()
    print(f"  iters={st['iterations']} host max|Du|={div:.2e}")
    assert st["status"] == 0 and div <= 1e-10 and _walls_zero(u, v)
    p = ctx.pressure(mean_zero=True)
    assert abs(p.mean()) <= 1e-12 * np.abs(p).max()


def test_drops512_4096_at_its_cfl_step(P2D, W):
    """The 4096^2 case with 512 small drops (sigma 0.0728) at the step the reference's CFL rule
    allows for its state: p2d_cfl_dt returns the workload's dt; mixed and FP64 reach max|D u|
    <= 1e-10 (host-recomputed) with status OK, and agree to 1e-9."""
    w = W.drops512(4096)
    print(f"case full size drops512 4096 dt={w.dt:.6e}", flush=True)
    res = {}
    for prec in (P2D.MIXED, P2D.FP64):
        ctx = P2D.Context(w.n, w.n, w.h, precision=prec, rtol=1e-10, div_tol=1e-10, max_iterations=2000)
        if prec == P2D.FP64:
            dt = ctx.cfl_dt(w.phi, w.u, w.v, w.extra["cfl_number"], rho_minus=w.rho_minus, rho_plus=w.rho_plus,
                            sigma=w.sigma, mu_minus=0.0, mu_plus=0.0)
            assert abs(dt - w.dt) <= 1e-13 * w.dt, (dt, w.dt)
        ctx.set_fluid(w.phi, w.rho_minus, w.rho_plus, w.sigma)
        u, v, st = ctx.project(w.u, w.v, w.dt)
        div = np.abs(_div(u, v, w.h)).max()
        umax = max(np.abs(u).max(), np.abs(v).max())
        print(f"  prec={prec} status={st['status']} iters={st['iterations']} outer={st['outer']} "
              f"max|Du|={div:.2e} max|u|={umax:.2e}", flush=True)
        assert st["status"] == 0 and div <= 1e-10
        assert _walls_zero(u, v)
        res[prec] = (u, v, ctx.pressure(mean_zero=True))
    (um, vm, pm), (ud, vd, pd) = res[P2D.MIXED], res[P2D.FP64]
    scale = max(np.abs(ud).max(), np.abs(vd).max())
    assert max(np.abs(um - ud).max(), np.abs(vm - vd).max()) <= 1e-9 * scale
    assert np.linalg.norm(pm - pd) <= 1e-9 * np.linalg.norm(pd)


def test_drops512_4096_adjoint(P2D, W):
    """The adjoint leg of the 4096^2 case: the mixed forward at its CFL step, then p2d_adjoint
    with only the forward pressure kept, for 0.5 ||u_proj - u_t||^2 (u_t seed 3).  Central
    differences of the same mixed forward along 4 seeded directions of the velocity and 4
    relative perturbations of beta.  The loss is quadratic in u, so the velocity step can be
    large (1e-3) without truncation error; beta uses 1e-5."""
    import torch
    w = W.drops512(4096)
    print("case full size drops512 4096 adjoint", flush=True)
    dev = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    free0 = torch.cuda.mem_get_info(dev)[0]
    ctx = P2D.Context(w.n, w.n, w.h, precision=P2D.MIXED, rtol=1e-13, div_tol=0.0, max_iterations=2000)
    free1 = torch.cuda.mem_get_info(dev)[0]
    ctx.set_fluid(T(w.phi), w.rho_minus, w.rho_plus, w.sigma)
    u_d, v_d = T(w.u), T(w.v)
    up, vp = u_d.clone(), v_d.clone()
    _, _, st = ctx.project(up, vp, w.dt, allow_not_converged=True)
    bx, by, jc = ctx.coefficients()
    ut, vt = W.target_velocity(w.n)
    ut_d, vt_d = T(ut), T(vt)
    gu, gv = up - ut_d, vp - vt_d
    gu[:, 0] = 0; gu[:, -1] = 0; gv[0] = 0; gv[-1] = 0
    du, dv, dbx, dby, sta = ctx.adjoint(gu.contiguous(), gv.contiguous())
    torch.cuda.synchronize()
    free2 = torch.cuda.mem_get_info(dev)[0]
    print(f"  forward iters={st['iterations']} rel_res={st['rel_residual']:.1e}; adjoint status={sta['status']} "
          f"iters={sta['iterations']}; context {(free0 - free1) / 1e9:.2f} GB, device use after the adjoint "
          f"{(free0 - free2) / 1e9:.2f} GB", flush=True)
    assert sta["status"] == 0
    assert free0 - free1 <= 8e9          # ~450 B per cell of 16.8 M cells
    phi_d, jc_d, bx_d, by_d = T(w.phi), T(jc), T(bx), T(by)

    def loss(u, v, bxx, byy):
        ctx.set_coefficients(bxx, byy, phi_d, jc_d)
        uu, vv = u.clone(), v.clone()
        ctx.project(uu, vv, w.dt, allow_not_converged=True)
        return 0.5 * (torch.sum((uu - ut_d) ** 2) + torch.sum((vv - vt_d) ** 2)).item()

    rng = np.random.default_rng(41)

    def dirs():
        a = rng.standard_normal(w.u.shape); a[:, 0] = a[:, -1] = 0
        b = rng.standard_normal(w.v.shape); b[0] = b[-1] = 0
        return T(a), T(b)

    worst_u = worst_b = 0.0
    for k in range(4):
        a, b = dirs()
        eps = 1e-3
        fd = (loss(u_d + eps * a, v_d + eps * b, bx_d, by_d) - loss(u_d - eps * a, v_d - eps * b, bx_d, by_d)) / (2 * eps)
        an = (torch.sum(du * a) + torch.sum(dv * b)).item()
        worst_u = max(worst_u, abs(fd - an) / abs(an))
        sa, sb = dirs()
        sa, sb = sa * bx_d, sb * by_d
        eps = 1e-5
        fd = (loss(u_d, v_d, bx_d + eps * sa, by_d + eps * sb) - loss(u_d, v_d, bx_d - eps * sa, by_d - eps * sb)) / (2 * eps)
        an = (torch.sum(dbx * sa) + torch.sum(dby * sb)).item()
        worst_b = max(worst_b, abs(fd - an) / abs(an))
        print(f"  direction {k}: worst relative error velocity {worst_u:.2e}, beta {worst_b:.2e}", flush=True)
    assert worst_u <= 1e-6 and worst_b <= 1e-4, (worst_u, worst_b)
