"""GPU adjoint of the projection (p2d_adjoint): against the oracle's FP64 CPU
gradient, against central finite differences of the GPU forward, and through
the test-only torch.autograd binding.

Loss 0.5 ||u_proj - u_target||^2 (u_target ~ U[-1,1], seed 3); gradients with
respect to the input face velocities and the face coefficients beta on
interior faces (wall faces carry boundary data: gradient 0 there).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(O, W, name, n):
    w = W.make(name, n)
    bx, by = O.face_beta(w.phi, w.rho_minus, w.rho_plus)
    jc = w.sigma * O.curvature(w.phi, w.h)
    ut, vt = W.target_velocity(n)
    return w, bx, by, jc, ut, vt


def _walls_zero(a, b):
    a = a.copy(); b = b.copy()
    a[:, 0] = a[:, -1] = 0
    b[0] = b[-1] = 0
    return a, b


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("prec,n,div_tol,tol", [(0, 128, 1e-12, 1e-8), (1, 128, 1e-12, 1e-8), (0, 1024, 1e-10, 1e-7)])
def test_adjoint_matches_oracle(P2D, O, W, prec, n, div_tol, tol):
    """At 1024^2 the drops' max |D u| floor (rounding of u) is ~2e-11: div_tol 1e-10 there."""
    w, bx, by, jc, ut, vt = _case(O, W, "drops64", n)
    print(f"case adjoint vs oracle drops64 n={n} prec={prec}", flush=True)
    ctx = P2D.Context(n, n, w.h, precision=prec, rtol=1e-13, div_tol=div_tol)
    ctx.set_coefficients(bx, by, w.phi, jc)
    up, vp, _ = ctx.project(w.u, w.v, w.dt)
    gu, gv = _walls_zero(up - ut, vp - vt)
    du, dv, dbx, dby, st = ctx.adjoint(gu, gv)
    assert st["status"] == 0
    p_r, up_r, vp_r = O.project_beta(w.phi, jc, bx, by, w.u, w.v, w.h, rtol=1e-14)
    gu_r, gv_r = _walls_zero(up_r - ut, vp_r - vt)
    du_r, dv_r, dbx_r, dby_r, _ = O.adjoint(w.phi, jc, bx, by, p_r, gu_r, gv_r, w.h, rtol=1e-14)
    for a, b in ((du, du_r), (dv, dv_r), (dbx, dbx_r), (dby, dby_r)):
        assert _rel(a, b) <= tol, _rel(a, b)


def _interior_dirs(w, rng):
    a = rng.standard_normal(w.u.shape); a[:, 0] = a[:, -1] = 0
    b = rng.standard_normal(w.v.shape); b[0] = b[-1] = 0
    return a, b


@pytest.mark.parametrize("name,n,ndir", [("drops64", 64, 4), ("drops64", 1024, 32)])
def test_adjoint_central_differences(P2D, O, W, name, n, ndir):
    """Central differences (step 1e-6, FP64) of the GPU forward along seeded random
    directions, for the velocities and (relative perturbations of) beta."""
    import torch
    w, bx, by, jc, ut, vt = _case(O, W, name, n)
    print(f"case adjoint FD {name} n={n} dirs={ndir}", flush=True)
    dev = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    ctx = P2D.Context(n, n, w.h, precision=P2D.FP64, rtol=1e-14, div_tol=1e-13, max_iterations=400)
    phi_d, jc_d, bx_d, by_d = T(w.phi), T(jc), T(bx), T(by)
    u_d, v_d, ut_d, vt_d = T(w.u), T(w.v), T(ut), T(vt)

    def loss(u, v, bxx, byy):
        ctx.set_coefficients(bxx, byy, phi_d, jc_d)
        uu, vv = u.clone(), v.clone()
        ctx.project(uu, vv, w.dt, allow_not_converged=True)
        return 0.5 * (torch.sum((uu - ut_d) ** 2) + torch.sum((vv - vt_d) ** 2)).item(), uu, vv

    _, up, vp = loss(u_d, v_d, bx_d, by_d)
    gu, gv = up - ut_d, vp - vt_d
    gu[:, 0] = 0; gu[:, -1] = 0; gv[0] = 0; gv[-1] = 0
    du, dv, dbx, dby, _ = ctx.adjoint(gu.contiguous(), gv.contiguous())
    rng = np.random.default_rng(11)
    eps = 1e-6
    worst_u = worst_b = 0.0
    for k in range(ndir):
        a, b = _interior_dirs(w, rng)
        a_d, b_d = T(a), T(b)
        fd = (loss(u_d + eps * a_d, v_d + eps * b_d, bx_d, by_d)[0] -
              loss(u_d - eps * a_d, v_d - eps * b_d, bx_d, by_d)[0]) / (2 * eps)
        an = (torch.sum(du * a_d) + torch.sum(dv * b_d)).item()
        worst_u = max(worst_u, abs(fd - an) / abs(an))
        sa, sb = _interior_dirs(w, rng)
        sa_d, sb_d = T(sa * bx), T(sb * by)
        fd = (loss(u_d, v_d, bx_d + eps * sa_d, by_d + eps * sb_d)[0] -
              loss(u_d, v_d, bx_d - eps * sa_d, by_d - eps * sb_d)[0]) / (2 * eps)
        an = (torch.sum(dbx * sa_d) + torch.sum(dby * sb_d)).item()
        worst_b = max(worst_b, abs(fd - an) / abs(an))
    print(f"  worst relative error: velocity {worst_u:.2e}, beta {worst_b:.2e}", flush=True)
    assert worst_u <= 1e-6 and worst_b <= 1e-5, (worst_u, worst_b)


def test_autograd_binding(P2D, O, W):
    import torch
    from torch_binding import project
    n = 128
    w, bx, by, jc, ut, vt = _case(O, W, "drops64", n)
    print(f"case autograd drops64 n={n}", flush=True)
    dev = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    ctx = P2D.Context(n, n, w.h, precision=P2D.FP64, rtol=1e-13, div_tol=1e-12)
    u = T(w.u).requires_grad_(True)
    v = T(w.v).requires_grad_(True)
    bxt = T(bx).requires_grad_(True)
    byt = T(by).requires_grad_(True)
    up, vp = project(ctx, T(w.phi), T(jc), u, v, bxt, byt, w.dt)
    loss = 0.5 * (torch.sum((up - T(ut)) ** 2) + torch.sum((vp - T(vt)) ** 2))
    loss.backward()
    p_r, up_r, vp_r = O.project_beta(w.phi, jc, bx, by, w.u, w.v, w.h, rtol=1e-14)
    assert _rel(up.detach().cpu().numpy(), up_r) <= 1e-9
    # the loss gradient also flows through the (untouched) wall faces; the binding
    # reports the interior part, as the oracle
    gu_r, gv_r = _walls_zero(up_r - ut, vp_r - vt)
    du_r, dv_r, dbx_r, dby_r, _ = O.adjoint(w.phi, jc, bx, by, p_r, gu_r, gv_r, w.h, rtol=1e-14)
    for g, r in ((u.grad, du_r), (v.grad, dv_r), (bxt.grad, dbx_r), (byt.grad, dby_r)):
        g = g.cpu().numpy()
        assert _rel(g, r) <= 1e-8, _rel(g, r)
